// The back-pointer walk of stage 2 (one wave per path segment), in a file of
// its own: it is built with uniform control flow left unstructured
// (-structurizecfg-skip-uniform-regions, see the Makefile), which keeps the
// serial step loop free of exec-mask bookkeeping.  The rest of stage 2 is in
// pf_stage2.hip.
#include "pf_common.h"

namespace pf {

// _nb_backtrace_within_chunk (Parflex.py:976-1012) for chain segment
// blockIdx.x of request blockIdx.y, one wave.  The walk itself is serial
// and keeps all of its state in scalar registers: the wave holds a window
// of packed back-pointers (128 rows x 128 columns ending at the current
// cell; lane l holds columns wc0+l and wc0+64+l, one word per 16-row block)
// and reads each step's word with v_readlane.  A step only appends its 2-bit
// code to a scalar word (16 codes per word, kept in LDS); the points are
// rebuilt afterwards 64 at a time, each lane one point, from a wave prefix
// sum of the steps, and written with coalesced stores.
constexpr int kWalkRB = 8;                       // row blocks in the window
constexpr int kWalkMaxCodes = 2 * 8192 + 16;     // steps of one walk, L <= 8192

__global__ __launch_bounds__(64) void k_s2_walk(const TileDesc *__restrict__ tiles, int64_t L,
                                                StepSet ss, const uint32_t *__restrict__ Bp,
                                                int64_t *__restrict__ chain,
                                                const int64_t *__restrict__ cnt,
                                                int64_t *__restrict__ pts) {
  // code word k (16 steps) at codes[k]; words are gathered in one VGPR
  // (word k in lane k % 64) and stored 64 at a time
  __shared__ uint32_t codes[(kWalkMaxCodes + 1023) / 1024 * 64];
  const int64_t q = blockIdx.y, s = blockIdx.x;
  if (s >= cnt[q]) return;
  int64_t *en = chain + (q * 101 + s) * 8;
  const int64_t t = en[0];
  const TileDesc td = wave_uniform(tiles[t]);
  // Walk state is wave-uniform; readfirstlane keeps it in scalar registers.
  const int cols = (int)td.W;
  const int sr = __builtin_amdgcn_readfirstlane((int)en[3]);
  const int sc = __builtin_amdgcn_readfirstlane((int)en[4]);
  const int64_t gro = en[5], gco = en[6];
  int64_t *out = pts + (q * 101 + s) * (2 * L + 8) * 2;
  const int lane = threadIdx.x;
  const int r_init = __builtin_amdgcn_readfirstlane((int)en[1]);
  const int c_init = __builtin_amdgcn_readfirstlane((int)en[2]);
  int r = r_init, c = c_init;
  // step table packed 4 bits per step: dr in bits 0-1, dc in bits 2-3
  int steptab = 0;
  for (int z = 0; z < 4; ++z) steptab |= (ss.dr[z] | (ss.dc[z] << 2)) << (4 * z);
  steptab = __builtin_amdgcn_readfirstlane(steptab);
  // the same per code as a byte for the fast steps: the row step doubled in
  // bits 0-5 (63 for a code that is no step) and the column step in bits 6-7
  uint32_t ftab = 0;
  for (int z = 0; z < 4; ++z)
    ftab |= (uint32_t)(z < ss.K ? (2 * ss.dr[z]) | (ss.dc[z] << 6) : 63) << (8 * z);
  ftab = __builtin_amdgcn_readfirstlane(ftab);
  // Codes enter acc from the top (two bits per step), so after 16 steps code
  // k of the word sits at bits 2k.  nw words are complete; rem + 1 codes are
  // still missing from the current one.
  int nw = 0, rem = 15;
  uint32_t acc = 0, cw = 0;
  auto flush = [&]() __attribute__((always_inline)) {
    cw = (lane == (nw & 63)) ? acc : cw;
    if ((nw & 63) == 63) codes[(nw >> 6) * 64 + lane] = cw;
    ++nw;
    acc = 0;
    rem = 15;
  };
  const int sr1 = sr + 1;
  // Window of packed back-pointers: word rows rb0 .. rb0 + kWalkRB - 1,
  // columns wc0 .. wc0 + 127 (two registers per word row).
  int rb0, wc0;
  uint32_t w[2 * kWalkRB];
  auto load_window = [&]() __attribute__((always_inline)) {
    rb0 = (r >> 4) - (kWalkRB - 1) > 0 ? (r >> 4) - (kWalkRB - 1) : 0;
    wc0 = c - 127 > 0 ? c - 127 : 0;
    const int rmax = r >> 4;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int col = wc0 + h * 64 + lane <= c ? wc0 + h * 64 + lane : c;  // never past c
      const uint32_t *base = Bp + td.b_off + col;
#pragma unroll
      for (int b = 0; b < kWalkRB; ++b)
        w[b * 2 + h] = base[(int64_t)(rb0 + b <= rmax ? rb0 + b : rmax) * cols];
    }
  };
  // The block the steps read: 64 columns from c_lo (lane = column - c_lo)
  // and 32 rows from r_lo, word rows r_lo / 16 (wlo) and r_lo / 16 + 1
  // (whi).  The cell is in the upper word row unless the window's bottom
  // word row is the tile's first; the walk only goes down and left, so only
  // the lower bounds are ever crossed.
  int r_lo, c_lo;
  uint32_t wlo, whi;
  auto select_block = [&]() __attribute__((always_inline)) {
    const int qi = (r >> 4) - rb0, h = (c - wc0) >> 6;
    const int ql = qi > 0 ? qi - 1 : 0;
    const int nh = qi * 2 + h, nl = ql * 2 + h;
    whi = w[0];
    wlo = w[0];
#pragma unroll
    for (int z = 1; z < 2 * kWalkRB; ++z) {
      whi = (nh == z) ? w[z] : whi;
      wlo = (nl == z) ? w[z] : wlo;
    }
    r_lo = (rb0 + ql) * 16;
    c_lo = wc0 + h * 64;
  };
  // (r, c) into the block; the window is reloaded when the cell is outside
  // it or in its bottom word row with more rows below
  auto ensure_block = [&]() __attribute__((always_inline)) {
    if (__builtin_expect(((r - r_lo) | (c - c_lo)) < 0, 0)) {
      const int qi = (r >> 4) - rb0;
      if ((qi | (c - wc0)) < 0 || (qi == 0 && rb0 > 0)) load_window();
      select_block();
    }
  };
  load_window();
  select_block();
  // The loop's only exits are stops.  Every step lowers the row (1 <= dr <= 2,
  // enforced when the step set is made), so a step off the tile (pr < 0)
  // ends the walk within R steps (R: the tile's rows).
  for (;;) {
    // One step with the exact tests: stop at the start, on an invalid code or
    // on a step off the tile (dr >= 1 and dc >= 0: only lower bounds).
    ensure_block();
    {
      const int dl = c - c_lo;
      const uint64_t wv = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)whi, dl) << 32) |
                          (uint32_t)__builtin_amdgcn_readlane((int)wlo, dl);
      const int code = (int)((wv >> (2 * (r - r_lo))) & 3u);
      const int e = steptab >> (4 * code);
      const int pr = r - (e & 3), pc = c - ((e >> 2) & 3);
      if ((r == sr && c == sc) || code >= ss.K || (pr | pc) < 0) break;
      acc = (acc >> 2) | ((uint32_t)code << 30);
      if (--rem < 0) flush();
      r = pr;
      c = pc;
    }
    // Fast steps while the cell is above the start row (there the start
    // cannot be the current cell).  The block is shifted so that its row
    // rbase = max(r_lo, sr + 1) is bit 0, so one sign test of the row and
    // column offsets and of rem covers leaving the block, falling to the
    // start row, stepping off the tile (offsets below zero), an invalid code
    // (its row step of 63 takes the row offset below zero) and a full code
    // word.  A step that was no step (invalid code, off the tile) is taken
    // back after the loop and ends the walk.
    while (r >= sr1) {
      ensure_block();
      const int rbase = r_lo > sr1 ? r_lo : sr1;
      const uint64_t f = (((uint64_t)whi << 32) | wlo) >> (2 * (rbase - r_lo));
      const uint32_t flo = (uint32_t)f, fhi = (uint32_t)(f >> 32);
      int dr2 = 2 * (r - rbase), dcl = c - c_lo;
      uint32_t t;  // the step's code in bits 0-1
      for (;;) {
        const uint64_t wv = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)fhi, dcl) << 32) |
                            (uint32_t)__builtin_amdgcn_readlane((int)flo, dcl);
        t = (uint32_t)(wv >> dr2);
        const uint32_t e = ftab >> ((t << 3) & 31);
        dr2 -= (int)(e & 63u);
        dcl -= (int)((e >> 6) & 3u);
        acc = __builtin_amdgcn_alignbit(t, acc, 2);
        --rem;
        if ((dr2 | dcl | rem) < 0) break;
      }
      const int last = (int)(t & 3u);
      c = c_lo + dcl;
      if (last >= ss.K) {
        r = rbase + ((dr2 + 63) >> 1);
        acc <<= 2;
        ++rem;
        goto walked;
      }
      r = rbase + (dr2 >> 1);
      if ((r | c) < 0) {
        const int e = steptab >> (4 * last);
        r += e & 3;
        c += (e >> 2) & 3;
        acc <<= 2;
        ++rem;
        goto walked;
      }
      if (rem < 0) flush();
    }
  }
walked:
  const int nc = nw * 16 + (15 - rem);  // step codes appended
  const int m = nc + 1;                 // cells visited, the last one included
  // (sr, sc) is appended after the last visited cell unless the walk ended
  // on it (it stopped on an invalid code or a step off the tile).
  const int tail = (r != sr || c != sc) ? 1 : 0;
  // a partial last word: its codes down to bit 0
  if ((nc & 15) != 0) cw = (lane == ((nc >> 4) & 63)) ? (acc >> (32 - 2 * (nc & 15))) : cw;
  if ((nc & 1023) != 0) codes[(nc >> 10) * 64 + lane] = cw;
  __syncthreads();
  // Cell i of the walk is (r_init, c_init) minus the steps of codes 0..i-1.
  int br = r_init, bc = c_init;
  for (int b0 = 0; b0 < m; b0 += 64) {
    const int j = b0 + lane;  // this lane's cell, and the code of its step out
    int dr = 0, dc = 0;
    if (j < nc) {
      const int code = (int)((codes[j >> 4] >> (2 * (j & 15))) & 3u);
      const int e = steptab >> (4 * code);
      dr = e & 3;
      dc = (e >> 2) & 3;
    }
    // inclusive prefix sums over the lanes, then exclusive by subtraction
    int sr_ = dr, sc_ = dc;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const int ur = __shfl_up(sr_, d), uc = __shfl_up(sc_, d);
      if (lane >= d) {
        sr_ += ur;
        sc_ += uc;
      }
    }
    if (j < m) {
      out[(int64_t)j * 2] = (br - (sr_ - dr)) + gro;
      out[(int64_t)j * 2 + 1] = (bc - (sc_ - dc)) + gco;
    }
    br -= __shfl(sr_, 63);
    bc -= __shfl(sc_, 63);
  }
  if (lane == 0) {
    if (tail) {
      out[(int64_t)m * 2] = sr + gro;
      out[(int64_t)m * 2 + 1] = sc + gco;
    }
    en[7] = m + tail;
  }
}

void launch_s2_walk(dim3 grid, hipStream_t st, const TileDesc *tiles, int64_t L, StepSet ss,
                    const uint32_t *Bp, int64_t *chain, const int64_t *cnt, int64_t *pts) {
  hipLaunchKernelGGL(k_s2_walk, grid, dim3(64), 0, st, tiles, L, ss, Bp, chain, cnt, pts);
}

}  // namespace pf
