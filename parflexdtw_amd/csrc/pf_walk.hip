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
  int nc = 0;       // step codes appended
  // codes enter acc from the top (two bits per step), so after 16 steps
  // code k of the word sits at bits 2k
  uint32_t acc = 0, cw = 0;
  const int sr1 = sr + 1;
  // Window of packed back-pointers: word rows rb0 .. rb0 + kWalkRB - 1,
  // columns wc0 .. wc0 + 127 (two registers per word row); the block of 16
  // rows x 64 columns holding (r, c) is word row r >> 4 in lanes
  // c_lo .. c_lo + 63 of wcur.
  int rb0, wc0, r_lo, c_lo;
  uint32_t w[2 * kWalkRB], wcur;
  auto load_window = [&]() {
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
  auto select_block = [&]() {
    const int qi = (r >> 4) - rb0, ci = c - wc0;
    const int ns = qi * 2 + (ci >> 6);
    wcur = w[0];
#pragma unroll
    for (int z = 1; z < 2 * kWalkRB; ++z) wcur = (ns == z) ? w[z] : wcur;
    r_lo = r & ~15;
    c_lo = wc0 + (ci & ~63);
  };
  load_window();
  select_block();
  // The loop's only exit is a stop.  Every step lowers the row (1 <= dr <= 2,
  // enforced when the step set is made), so a step off the tile (pr < 0)
  // ends the walk within R steps (R: the tile's rows).
  for (;;) {
    int dr_ = r - r_lo, dc_ = c - c_lo;
    if (__builtin_expect((dr_ | dc_) < 0, 0)) {
      if ((((r >> 4) - rb0) | (c - wc0)) < 0) load_window();
      select_block();
      dr_ = r - r_lo;
      dc_ = c - c_lo;
    }
    const uint32_t word = __builtin_amdgcn_readlane(wcur, dc_);
    const int code = (int)((word >> (2 * dr_)) & 3u);
    const int e = steptab >> (4 * code);
    const int pr = r - (e & 3), pc = c - ((e >> 2) & 3);
    // Stop at the start, on an invalid code or on a step off the tile
    // (dr >= 1 and dc >= 0: only lower bounds).  Above the start row
    // (r > sr) the start cannot be the current cell, so one sign test
    // covers the common case; the exact test runs from the start row on.
    if (__builtin_expect(((r - sr1) | (ss.K - 1 - code) | pr | pc) < 0, 0)) {
      // integer form (no compares): the top bit of (d - 1) & ~d is set
      // iff d == 0 for d >= 0, here d = 0 iff (r, c) == (sr, sc)
      const int d = (r ^ sr) | (c ^ sc);
      const int stop = ((d - 1) & ~d) | (ss.K - 1 - code) | pr | pc;
      if (stop < 0) break;
    }
    acc = (acc >> 2) | ((uint32_t)code << 30);
    if (__builtin_expect((nc & 15) == 15, 0)) {
      cw = (lane == ((nc >> 4) & 63)) ? acc : cw;
      acc = 0;
      if (((nc >> 4) & 63) == 63) codes[(nc >> 10) * 64 + lane] = cw;
    }
    ++nc;
    r = pr;
    c = pc;
  }
  const int m = nc + 1;  // cells visited, the last one included
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
