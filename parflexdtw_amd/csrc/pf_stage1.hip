// Stage 1 of ParFlex on gfx950: per-tile FlexDTW dynamic programme.
//
// Semantics follow FlexDTW.flexdtw (FlexDTW.py:1320-1364) exactly:
//   boundary: D[0,:] = C[0,:], D[:,0] = C[:,0], P = B = 0 there;
//   interior: for step k in order, pathcost = D[pred] + C*w_k (a multiply and
//             an add, each rounded: the library is built with
//             -ffp-contract=off), mdist = row + col - |P[pred]|,
//             cpm = pathcost / mdist, strict '<' keeps the first minimum;
//             D = pathcost of the winner, B = k, P = start column (pred on
//             row 0), -start row (pred on column 0) or P[pred].
//
// Comparisons without division.  Only the ORDER of the rounded quotients
// matters (cpm itself is never stored), so RN(a1/m1) < RN(a2/m2) is decided
// from the cross products a1*m2 and a2*m1: when they differ by more than
// 2^-49 of their magnitude the rounded quotients are more than two ulps apart
// and the sign of the difference is the answer; otherwise (a near tie) both
// quotients are computed with IEEE division and compared, exactly as the
// reference does.  The result is bit-identical to dividing every time.
// The FP32 DP of the precision form (BF16 or FP32 cost) has no bit-exact
// target: there path costs are fused multiply-adds and the order is taken
// from the cross products alone.
//
// Mapping.  A cell depends only on rows r-1 and r-2 (every step has
// dr >= 1), so a row is computed in parallel across its columns.  A tile's
// columns are cut into S strips of SW = NW*64*CPL columns; each (tile, strip)
// is one workgroup of NW waves that sweeps the rows top to bottom, CPL
// consecutive columns per lane, rows r-1 and r-2 in registers.  A lane takes
// its two left neighbour columns of row r-1 from lane-1 (DPP wave_shr), lane
// 0 of each wave from the wave on its left through an 8-row LDS ring with a
// per-wave progress counter (waves run decoupled, no barrier per row: a wave
// waits only for its left neighbour's previous row and, before overwriting a
// ring slot, for its right neighbour to have read it), and lane 0 of wave 0
// from the previous strip through a global exchange buffer: the strip to the
// left publishes its last two columns row by row and bumps a per-strip row
// counter every G rows; the strip to the right waits on it before each
// block of G rows.  The exchanged rows and the counter are relaxed atomics
// at agent scope, each access coherent across the XCDs' L2 caches; the
// producer stores the counter only after its row stores have completed
// (s_waitcnt), and the consumer reads the rows only after it has seen the
// counter.  (A release / acquire pair at agent scope would write back or
// invalidate the whole L2 of the XCD at every block.)
// Work items are taken from a ticket counter in strip order, so a strip only
// ever waits for an item that is already running: no deadlock whatever the
// grid size or residency.
//
// Outputs (packed mode): 2-bit back-pointers packed 16 rows per uint32 per
// column, and D/P on the tile's last row and last column only -- everything
// stage 2 needs (Parflex.py:509-564, :1063-1064).  Full mode writes D/P/B of
// every cell (the 6-tuple of gpu_flexdtw.flexdtw_chunk_from_global_C).
#include "pf_common.h"

#include <mutex>

namespace pf {

template <typename T>
__device__ __forceinline__ T dp_inf();
template <>
__device__ __forceinline__ double dp_inf<double>() { return __builtin_huge_val(); }
template <>
__device__ __forceinline__ float dp_inf<float>() { return __builtin_huge_valf(); }

template <typename DT, typename CT>
__device__ __forceinline__ DT load_c(const CT *p) {
  if constexpr (sizeof(DT) == 8) {
    return (DT)to_f64(*p);
  } else {
    return (DT)to_f32(*p);
  }
}

// RN(a1/m1) < RN(a2/m2) for positive integer-valued m (see the header).
__device__ __forceinline__ bool quot_less(double a1, double m1, double a2, double m2) {
  const double x = a1 * m2, y = a2 * m1;
  const double d = x - y;
  const double mag = fmax(fabs(x), fabs(y));
  if (fabs(d) > mag * 0x1p-49 && mag > 0x1p-900) return d < 0.0;
  return (a1 / m1) < (a2 / m2);
}
__device__ __forceinline__ bool quot_less(float a1, float m1, float a2, float m2) {
  const float x = a1 * m2, y = a2 * m1;
  const float d = x - y;
  const float mag = fmaxf(fabsf(x), fabsf(y));
  if (fabsf(d) > mag * 0x1p-20f && mag > 0x1p-100f) return d < 0.0f;
  return (a1 / m1) < (a2 / m2);
}

// Branch-free form of quot_less: returns the fast decision and sets ok when
// it is certain; callers redo the whole cell with quot_less when !ok.
__device__ __forceinline__ bool fast_less(double a1, double m1, double a2, double m2, bool &ok) {
  const double x = a1 * m2, y = a2 * m1;
  const double d = x - y;
  const double mag = fmax(fabs(x), fabs(y));
  ok = (fabs(d) > mag * 0x1p-49) && (mag > 0x1p-900);
  return d < 0.0;
}
__device__ __forceinline__ bool fast_less(float a1, float m1, float a2, float m2, bool &ok) {
  const float x = a1 * m2, y = a2 * m1;
  const float d = x - y;
  const float mag = fmaxf(fabsf(x), fabsf(y));
  ok = (fabsf(d) > mag * 0x1p-20f) && (mag > 0x1p-100f);
  return d < 0.0f;
}

// Boundary values a strip hands to the strip on its right, per row: D and
// the signed path length X of its last two columns.
template <typename DT>
struct Xch {
  DT d0, d1, x0, x1;
};

static_assert(sizeof(Xch<double>) == 32 && sizeof(Xch<float>) == 16,
              "pf_stage1_workspace reports these entry sizes");

// Exchange rows are stored and loaded as relaxed 8-byte atomics at agent
// scope (coherent across the XCDs without cache-wide write-backs).
template <typename DT>
__device__ __forceinline__ void xch_put(Xch<DT> *p, const Xch<DT> &v) {
  constexpr int n = (int)(sizeof(Xch<DT>) / 8);
  uint64_t w[n];
  __builtin_memcpy(w, &v, sizeof(w));
  uint64_t *q = reinterpret_cast<uint64_t *>(p);
#pragma unroll
  for (int k = 0; k < n; ++k)
    __hip_atomic_store(q + k, w[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename DT>
__device__ __forceinline__ Xch<DT> xch_get(const Xch<DT> *p) {
  constexpr int n = (int)(sizeof(Xch<DT>) / 8);
  uint64_t w[n];
  const uint64_t *q = reinterpret_cast<const uint64_t *>(p);
#pragma unroll
  for (int k = 0; k < n; ++k)
    w[k] = __hip_atomic_load(q + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  Xch<DT> v;
  __builtin_memcpy(&v, w, sizeof(w));
  return v;
}
// Publishes a strip's row counter once this wave's exchange-row stores have
// completed.
__device__ __forceinline__ void xch_publish(int *f, int rows) {
  __builtin_amdgcn_s_waitcnt(0);
  __hip_atomic_store(f, rows, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Set by a stage-1 wave whose wait for a neighbour gave up (see k_stage1);
// read and cleared by pf_device_status.
__device__ int g_s1_wait_error;
// Polls a stage-1 wait makes before it gives up (pf_set_wait_limit; passed
// to every launch): about 4 M polls of 0.1-1 us, seconds, while a wait in a
// correct run lasts at most the few ms a neighbour needs for its next rows.
static int g_wait_limit = 1 << 22;

constexpr int kMaxG = 64;   // rows per exchange block, upper bound
// Rows of C prefetched ahead, and rows per unrolled step of the sweep: 8
// for the packed DP with the standard step set in FP32 (up to 4 columns per
// lane) and in FP64 built for at most 256 threads (2 columns per lane, up to
// 512 VGPRs), 4 otherwise (the wider state of the other variants would not
// fit their register budget with 8).
template <typename DT, int CPL, bool FULL, bool STD, int MAXT>
constexpr int prefetch_rows() {
  return (!FULL && STD &&
          ((sizeof(DT) == 4 && CPL <= 4) || (sizeof(DT) == 8 && CPL == 2 && MAXT <= 256)))
             ? 8
             : 4;
}
constexpr int kRing = 8;    // rows of wave-to-wave hand-over kept in LDS
// Left context of every wave in one LDS array of column pairs (the two
// columns left of a wave), the D pairs in ctx[0] and the X pairs in ctx[1]
// at the same index:
//   ring slot s of wave w (its last two columns of a row r, s = r % kRing)
//     at s * 16 + w (up to 16 waves);
//   exchange row q of the strip on the left at kCtxLx + q % kMaxG;
//   the tile's missing columns -2, -1 (D = +inf, X = 0) at kCtxNone.
constexpr int kCtxLx = kRing * 16;
constexpr int kCtxNone = kCtxLx + kMaxG;
template <typename DT>
struct alignas(2 * sizeof(DT)) Pair {
  DT a, b;
};

// Lane i receives lane i-1's value (DPP wave_shr:1); lane 0 keeps `old`.
__device__ __forceinline__ int shr1(int v, int old) {
  return __builtin_amdgcn_update_dpp(old, v, 0x138, 0xf, 0xf, false);
}
__device__ __forceinline__ float shr1(float v, float old) {
  return __int_as_float(shr1(__float_as_int(v), __float_as_int(old)));
}
__device__ __forceinline__ double shr1(double v, double old) {
  const long long vi = __double_as_longlong(v), oi = __double_as_longlong(old);
  const int lo = shr1((int)(vi & 0xffffffffll), (int)(oi & 0xffffffffll));
  const int hi = shr1((int)(vi >> 32), (int)(oi >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__device__ __forceinline__ double dabs(double v) { return fabs(v); }
__device__ __forceinline__ float dabs(float v) { return fabsf(v); }
// two floats as bf16 (round to nearest even, as __float2bfloat16) in one
// packed conversion, lo in the low half
typedef __bf16 s1bf16x2_t __attribute__((ext_vector_type(2)));
typedef float s1floatx2_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t bf16_pair(float lo, float hi) {
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(s1floatx2_t{lo, hi}, s1bf16x2_t));
}
__device__ __forceinline__ double copy_sign(double mag, double sgn) { return copysign(mag, sgn); }
__device__ __forceinline__ float copy_sign(float mag, float sgn) { return copysignf(mag, sgn); }

// 2 * acc + (this lane's bit of mask): one v_addc_co_u32 with the lane mask
// as carry-in.  Appending bits this way, MSB first, and reversing the word
// (v_bfrev_b32) puts the first appended bit at bit 0.
__device__ __forceinline__ uint32_t shift_in(uint32_t acc, uint64_t mask) {
  uint32_t out;
  uint64_t carry;
  asm("v_addc_co_u32_e64 %0, %1, %2, %2, %3" : "=v"(out), "=s"(carry) : "v"(acc), "s"(mask));
  return out;
}

// Start encoding P of a cell from its signed path length X (see below).
template <typename DT>
__device__ __forceinline__ int p_of(DT X, int rc) {
  return (X >= (DT)0) ? rc - (int)X : -(rc + (int)X);
}

// State of a cell: D and X = +-M, where M = row + col - |P| is the
// Manhattan length of the path from its start (all small integers, exact in
// floating point) and the sign is that of P (+ for a start on row 0, - for a
// start on column 0).  The reference's mdist of a step from a predecessor is
// then |X_pred| + dr + dc, and the new cell inherits X = sign(X_pred) * mdist
// except when the predecessor is a boundary cell, where P is re-derived from
// the predecessor's coordinates (FlexDTW.py:1359-1364).
// Threads per workgroup the packed kernels are compiled for: 2 columns per
// lane up to 16 waves, 4 columns up to 8, 8 columns up to 4 (larger register
// budgets per lane).  The full-output kernels (a single tile, API path) take
// up to 16 waves for every width.
template <int CPL, bool FULL>
struct MaxThreads {
  static constexpr int value = (FULL || CPL <= 2) ? 1024 : (CPL == 4 ? 512 : 256);
};

typedef __bf16 s1bf16x8_t __attribute__((ext_vector_type(8)));
typedef float s1floatx16_t __attribute__((ext_vector_type(16)));

// FEAT: the cost is not read from HBM but computed inside the sweep from
// BF16 frame-major features (Fa: rows' frames, Fb: columns' frames, Kp
// padded to 16; tile table columns 0/1 hold the tile's first row / column
// frame).  Every 32 rows each wave computes the 32 x (64*CPL) block of C of
// its columns on v_mfma_f32_32x32x16_bf16 (the operand layout and rounding
// of the MFMA cost kernel, so the values are those of the BF16 cost matrix)
// into a wave-private LDS block, which the row sweep then reads.  One block
// buffer suffices: the C prefetch kPF rows ahead finishes reading block b
// one row before it starts reading block b+1, and block b+1 is computed in
// between.
// MAXT bounds the workgroup size, and with it the registers per lane: a
// 1024-thread workgroup (16 waves, 4 per SIMD) leaves 128 VGPRs a lane.
template <int CPL, typename CT, typename DT, bool FULL, bool STD, bool FEAT = false,
          int MAXT = MaxThreads<CPL, FULL>::value>
__global__ __launch_bounds__(MAXT) void k_stage1(
    const CT *__restrict__ Cb, const TileDesc *__restrict__ tiles, TileDesc td0, StepSet ss,
    int64_t Lmax, int S, int G, int *__restrict__ flags, Xch<DT> *__restrict__ xch,
    uint32_t *__restrict__ Bp, double *__restrict__ eD, int32_t *__restrict__ eP,
    double *__restrict__ eC, double *__restrict__ Df, int32_t *__restrict__ Pf,
    int8_t *__restrict__ Bf, const __hip_bfloat16 *__restrict__ Fa,
    const __hip_bfloat16 *__restrict__ Fb, int Kp, int spin_limit) {
  static_assert(CPL >= 2, "a lane needs two columns of left context");
  static_assert(!FEAT || (!FULL && sizeof(DT) == 4), "features form: packed FP32 DP only");
  extern __shared__ __attribute__((aligned(16))) __hip_bfloat16 cdyn[];
  __shared__ Pair<DT> ctx[2][kCtxNone + 1];
  __shared__ int prog[16];
  __shared__ int sh_item;

  // Work item from the ticket counter in strip order.
  int64_t item;
  if (S > 1) {
    // flags[-1] is the ticket counter (see pf_stage1).
    if (threadIdx.x == 0) sh_item = atomicAdd(&flags[-1], 1);
    __syncthreads();
    item = __builtin_amdgcn_readfirstlane(sh_item);
  } else {
    item = blockIdx.x;
  }
  const int64_t t = item / S;
  const int s = (int)(item - t * S);
  // the tile's geometry is the same for every lane: scalar registers, so
  // per-row conditions on R and W are scalar branches
  const TileDesc td = wave_uniform(tiles ? tiles[t] : td0);
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int SW = (int)blockDim.x * CPL;
  const int R = (int)td.R, W = (int)td.W;
  const int cbeg = s * SW;
  if (cbeg >= W) return;  // idle strip of a narrow tile (nobody waits on it)
  const bool producer = (S > 1) && (cbeg + SW < W) && (tid == (int)blockDim.x - 1);
  const bool consumer = (s > 0);
  int *fin = consumer ? flags + t * S + s - 1 : nullptr;
  int *fout = producer ? flags + t * S + s : nullptr;
  const Xch<DT> *xin = consumer ? xch + (t * S + s - 1) * Lmax : nullptr;
  Xch<DT> *xout = producer ? xch + (t * S + s) * Lmax : nullptr;

  const CT *__restrict__ C = Cb + (FEAT ? 0 : td.c_off);
  const int64_t ldc = td.ldc;
  // This tile's edge records: the row edge (last row of D and P, first row
  // of C) at e_off + c, the column edge (last column of D and P, first
  // column of C) at e_off + Lmax + r. Addressed from the column-edge base,
  // so one pointer per array stays live in the sweep.
  double *const eDc = FULL ? nullptr : eD + td.e_off + Lmax;
  int32_t *const ePc = FULL ? nullptr : eP + td.e_off + Lmax;
  double *const eCc = FULL ? nullptr : eC + td.e_off + Lmax;
  uint32_t *const Bt = FULL ? nullptr : Bp + td.b_off;
  // features of this tile's first row and column
  const __hip_bfloat16 *const Fat = FEAT ? Fa + td.c_off * (int64_t)Kp : nullptr;
  const __hip_bfloat16 *const Fbt = FEAT ? Fb + td.ldc * (int64_t)Kp : nullptr;
  const int c0 = cbeg + tid * CPL;
  const int jlast = W - 1 - c0;  // this lane's index of the tile's last column
  const bool own_last = (jlast >= 0 && jlast < CPL);
  // the wave holding the tile's last column (wave-uniform)
  const int wcol0 = cbeg + wave * 64 * CPL;
  const bool own_last_wave = (W - 1 >= wcol0) && (W - 1 < wcol0 + 64 * CPL);
  const DT INF = dp_inf<DT>();

  DT wk[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) wk[k] = (DT)ss.w[k];

  DT D1[CPL], X1[CPL], D2[CPL], X2[CPL];
  uint32_t bacc[CPL];
  constexpr int kPF = prefetch_rows<DT, CPL, FULL, STD, MAXT>();
  DT Cq[kPF][CPL];
  // Left context: columns c0-2 (a), c0-1 (b) of rows r-1 (1) and r-2 (2).
  DT LD2a = INF, LD2b = INF, LX2a = 0, LX2b = 0;

  // Column of each lane slot clamped into the tile: loads are unconditional
  // (no branch around them), values of columns past the tile are unused.
  int cl[CPL];
#pragma unroll
  for (int j = 0; j < CPL; ++j) cl[j] = (c0 + j < W) ? c0 + j : W - 1;

  // features form: this wave's block of C, 32 rows x (64*CPL) columns at a
  // row stride of kCS = 64*CPL + 4 elements (an odd multiple of 8 bytes:
  // 32 lanes writing 8 bytes each to 32 rows hit distinct banks)
  constexpr int kWC = 64 * CPL;
  constexpr int kCS = kWC + 4;
  const bool has_col0 = (cbeg == 0) && (wave == 0);  // wave-uniform
  __hip_bfloat16 *cblk = cdyn + (FEAT ? (int64_t)(tid >> 6) * 32 * kCS : 0);
  // Kp == 16: one MFMA per 32 x 32 sub-block, with the column features as
  // the A operand (the same for every block of the tile: loaded once) and
  // the row features as B, so that a lane holds 4 adjacent columns of one
  // row; the row features of the next block are loaded a block ahead.
  const int f_lr = lane & 31, f_kh = (lane >> 5) * 8;
  s1bf16x8_t fcol[FEAT ? kWC / 32 : 1], frow_next;
  auto load_frow = [&](const int b) {
    const int ar = (32 * b + f_lr < R) ? 32 * b + f_lr : R - 1;
    return *reinterpret_cast<const s1bf16x8_t *>(Fat + ar * (int64_t)Kp + f_kh);
  };
  if constexpr (FEAT) {
    if (Kp == 16) {
      const int wc0 = cbeg + (tid >> 6) * kWC;
#pragma unroll
      for (int cb = 0; cb < kWC / 32; ++cb) {
        const int gc = (wc0 + cb * 32 + f_lr < W) ? wc0 + cb * 32 + f_lr : W - 1;
        // negated: the MFMA started at 1 then yields 1 - a.b (see fill_block)
        const uint4 u = *reinterpret_cast<const uint4 *>(Fbt + gc * (int64_t)16 + f_kh);
        fcol[cb] = __builtin_bit_cast(
            s1bf16x8_t, uint4{u.x ^ 0x80008000u, u.y ^ 0x80008000u, u.z ^ 0x80008000u,
                              u.w ^ 0x80008000u});
      }
      frow_next = load_frow(0);
    }
  }
  auto fill_block = [&](const int b) {
    if constexpr (FEAT) {
      const int lr = f_lr, kh = f_kh;
      if (Kp == 16) {
        const s1bf16x8_t a = frow_next;
        if (32 * (b + 1) < R) frow_next = load_frow(b + 1);
#pragma unroll
        for (int cb = 0; cb < kWC / 32; ++cb) {
          // D[column][row]: lane holds columns 8g + 4(lane>>5) + 0..3 of
          // sub-block row lr, g = 0..3
          s1floatx16_t one;
#pragma unroll
          for (int v = 0; v < 16; ++v) one[v] = 1.0f;
          const s1floatx16_t acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fcol[cb], a, one, 0, 0, 0);
#pragma unroll
          for (int g = 0; g < 4; ++g) {
            const uint32_t lo = bf16_pair(acc[4 * g], acc[4 * g + 1]);
            const uint32_t hi = bf16_pair(acc[4 * g + 2], acc[4 * g + 3]);
            *reinterpret_cast<uint2 *>(&cblk[lr * kCS + cb * 32 + g * 8 + (lane >> 5) * 4]) =
                uint2{lo, hi};
          }
          // one accumulator live at a time (register budget of the sweep)
          __builtin_amdgcn_sched_barrier(0);
        }
      } else {
        const int ar = (32 * b + lr < R) ? 32 * b + lr : R - 1;
        const __hip_bfloat16 *ap = Fat + ar * (int64_t)Kp + kh;
        const int wc0 = cbeg + (tid >> 6) * kWC;
#pragma unroll
        for (int cb = 0; cb < kWC / 32; ++cb) {
          const int gc = (wc0 + cb * 32 + lr < W) ? wc0 + cb * 32 + lr : W - 1;
          const __hip_bfloat16 *bp = Fbt + gc * (int64_t)Kp + kh;
          s1floatx16_t acc;
#pragma unroll
          for (int v = 0; v < 16; ++v) acc[v] = 1.0f;
          // k-chunks four at a time: the 8 operand loads of a group are in
          // flight together (chunks past Kp read chunk 0 and are skipped)
          for (int kk0 = 0; kk0 < Kp; kk0 += 64) {
            uint4 u[4];
            s1bf16x8_t bv[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              const int kk = (kk0 + 16 * q < Kp) ? kk0 + 16 * q : 0;
              u[q] = *reinterpret_cast<const uint4 *>(ap + kk);
              bv[q] = *reinterpret_cast<const s1bf16x8_t *>(bp + kk);
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              if (kk0 + 16 * q < Kp) {
                const s1bf16x8_t av = __builtin_bit_cast(
                    s1bf16x8_t, uint4{u[q].x ^ 0x80008000u, u[q].y ^ 0x80008000u,
                                      u[q].z ^ 0x80008000u, u[q].w ^ 0x80008000u});
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv[q], acc, 0, 0, 0);
              }
            }
          }
#pragma unroll
          for (int v = 0; v < 16; ++v)
            cblk[((v >> 2) * 8 + (lane >> 5) * 4 + (v & 3)) * kCS + cb * 32 + lr] =
                __float2bfloat16(acc[v]);
        }
      }
      // the block is read by other lanes of this wave
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      // column 0 of the tile (stage 2's start cells): one 32-lane store per
      // block instead of one single-lane store per row
      if (has_col0 && lane < 32 && 32 * b + lane < R)
        eCc[32 * b + lane] = (double)__bfloat162float(cblk[lane * kCS]);
    }
  };
  // C of row `row`, lane slot j (a column clamped into the tile)
  auto c_at = [&](const int row, const int j) -> DT {
    if constexpr (FEAT) {
      return (DT)__bfloat162float(cblk[(row & 31) * kCS + lane * CPL + j]);
    } else {
      return load_c<DT>(C + (int64_t)row * ldc + cl[j]);
    }
  };

  // Row 0: boundary cells, P = 0, X = +col.
  fill_block(0);
#pragma unroll
  for (int j = 0; j < CPL; ++j) {
    const int c = c0 + j;
    const DT v = c_at(0, j);
    D1[j] = (c < W) ? v : INF;
    X1[j] = (DT)c;
    D2[j] = INF;
    X2[j] = 0;
    bacc[j] = 0u;
    if (c < W) {
      if constexpr (FULL) {
        Df[c] = (double)D1[j];
        Pf[c] = 0;
        Bf[c] = 0;
      } else {
        // first row and first column of C (the start cells stage 2 reads)
        eCc[c - Lmax] = (double)v;
        if (c == 0) eCc[0] = (double)v;
        if (c == W - 1) {
          eDc[0] = (double)D1[j];
          ePc[0] = 0;
        }
        if (R == 1) {
          eDc[c - Lmax] = (double)D1[j];
          ePc[c - Lmax] = 0;
          Bt[c] = 0u;
        }
      }
    }
  }
#pragma unroll
  for (int u = 0; u < kPF; ++u) {
    const int row = (1 + u < R) ? 1 + u : R - 1;
#pragma unroll
    for (int j = 0; j < CPL; ++j) Cq[u][j] = c_at(row, j);
  }
  if (producer) {
    xch_put(xout, Xch<DT>{D1[CPL - 2], D1[CPL - 1], X1[CPL - 2], X1[CPL - 1]});
    if (R == 1) xch_publish(fout, 1);
  }
  if (tid == 0) {
    ctx[0][kCtxNone] = Pair<DT>{INF, INF};
    ctx[1][kCtxNone] = Pair<DT>{0, 0};
  }
  if (lane == 63) {
    ctx[0][wave] = Pair<DT>{D1[CPL - 2], D1[CPL - 1]};
    ctx[1][wave] = Pair<DT>{X1[CPL - 2], X1[CPL - 1]};
    prog[wave] = 0;
  }
  const int nwaves = (int)blockDim.x >> 6;
  const bool producer_wave = (S > 1) && (cbeg + SW < W) && (wave == nwaves - 1);
  int seen_left = 0, seen_right = 0;  // last observed progress of the neighbour waves
  // Rows up to which no synchronisation is needed. Row r needs row r-1 of
  // the wave on the left (waves > 0); wave 0 of a consumer strip stages the
  // next exchange block at its first row (rows 1 .. G are staged below);
  // row r overwrites the ring slot of row r - kRing, read by the wave on the
  // right. Waves without that neighbour never wait.
  constexpr int kNever = 0x7fffffff;
  int sync_ok = (wave > 0) ? seen_left + 1 : ((consumer && 1 + G < R) ? G : kNever);
  int right_ok = (wave + 1 < nwaves) ? seen_right + kRing - 1 : kNever;
  // rows 1-2 of every wave and all rows of the wave holding columns 0-2
  // take the boundary rules
  const int border_rows = (wcol0 <= 2) ? kNever : 2;
  // Left context of row r (row r-1 of the two columns left of the wave):
  // for waves > 0 the ring slot (r-1) % kRing of the wave on the left, at
  // ((r-1) % kRing) * 16 + ctx_base; otherwise at ((r-1) & ctx_mask) +
  // ctx_base, for wave 0 of a strip with a strip on its left the exchange
  // row r-1, else the tile's missing columns.
  int ctx_mask, ctx_base;
  if (wave > 0) {
    ctx_mask = 0;  // ring slot (r-1) % kRing, from the row's slot
    ctx_base = wave - 1;
  } else if (consumer) {
    ctx_mask = kMaxG - 1;
    ctx_base = kCtxLx;
  } else {
    ctx_mask = 0;
    ctx_base = kCtxNone;
  }
  // next row r with (r + 1) % G == 0, from row 1 on: the producer publishes
  // its progress there (a running value, no division per row)
  int next_pub = G - 1;
  while (next_pub < 1) next_pub += G;
  // Rows 0 .. G-1 of the strip on the left feed rows 1 .. G.
  // Every wait polls at most spin_limit times.  A wave whose wait gives up
  // marks sync_ok and right_ok with kGaveUp (no wait is entered again: rows
  // stay below it) and finishes its rows on whatever values it has, so every
  // wave still exits (neighbours waiting on it see its progress, or give up
  // too); at its end it records the error (g_s1_wait_error, read by
  // pf_device_status).
  constexpr int kGaveUp = kNever - 1;
  auto give_up = [&]() {
    sync_ok = kGaveUp;
    right_ok = kGaveUp;
  };
  if (consumer && wave == 0 && R > 1) {
    const int need = (G < R) ? G : R;
    for (int n = spin_limit;
         __builtin_amdgcn_readfirstlane(
             __hip_atomic_load(fin, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) < need;
         --n) {
      if (n <= 0) {
        give_up();
        break;
      }
      __builtin_amdgcn_s_sleep(2);
    }
    for (int q0 = lane; q0 < G; q0 += 64) {
      if (q0 < R) {
        const Xch<DT> v = xch_get(xin + q0);
        ctx[0][kCtxLx + q0] = Pair<DT>{v.d0, v.d1};
        ctx[1][kCtxLx + q0] = Pair<DT>{v.x0, v.x1};
      }
    }
  }
  __syncthreads();

  // One row of the sweep; Cc holds C of row r.
  // slot = r % kRing (a constant in the unrolled steps when kPF is a
  // multiple of kRing).
  auto row_step = [&](const int r, const int slot, const DT (&Cc)[CPL]) {
      // Row r-1 of columns c0-2, c0-1: from lane-1 by DPP; lane 0 from the
      // previous wave (LDS ring), the previous strip (exchange rows) or the
      // missing columns -2, -1 of the tile.
      if (r > sync_ok) {
        if (wave > 0) {
          // wait until the wave on the left has handed over row r-1
          for (int n = spin_limit;; --n) {
            if (n <= 0) {
              give_up();
              break;
            }
            seen_left = __builtin_amdgcn_readfirstlane(__hip_atomic_load(
                &prog[wave - 1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP));
            if (seen_left >= r - 1) break;
            __builtin_amdgcn_s_sleep(1);
          }
          if (sync_ok != kGaveUp) sync_ok = seen_left + 1;
        } else {
          // first row of an exchange block: wait for the strip on the left
          // and stage its rows r-1 .. r+G-2 (row r'-1 of every row r' of the
          // block)
          const int last = r + G - 1;
          const int need = (last < R) ? last : R;
          for (int n = spin_limit;
               __builtin_amdgcn_readfirstlane(
                   __hip_atomic_load(fin, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) < need;
               --n) {
            if (n <= 0) {
              give_up();
              break;
            }
            __builtin_amdgcn_s_sleep(2);
          }
          for (int q0 = lane; q0 < G; q0 += 64) {
            const int q = r - 1 + q0;
            if (q < R) {
              const Xch<DT> v = xch_get(xin + q);
              const int i = kCtxLx + (q & (kMaxG - 1));
              ctx[0][i] = Pair<DT>{v.d0, v.d1};
              ctx[1][i] = Pair<DT>{v.x0, v.x1};
            }
          }
          // lane 0 reads the staged rows from this row on
          __threadfence_block();
          if (sync_ok != kGaveUp) sync_ok = (r + G < R) ? r + G - 1 : kNever;
        }
      }
      // The source is the same for the whole wave (only lane 0's copy is
      // used): an unconditional LDS read at a wave-uniform index.
      const int ci = (wave > 0) ? ((slot - 1) & (kRing - 1)) * 16 + ctx_base
                                : ((r - 1) & ctx_mask) + ctx_base;
      const Pair<DT> oD = ctx[0][ci], oX = ctx[1][ci];
      const DT oD1a = oD.a, oD1b = oD.b, oX1a = oX.a, oX1b = oX.b;
      const DT LD1a = shr1(D1[CPL - 2], oD1a), LD1b = shr1(D1[CPL - 1], oD1b);
      const DT LX1a = shr1(X1[CPL - 2], oX1a), LX1b = shr1(X1[CPL - 1], oX1b);

      DT ND[CPL], NX[CPL];
      int code[CPL];
      // FP32 standard step set, packed output: the step codes as lane masks
      // (code 1: mb1, code 2: mb2), shifted into bacc without forming code
      constexpr bool kMaskCodes = sizeof(DT) == 4 && STD && !FULL;
      uint64_t mb1[CPL], mb2[CPL];
      bool amb = false;  // some cell of this lane needs the exact comparison
      // rows 1-2 and columns 0-2 of the tile: predecessors on the boundary
      const bool border = r <= border_rows;
#pragma unroll
      for (int j = 0; j < CPL; ++j) {
        const int c = c0 + j;
        const DT cval = Cc[j];
        DT ba, bm, bx;
        int bk;
        if constexpr (STD) {
          // Steps (1,1) w0, (1,2) w1, (2,1) w2, decided without division.
          const DT pD0 = j >= 1 ? D1[j - 1] : LD1b, pX0 = j >= 1 ? X1[j - 1] : LX1b;
          const DT pD1 = j >= 2 ? D1[j - 2] : (j == 1 ? LD1b : LD1a);
          const DT pX1 = j >= 2 ? X1[j - 2] : (j == 1 ? LX1b : LX1a);
          const DT pD2 = j >= 1 ? D2[j - 1] : LD2b, pX2 = j >= 1 ? X2[j - 1] : LX2b;
          const bool v1 = c >= 2, v2 = r >= 2;
          if constexpr (sizeof(DT) == 4) {
            // FP32 DP (precision form): fused multiply-add path costs and the
            // plain cross-product order a1/m1 < a0/m0 <=> a1*m0 < a0*m1.
            // A predecessor outside the tile (column -2/-1, row -1) has
            // D = +inf, so its cost a is +inf and a*m < a'*m' is false: the
            // v1 / v2 masks are implied and not evaluated.
            (void)v1;
            (void)v2;
            const DT a0 = __builtin_fmaf(cval, wk[0], pD0), m0 = dabs(pX0) + (DT)2;
            const DT a1 = __builtin_fmaf(cval, wk[1], pD1), m1 = dabs(pX1) + (DT)3;
            const DT a2 = __builtin_fmaf(cval, wk[2], pD2), m2 = dabs(pX2) + (DT)3;
            const bool t1 = a1 * m0 < a0 * m1;
            ba = t1 ? a1 : a0;
            bm = t1 ? m1 : m0;
            bx = t1 ? pX1 : pX0;
            bk = t1 ? 1 : 0;
            const bool t2 = a2 * bm < ba * m2;
            if constexpr (kMaskCodes) {
              mb2[j] = __builtin_amdgcn_ballot_w64(t2);
              mb1[j] = __builtin_amdgcn_ballot_w64(t1) & ~mb2[j];
            }
            ba = t2 ? a2 : ba;
            bm = t2 ? m2 : bm;
            bx = t2 ? pX2 : bx;
            bk = t2 ? 2 : bk;
          } else {
            const DT a0 = pD0 + cval * wk[0], m0 = dabs(pX0) + (DT)2;
            const DT a1 = pD1 + cval * wk[1], m1 = dabs(pX1) + (DT)3;
            const DT a2 = pD2 + cval * wk[2], m2 = dabs(pX2) + (DT)3;
            bool ok1, ok2;
            const bool t1 = fast_less(a1, m1, a0, m0, ok1) && v1;
            ba = t1 ? a1 : a0;
            bm = t1 ? m1 : m0;
            bx = t1 ? pX1 : pX0;
            bk = t1 ? 1 : 0;
            const bool t2 = fast_less(a2, m2, ba, bm, ok2) && v2;
            ba = t2 ? a2 : ba;
            bm = t2 ? m2 : bm;
            bx = t2 ? pX2 : bx;
            bk = t2 ? 2 : bk;
            // Column 0 and columns past the tile are not computed here.
            amb = amb || ((c >= 1) && (c < W) && ((v1 && !ok1) || (v2 && !ok2)));
          }
        } else {
          // Generic step set (runtime, K <= 4): exact comparisons.
          DT E1[CPL + 2], F1[CPL + 2], E2[CPL + 2], F2[CPL + 2];
          E1[0] = LD1a;
          E1[1] = LD1b;
          F1[0] = LX1a;
          F1[1] = LX1b;
          E2[0] = LD2a;
          E2[1] = LD2b;
          F2[0] = LX2a;
          F2[1] = LX2b;
#pragma unroll
          for (int q = 0; q < CPL; ++q) {
            E1[q + 2] = D1[q];
            F1[q + 2] = X1[q];
            E2[q + 2] = D2[q];
            F2[q + 2] = X2[q];
          }
          ba = INF;
          bm = 1;
          bx = 0;
          bk = 0;
          bool have = false;
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            if (k < ss.K) {
              const int dr = ss.dr[k], dc = ss.dc[k];
              const DT v1 = (dc == 0) ? E1[j + 2] : ((dc == 1) ? E1[j + 1] : E1[j]);
              const DT v2 = (dc == 0) ? E2[j + 2] : ((dc == 1) ? E2[j + 1] : E2[j]);
              const DT y1 = (dc == 0) ? F1[j + 2] : ((dc == 1) ? F1[j + 1] : F1[j]);
              const DT y2 = (dc == 0) ? F2[j + 2] : ((dc == 1) ? F2[j + 1] : F2[j]);
              const DT px = (dr == 1) ? y1 : y2;
              const DT a = ((dr == 1) ? v1 : v2) + cval * wk[k];
              const DT m = dabs(px) + (DT)(dr + dc);
              if (r - dr >= 0 && c - dc >= 0 && (!have || quot_less(a, m, ba, bm))) {
                ba = a;
                bm = m;
                bx = px;
                bk = k;
                have = true;
              }
            }
          }
        }
        ND[j] = ba;
        NX[j] = copy_sign(bm, bx);
        code[j] = bk;
      }
      if constexpr (STD && sizeof(DT) == 8) {
        if (amb) {
          // Near tie (rare): redo this lane's cells with the reference's
          // comparisons, IEEE division included.
#pragma unroll
          for (int j = 0; j < CPL; ++j) {
            const int c = c0 + j;
            const DT cval = Cc[j];
            const DT pD0 = j >= 1 ? D1[j - 1] : LD1b, pX0 = j >= 1 ? X1[j - 1] : LX1b;
            const DT pD1 = j >= 2 ? D1[j - 2] : (j == 1 ? LD1b : LD1a);
            const DT pX1 = j >= 2 ? X1[j - 2] : (j == 1 ? LX1b : LX1a);
            const DT pD2 = j >= 1 ? D2[j - 1] : LD2b, pX2 = j >= 1 ? X2[j - 1] : LX2b;
            const DT a0 = pD0 + cval * wk[0], m0 = dabs(pX0) + (DT)2;
            const DT a1 = pD1 + cval * wk[1], m1 = dabs(pX1) + (DT)3;
            const DT a2 = pD2 + cval * wk[2], m2 = dabs(pX2) + (DT)3;
            DT ba = a0, bm = m0, bx = pX0;
            int bk = 0;
            if (c >= 2 && quot_less(a1, m1, ba, bm)) {
              ba = a1;
              bm = m1;
              bx = pX1;
              bk = 1;
            }
            if (r >= 2 && quot_less(a2, m2, ba, bm)) {
              ba = a2;
              bm = m2;
              bx = pX2;
              bk = 2;
            }
            ND[j] = ba;
            NX[j] = copy_sign(bm, bx);
            code[j] = bk;
          }
        }
      }
      if constexpr (!FULL && !FEAT) {
        if (c0 == 0) eCc[r] = (double)Cc[0];
      }
      if (border) {
        // Cells whose predecessor is on row 0 or column 0, and column 0
        // itself (FlexDTW.py:1326-1327, :1359-1364).
#pragma unroll
        for (int j = 0; j < CPL; ++j) {
          const int c = c0 + j;
          int bdr, bdc;
          if constexpr (STD) {
            bdr = (code[j] == 2) ? 2 : 1;
            bdc = (code[j] == 1) ? 2 : 1;
          } else {
            bdr = ss.dr[code[j]];
            bdc = ss.dc[code[j]];
          }
          const int bpr = r - bdr, bpc = c - bdc;
          if (c == 0) {
            ND[j] = Cc[j];
            NX[j] = (DT)r;
            code[j] = 0;
          } else if (bpr == 0) {
            NX[j] = (DT)(r + c - bpc);
          } else if (bpc == 0) {
            NX[j] = -(DT)(r + c - bpr);
          }
        }
      }

      LD2a = LD1a;
      LD2b = LD1b;
      LX2a = LX1a;
      LX2b = LX1b;
#pragma unroll
      for (int j = 0; j < CPL; ++j) {
        D2[j] = D1[j];
        X2[j] = X1[j];
        D1[j] = ND[j];
        X1[j] = NX[j];
      }

      if constexpr (FULL) {
        const int64_t base = (int64_t)r * W;
#pragma unroll
        for (int j = 0; j < CPL; ++j) {
          const int c = c0 + j;
          if (c < W) {
            Df[base + c] = (double)ND[j];
            Pf[base + c] = p_of(NX[j], r + c);
            Bf[base + c] = (int8_t)code[j];
          }
        }
      } else {
        if constexpr (kMaskCodes) {
          // row r's code bits, low bit first: after 16 rows and a bit
          // reversal the code of row k sits at bits 2k, 2k+1.  Column 0
          // (code 0 by the boundary rule) is zeroed when its word is stored.
#pragma unroll
          for (int j = 0; j < CPL; ++j) bacc[j] = shift_in(shift_in(bacc[j], mb1[j]), mb2[j]);
        } else {
          const int sh = 2 * (r & 15);
#pragma unroll
          for (int j = 0; j < CPL; ++j) bacc[j] |= (uint32_t)code[j] << sh;
        }
        // full words here; the last partial word after the sweep
        if ((r & 15) == 15) {
          uint32_t *dst = Bt + (int64_t)(r >> 4) * W;
#pragma unroll
          for (int j = 0; j < CPL; ++j) {
            const int c = c0 + j;
            const uint32_t wv = kMaskCodes ? (c == 0 ? 0u : __builtin_bitreverse32(bacc[j])) : bacc[j];
            if (c < W) dst[c] = wv;
            bacc[j] = 0u;
          }
        }
        if (own_last_wave) {
          if (own_last) {
#pragma unroll
            for (int j = 0; j < CPL; ++j)
              if (j == jlast) {
                eDc[r] = (double)ND[j];
                ePc[r] = p_of(NX[j], r + W - 1);
              }
          }
        }
      }

      // the producer is lane 63 of the last wave; the branch on the wave
      // keeps the publish counter in a scalar register
      if (producer_wave) {
        if (lane == 63) {
          xch_put(xout + r, Xch<DT>{D1[CPL - 2], D1[CPL - 1], X1[CPL - 2], X1[CPL - 1]});
          if (r == next_pub || r == R - 1) xch_publish(fout, r + 1);
        }
        if (r == next_pub) next_pub += G;
      }
      // Hand row r to the wave on the right once it has read row r - kRing.
      if (r > right_ok) {
        for (int n = spin_limit;; --n) {
          if (n <= 0) {
            give_up();
            break;
          }
          seen_right = __builtin_amdgcn_readfirstlane(__hip_atomic_load(
              &prog[wave + 1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP));
          if (seen_right >= r - kRing + 1) break;
          __builtin_amdgcn_s_sleep(1);
        }
        if (right_ok != kGaveUp) right_ok = seen_right + kRing - 1;
      }
      if (lane == 63) {
        const int s0 = slot * 16 + wave;
        ctx[0][s0] = Pair<DT>{D1[CPL - 2], D1[CPL - 1]};
        ctx[1][s0] = Pair<DT>{X1[CPL - 2], X1[CPL - 1]};
        __hip_atomic_store(&prog[wave], r, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
      }
  };

  // Main sweep in blocks of kPF rows with no early exit, so the C loads for
  // kPF rows ahead stay in flight across rows (a branch or loop exit around
  // them would make the compiler wait for all outstanding loads).
  int rb = 1;
  for (; rb + kPF <= R; rb += kPF) {
    // rb = 1 (mod kPF): the row r = rb + u of each unrolled step has known
    // low bits, so row conditions with period 16 or 32 (back-pointer word
    // flush, C block fill, ring slot) fold to constants in most steps
    __builtin_assume((rb & (kPF - 1)) == 1);
#pragma unroll
    for (int u = 0; u < kPF; ++u) {
      const int r = rb + u;
      row_step(r, (kPF % kRing == 0) ? (1 + u) % kRing : r % kRing, Cq[u]);
      // Refill the slot only after its row is done, so the load targets the
      // same registers (no copy of an in-flight load at the loop edge).
      // Clamped to a valid row for global loads; the features form reads
      // its LDS block at row & 31, in bounds for any row (rows past the
      // tile are not used).
      const int rp = (FEAT || r + kPF < R) ? r + kPF : R - 1;
      if constexpr (FEAT) {
        // rows of the previous block were all read by earlier refills
        if (((r + kPF) & 31) == 0 && r + kPF < R) fill_block((r + kPF) >> 5);
      }
#pragma unroll
      for (int j = 0; j < CPL; ++j) Cq[u][j] = c_at(rp, j);
    }
  }
  // Remaining rows (fewer than kPF), already prefetched.
#pragma unroll
  for (int u = 0; u < kPF; ++u) {
    const int r = rb + u;
    if (r < R) row_step(r, r % kRing, Cq[u]);
  }
  // every lane stores the same word: one vector store
  if (sync_ok == kGaveUp) g_s1_wait_error = 1;
  if constexpr (!FULL) {
    // Last back-pointer word when it holds fewer than 16 rows.
    if (R > 1 && ((R - 1) & 15) != 15) {
      uint32_t *dst = Bt + (int64_t)((R - 1) >> 4) * W;
      // rows of the word: n = ((R-1) & 15) + 1, 2n appended bits
      const int n = ((R - 1) & 15) + 1;
#pragma unroll
      for (int j = 0; j < CPL; ++j) {
        const int c = c0 + j;
        const uint32_t wv =
            (sizeof(DT) == 4 && STD)
                ? (c == 0 ? 0u : __builtin_bitreverse32(bacc[j]) >> (32 - 2 * n))
                : bacc[j];
        if (c < W) dst[c] = wv;
      }
    }
    // Last row of the tile (edge 0); R == 1 was written with row 0.
    if (R > 1) {
#pragma unroll
      for (int j = 0; j < CPL; ++j) {
        const int c = c0 + j;
        if (c < W) {
          eDc[c - Lmax] = (double)D1[j];
          ePc[c - Lmax] = p_of(X1[j], R - 1 + c);
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Features-form sweep with four columns per lane (k_sweep4): the standard
// step set, FP32 DP, packed outputs, one strip per tile (W <= 256 * waves),
// K <= 16 KC (KC chunks of 16 features, up to 8).  It computes what k_stage1<2 or 4, bf16, float, false, true, true>
// computes, bit for bit (the same cell arithmetic and the same C values),
// with less work per row:
//  * four columns per lane: the four DPP moves and the per-row bookkeeping
//    serve four cells instead of two;
//  * C in 16-row blocks: a chain of KC v_mfma_f32_16x16x16_bf16 per 16
//    columns (the same dot products and chains as the 32x32x16 form, bit
//    for bit), accumulator started at 1 and the column features negated, so
//    the MFMA yields 1 - a.b and the epilogue only rounds to BF16; a wave's
//    block is 8.25 KB (two 8-wave workgroups per CU); with KC = 1 the column
//    features stay in registers, otherwise each fill loads them (L2);
//  * every LDS access is a per-lane base plus an immediate offset: rows run
//    16 to an unrolled step group, r % 16 known in each step;
//  * the left wave's last two columns of a row come as two 8-byte records
//    {D, D} and {X, X} of an 8-slot ring (column 0 of the ring holds the
//    tile's missing columns -2, -1 for wave 0), stored from the registers
//    they live in; progress is published every 2 rows;
//  * the tile's last column (D, X) is staged in LDS and written 16 rows at a
//    time; the first column of C is written from the block fill.
// Rows 1 and 2 take the boundary rules of k_stage1 in full; from row 3 on
// only columns 0..2 (lane 0 of wave 0) can have a predecessor on column 0,
// and their start encodings are constants of the chosen step.
typedef __bf16 s4bf16x4_t __attribute__((ext_vector_type(4)));
typedef float s4floatx4_t __attribute__((ext_vector_type(4)));

constexpr int kS4Rows = 16;                // rows per block of C
constexpr int kS4Cols = 256;               // columns per wave (64 lanes x 4)
constexpr int kS4Stride = kS4Cols + 8;     // bf16 per block row in LDS (528 B = 16 mod 256)
constexpr int kS4Ring = 8;                 // rows of wave-to-wave hand-over
constexpr int kS4MaxWaves = 16;
constexpr int kS4RingW = kS4MaxWaves + 1;  // ring columns; column w + 1 is written by wave w


__device__ __forceinline__ s4bf16x4_t s4_load4(const __hip_bfloat16 *p, uint32_t flip) {
  uint2 v = *reinterpret_cast<const uint2 *>(p);
  v.x ^= flip;
  v.y ^= flip;
  return __builtin_bit_cast(s4bf16x4_t, v);
}

template <int MAXT, int KC>
__global__ __launch_bounds__(MAXT) void k_sweep4(const TileDesc *__restrict__ tiles, StepSet ss,
                                                 int64_t Lmax, uint32_t *__restrict__ Bp,
                                                 double *__restrict__ eD,
                                                 int32_t *__restrict__ eP,
                                                 double *__restrict__ eC,
                                                 const __hip_bfloat16 *__restrict__ Fa,
                                                 const __hip_bfloat16 *__restrict__ Fb,
                                                 int spin_limit) {
  extern __shared__ __attribute__((aligned(16))) __hip_bfloat16 cdyn[];
  __shared__ float2 ringD[kS4Ring][kS4RingW], ringX[kS4Ring][kS4RingW];
  __shared__ int prog[kS4MaxWaves];
  __shared__ float2 edge_st[kS4Rows];

  const TileDesc td = wave_uniform(tiles[blockIdx.x]);
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  // the wave index as a lane value: LDS bases built from it stay in VGPRs,
  // and per-row offsets fold into the instructions' immediates
  const int vwave = tid >> 6;
  const int nwaves = (int)blockDim.x >> 6;
  const int R = (int)td.R, W = (int)td.W;
  const int c0 = tid * 4;
  const int wcol0 = wave * kS4Cols;
  const float INF = __builtin_huge_valf();
  const float w0 = (float)ss.w[0], w1 = (float)ss.w[1], w2 = (float)ss.w[2];

  double *const eDc = eD + td.e_off + Lmax;
  int32_t *const ePc = eP + td.e_off + Lmax;
  double *const eCc = eC + td.e_off + Lmax;
  uint32_t *const Bt = Bp + td.b_off;
  constexpr int Kp = 16 * KC;
  const __hip_bfloat16 *const Fat = Fa + td.c_off * Kp;
  const __hip_bfloat16 *const Fbt = Fb + td.ldc * Kp;

  // this wave's block of C and the lane's read / fill positions in it
  __hip_bfloat16 *const cblk = cdyn + vwave * (kS4Rows * kS4Stride);
  const __hip_bfloat16 *const crd = cblk + lane * 4;
  const int fr = lane & 15, fk = (lane >> 4) * 4;
  __hip_bfloat16 *const cwr = cblk + fr * kS4Stride + fk;

  const bool active = wcol0 < W;  // wave-uniform
  // column features (negated) of the wave's 16 column blocks (KC = 1: in
  // registers, loaded once), row features of the next block (KC chunks)
  s4bf16x4_t fcol[KC == 1 ? kS4Cols / 16 : 1], frow_next[KC];
  auto load_frow = [&](const int b) {
    const int ar = (16 * b + fr < R) ? 16 * b + fr : R - 1;
#pragma unroll
    for (int kc = 0; kc < KC; ++kc) frow_next[kc] = s4_load4(Fat + ar * Kp + kc * 16 + fk, 0u);
  };
  auto col_of = [&](const int cb) {
    return (wcol0 + cb * 16 + fr < W) ? wcol0 + cb * 16 + fr : W - 1;
  };
  if (active) {
    if constexpr (KC == 1) {
#pragma unroll
      for (int cb = 0; cb < kS4Cols / 16; ++cb)
        fcol[cb] = s4_load4(Fbt + col_of(cb) * 16 + fk, 0x80008000u);
    }
    load_frow(0);
  }
  // C of rows 16b .. 16b + 15 into the wave's block (inlined: as a call its
  // captures would go through scratch memory)
  // KC = 1: the row features are copied and the next block's load issued
  // first; KC > 1: column features in groups of up to 4 chunks, the next
  // block's row features loaded after the MFMAs have read them (registers).
  auto fill = [&](const int b) __attribute__((always_inline)) {
    s4bf16x4_t a0 = frow_next[0];
    if constexpr (KC == 1) {
      if (16 * (b + 1) < R) load_frow(b + 1);
    }
    const s4floatx4_t one = {1.0f, 1.0f, 1.0f, 1.0f};
#pragma unroll
    for (int cb = 0; cb < kS4Cols / 16; cb += 2) {
      s4floatx4_t x = one, y = one;
      if constexpr (KC == 1) {
        x = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(fcol[cb], a0, x, 0, 0, 0);
        y = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(fcol[cb + 1], a0, y, 0, 0, 0);
      } else {
        constexpr int KH = KC < 4 ? KC : 4;
        const int gx = col_of(cb), gy = col_of(cb + 1);
#pragma unroll
        for (int k0 = 0; k0 < KC; k0 += KH) {
          s4bf16x4_t cx[KH], cy[KH];
#pragma unroll
          for (int kc = 0; kc < KH; ++kc) {
            cx[kc] = s4_load4(Fbt + gx * Kp + (k0 + kc) * 16 + fk, 0x80008000u);
            cy[kc] = s4_load4(Fbt + gy * Kp + (k0 + kc) * 16 + fk, 0x80008000u);
          }
#pragma unroll
          for (int kc = 0; kc < KH; ++kc) {
            x = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(cx[kc], frow_next[k0 + kc], x, 0, 0, 0);
            y = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(cy[kc], frow_next[k0 + kc], y, 0, 0, 0);
          }
        }
      }
      *reinterpret_cast<uint2 *>(cwr + cb * 16) = uint2{bf16_pair(x[0], x[1]), bf16_pair(x[2], x[3])};
      *reinterpret_cast<uint2 *>(cwr + cb * 16 + 16) =
          uint2{bf16_pair(y[0], y[1]), bf16_pair(y[2], y[3])};
      __builtin_amdgcn_sched_barrier(0);
    }
    if constexpr (KC > 1) {
      if (16 * (b + 1) < R) load_frow(b + 1);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // column 0 of the tile (stage 2's start cells), 16 rows at once
    if (wave == 0 && lane < 16 && 16 * b + lane < R)
      eCc[16 * b + lane] = (double)__bfloat162float(cdyn[lane * kS4Stride]);
  };
  auto unpack = [](const uint2 q, float (&cv)[4]) {
    cv[0] = __uint_as_float(q.x << 16);
    cv[1] = __uint_as_float(q.x & 0xffff0000u);
    cv[2] = __uint_as_float(q.y << 16);
    cv[3] = __uint_as_float(q.y & 0xffff0000u);
  };

  // the lane and slot holding the tile's last column (wave-uniform wave)
  const bool own_last_wave = (W - 1 >= wcol0) && (W - 1 < wcol0 + kS4Cols);
  const int jl = (W - 1) & 3;
  const bool own_last = own_last_wave && (lane == ((W - 1 - wcol0) >> 2));
  // back-pointer words as 16-byte stores when every row of them is aligned
  const bool vec_ok = ((td.b_off | (int64_t)W) & 3) == 0;

  float D1[4], X1[4], D2[4], X2[4];
  uint32_t bacc[4];
  // row r - 2 of column c0 - 1 (the left context of step (2, 1))
  float LD2b = INF, LX2b = 0.0f;

  // Row 0: boundary cells, P = 0, X = +col.
  uint2 cq = uint2{0u, 0u};
  if (active) {
    fill(0);
    cq = *reinterpret_cast<const uint2 *>(crd);
  }
  {
    float cv[4];
    unpack(cq, cv);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int c = c0 + j;
      D1[j] = (c < W) ? cv[j] : INF;
      X1[j] = (float)c;
      D2[j] = INF;
      X2[j] = 0.0f;
      bacc[j] = 0u;
      if (c < W) {
        eCc[c - Lmax] = (double)cv[j];
        if (R == 1) {
          eDc[c - Lmax] = (double)D1[j];
          ePc[c - Lmax] = 0;
          Bt[c] = 0u;
        }
      }
    }
    if (own_last) edge_st[0] = float2{D1[jl], X1[jl]};
  }
  for (int q = tid; q < kS4Ring; q += (int)blockDim.x) {
    ringD[q][0] = float2{INF, INF};
    ringX[q][0] = float2{0.0f, 0.0f};
  }
  if (lane == 63) {
    ringD[0][vwave + 1] = float2{D1[2], D1[3]};
    ringX[0][vwave + 1] = float2{X1[2], X1[3]};
    prog[wave] = 0;
  }
  __syncthreads();
  if (!active) return;

  const bool has_right = (wave + 1 < nwaves) && (wcol0 + kS4Cols < W);
  constexpr int kNever = 0x7fffffff;
  constexpr int kGaveUp = kNever - 1;
  int seen_left = 0, seen_right = 0;
  int sync_ok = (wave > 0) ? 1 : kNever;
  int right_ok = has_right ? kS4Ring - 1 : kNever;
  auto give_up = [&]() {
    sync_ok = kGaveUp;
    right_ok = kGaveUp;
  };
  // (D, X) of the last column of rows rb .. rb + n - 1, staged in edge_st
  auto flush_edges = [&](const int rb, const int n) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (lane < n) {
      const float2 v = edge_st[lane];
      eDc[rb + lane] = (double)v.x;
      ePc[rb + lane] = p_of(v.y, rb + lane + W - 1);
    }
  };
  if (R > 1) cq = *reinterpret_cast<const uint2 *>(crd + kS4Stride);

  // One row; U = r % 16, BORDER: rows 1 and 2 (the full boundary rules).
  auto step = [&](const int r, auto Uc, auto Bc) {
    constexpr int U = decltype(Uc)::value;
    constexpr bool BORDER = decltype(Bc)::value;
    if (r > sync_ok) {
      // wait until the wave on the left has handed over row r - 1
      for (int n = spin_limit;; --n) {
        if (n <= 0) {
          give_up();
          break;
        }
        seen_left = __builtin_amdgcn_readfirstlane(__hip_atomic_load(
            &prog[wave - 1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP));
        if (seen_left >= r - 1) break;
        __builtin_amdgcn_s_sleep(1);
      }
      if (sync_ok != kGaveUp) sync_ok = seen_left + 1;
    }
    const float2 rd = ringD[(U + kS4Ring - 1) % kS4Ring][vwave];
    const float2 rx = ringX[(U + kS4Ring - 1) % kS4Ring][vwave];
    const float LD1a = shr1(D1[2], rd.x), LD1b = shr1(D1[3], rd.y);
    const float LX1a = shr1(X1[2], rx.x), LX1b = shr1(X1[3], rx.y);
    float cv[4];
    unpack(cq, cv);
    float ND[4], NX[4];
    uint64_t mb1[4], mb2[4];
    bool T1[4], T2[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const float pD0 = j >= 1 ? D1[j - 1] : LD1b, pX0 = j >= 1 ? X1[j - 1] : LX1b;
      const float pD1 = j >= 2 ? D1[j - 2] : (j == 1 ? LD1b : LD1a);
      const float pX1 = j >= 2 ? X1[j - 2] : (j == 1 ? LX1b : LX1a);
      const float pD2 = j >= 1 ? D2[j - 1] : LD2b, pX2 = j >= 1 ? X2[j - 1] : LX2b;
      const float a0 = __builtin_fmaf(cv[j], w0, pD0), m0 = fabsf(pX0) + 2.0f;
      const float a1 = __builtin_fmaf(cv[j], w1, pD1), m1 = fabsf(pX1) + 3.0f;
      const float a2 = __builtin_fmaf(cv[j], w2, pD2), m2 = fabsf(pX2) + 3.0f;
      const bool t1 = a1 * m0 < a0 * m1;
      float ba = t1 ? a1 : a0, bm = t1 ? m1 : m0, bx = t1 ? pX1 : pX0;
      const bool t2 = a2 * bm < ba * m2;
      mb2[j] = __builtin_amdgcn_ballot_w64(t2);
      mb1[j] = __builtin_amdgcn_ballot_w64(t1) & ~mb2[j];
      T1[j] = t1;
      T2[j] = t2;
      ND[j] = t2 ? a2 : ba;
      bm = t2 ? m2 : bm;
      bx = t2 ? pX2 : bx;
      NX[j] = copysignf(bm, bx);
    }
    if constexpr (BORDER) {
      // Cells whose predecessor is on row 0 or column 0, and column 0 itself
      // (FlexDTW.py:1326-1327, :1359-1364).
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int c = c0 + j;
        const int code = T2[j] ? 2 : (T1[j] ? 1 : 0);
        const int bpr = r - ((code == 2) ? 2 : 1), bpc = c - ((code == 1) ? 2 : 1);
        if (c == 0) {
          ND[j] = cv[j];
          NX[j] = (float)r;
        } else if (bpr == 0) {
          NX[j] = (float)(r + c - bpc);
        } else if (bpc == 0) {
          NX[j] = -(float)(r + c - bpr);
        }
      }
    } else {
      if (wave == 0 && lane == 0) {
        // r >= 3: column 0 is C with X = r; column 1 always steps from
        // column 0 (P = -(r - dr), X = -(1 + dr)); column 2 does with step
        // (1, 2) (P = -(r - 1), X = -3)
        ND[0] = cv[0];
        NX[0] = (float)r;
        NX[1] = T2[1] ? -3.0f : -2.0f;
        if (T1[2] && !T2[2]) NX[2] = -3.0f;
      }
    }
    LD2b = LD1b;
    LX2b = LX1b;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      D2[j] = D1[j];
      X2[j] = X1[j];
      D1[j] = ND[j];
      X1[j] = NX[j];
      bacc[j] = shift_in(shift_in(bacc[j], mb1[j]), mb2[j]);
    }
    if (own_last_wave) {
      if (own_last) {
        float2 v;
        if (jl == 0) v = float2{ND[0], NX[0]};
        else if (jl == 1) v = float2{ND[1], NX[1]};
        else if (jl == 2) v = float2{ND[2], NX[2]};
        else v = float2{ND[3], NX[3]};
        edge_st[U] = v;
      }
    }
    if constexpr (U == 15) {
      // back-pointer words of rows r - 15 .. r (column 0: code 0)
      uint32_t *dst = Bt + (int64_t)(r >> 4) * W;
      uint32_t wv[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        wv[j] = (c0 + j == 0) ? 0u : __builtin_bitreverse32(bacc[j]);
        bacc[j] = 0u;
      }
      if (vec_ok) {
        if (c0 < W) *reinterpret_cast<uint4 *>(dst + c0) = uint4{wv[0], wv[1], wv[2], wv[3]};
      } else {
#pragma unroll
        for (int j = 0; j < 4; ++j)
          if (c0 + j < W) dst[c0 + j] = wv[j];
      }
      if (own_last_wave) flush_edges(r - 15, 16);
      if (r + 1 < R) fill((r + 1) >> 4);
    }
    cq = *reinterpret_cast<const uint2 *>(crd + ((U + 1) % kS4Rows) * kS4Stride);
    // Hand row r to the wave on the right once it has read row r - kS4Ring.
    if (r > right_ok) {
      for (int n = spin_limit;; --n) {
        if (n <= 0) {
          give_up();
          break;
        }
        seen_right = __builtin_amdgcn_readfirstlane(__hip_atomic_load(
            &prog[wave + 1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP));
        if (seen_right >= r - kS4Ring + 1) break;
        __builtin_amdgcn_s_sleep(1);
      }
      if (right_ok != kGaveUp) right_ok = seen_right + kS4Ring - 1;
    }
    if (lane == 63) {
      ringD[U % kS4Ring][vwave + 1] = float2{D1[2], D1[3]};
      ringX[U % kS4Ring][vwave + 1] = float2{X1[2], X1[3]};
      if constexpr (U % 2 == 1)
        __hip_atomic_store(&prog[vwave], r, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
  };

  using T = std::true_type;
  using F = std::false_type;
  if (R > 1) step(1, std::integral_constant<int, 1>{}, T{});
  if (R > 2) step(2, std::integral_constant<int, 2>{}, T{});
  // Full groups of 16 rows with no branch between the steps (the registers
  // then rotate through the group without copies), then the last rows of
  // the tile, each step guarded.  rb = 3 (mod 16): row rb + u has
  // r % 16 = (3 + u) % 16.
  int rb = 3;
  for (; rb + 16 <= R; rb += 16) {
    __builtin_assume((rb & 15) == 3);
#define PF_S4_STEP(u) step(rb + u, std::integral_constant<int, (3 + u) % 16>{}, F{});
    PF_S4_STEP(0) PF_S4_STEP(1) PF_S4_STEP(2) PF_S4_STEP(3) PF_S4_STEP(4) PF_S4_STEP(5)
    PF_S4_STEP(6) PF_S4_STEP(7) PF_S4_STEP(8) PF_S4_STEP(9) PF_S4_STEP(10) PF_S4_STEP(11)
    PF_S4_STEP(12) PF_S4_STEP(13) PF_S4_STEP(14) PF_S4_STEP(15)
#undef PF_S4_STEP
  }
  if (rb < R) {
    __builtin_assume((rb & 15) == 3);
#define PF_S4_STEP(u) \
  if (rb + u < R) step(rb + u, std::integral_constant<int, (3 + u) % 16>{}, F{});
    PF_S4_STEP(0) PF_S4_STEP(1) PF_S4_STEP(2) PF_S4_STEP(3) PF_S4_STEP(4) PF_S4_STEP(5)
    PF_S4_STEP(6) PF_S4_STEP(7) PF_S4_STEP(8) PF_S4_STEP(9) PF_S4_STEP(10) PF_S4_STEP(11)
    PF_S4_STEP(12) PF_S4_STEP(13) PF_S4_STEP(14)
#undef PF_S4_STEP
  }
  // the last row for the wave on the right
  if (lane == 63) __hip_atomic_store(&prog[wave], R - 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
  if (sync_ok == kGaveUp) g_s1_wait_error = 1;
  // the last-column edges of the last rows not yet written
  if (((R - 1) & 15) != 15 && own_last_wave) flush_edges((R - 1) & ~15, ((R - 1) & 15) + 1);
  if (R > 1) {
    // last back-pointer word when it holds fewer than 16 rows
    if (((R - 1) & 15) != 15) {
      uint32_t *dst = Bt + (int64_t)((R - 1) >> 4) * W;
      const int n = ((R - 1) & 15) + 1;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int c = c0 + j;
        const uint32_t wv = c == 0 ? 0u : __builtin_bitreverse32(bacc[j]) >> (32 - 2 * n);
        if (c < W) dst[c] = wv;
      }
    }
    // last row of the tile (edge 0)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int c = c0 + j;
      if (c < W) {
        eDc[c - Lmax] = (double)D1[j];
        ePc[c - Lmax] = p_of(X1[j], R - 1 + c);
      }
    }
  }
}

template <typename CT, typename DT, bool STD, int CPL>
static void launch_cpl(dim3 grid, dim3 block, hipStream_t st, const CT *C, const TileDesc *tiles,
                       const TileDesc &td0, const StepSet &ss, int64_t Lmax, int S, int G,
                       int *flags, void *xch, uint32_t *Bp, double *eD, int32_t *eP, double *eC,
                       double *Df, int32_t *Pf, int8_t *Bf, bool full) {
  Xch<DT> *x = static_cast<Xch<DT> *>(xch);
  if constexpr (sizeof(DT) == 8 && CPL == 2) {
    // FP64 sweep of up to 4 waves (the default shape): built for 256
    // threads, so its state fits in registers without spills
    if (!full && block.x <= 256) {
      hipLaunchKernelGGL((k_stage1<CPL, CT, DT, false, STD, false, 256>), grid, block, 0, st, C,
                         tiles, td0, ss, Lmax, S, G, flags, x, Bp, eD, eP, eC, Df, Pf, Bf,
                         nullptr, nullptr, 0, g_wait_limit);
      return;
    }
  }
  if (full)
    hipLaunchKernelGGL((k_stage1<CPL, CT, DT, true, STD>), grid, block, 0, st, C, tiles, td0, ss,
                       Lmax, S, G, flags, x, Bp, eD, eP, eC, Df, Pf, Bf, nullptr, nullptr, 0,
                       g_wait_limit);
  else
    hipLaunchKernelGGL((k_stage1<CPL, CT, DT, false, STD>), grid, block, 0, st, C, tiles, td0, ss,
                       Lmax, S, G, flags, x, Bp, eD, eP, eC, Df, Pf, Bf, nullptr, nullptr, 0,
                       g_wait_limit);
}

// Features form: LDS for the per-wave C blocks is dynamic (nw * 32 * (64*CPL + 4) bf16).
template <bool STD, int CPL>
static int launch_feat_cpl(dim3 grid, dim3 block, size_t lds, hipStream_t st,
                           const TileDesc *tiles, const StepSet &ss, int64_t Lmax, int S, int G,
                           int *flags, void *xch, uint32_t *Bp, double *eD, int32_t *eP,
                           double *eC, const __hip_bfloat16 *Fa, const __hip_bfloat16 *Fb, int Kp) {
  auto kern = k_stage1<CPL, __hip_bfloat16, float, false, STD, true>;
  if (lds > 64 * 1024)
    PF_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  const TileDesc td0{};
  hipLaunchKernelGGL(kern, grid, block, lds, st, static_cast<const __hip_bfloat16 *>(nullptr),
                     tiles, td0, ss, Lmax, S, G, flags, static_cast<Xch<float> *>(xch), Bp, eD, eP,
                     eC, nullptr, nullptr, nullptr, Fa, Fb, Kp, g_wait_limit);
  PF_CHECK_HIP(hipGetLastError());
  return PF_OK;
}

template <typename CT, typename DT>
static int launch_dp(int cpl, int nw, int S, int G, int64_t n_items, hipStream_t st, const CT *C,
                     const TileDesc *tiles, const TileDesc &td0, const StepSet &ss, int64_t Lmax,
                     int *flags, void *xch, uint32_t *Bp, double *eD, int32_t *eP, double *eC,
                     double *Df, int32_t *Pf, int8_t *Bf, bool full) {
  const bool std3 = ss.K == 3 && ss.dr[0] == 1 && ss.dc[0] == 1 && ss.dr[1] == 1 &&
                    ss.dc[1] == 2 && ss.dr[2] == 2 && ss.dc[2] == 1;
  const dim3 grid((unsigned)n_items), block((unsigned)(nw * 64));
#define PF_LAUNCH(STDV, CPLV)                                                                  \
  launch_cpl<CT, DT, STDV, CPLV>(grid, block, st, C, tiles, td0, ss, Lmax, S, G, flags, xch,  \
                                 Bp, eD, eP, eC, Df, Pf, Bf, full)
  if (std3) {
    if (cpl == 2) PF_LAUNCH(true, 2);
    else if (cpl == 4) PF_LAUNCH(true, 4);
    else PF_LAUNCH(true, 8);
  } else {
    if (cpl == 2) PF_LAUNCH(false, 2);
    else if (cpl == 4) PF_LAUNCH(false, 4);
    else PF_LAUNCH(false, 8);
  }
#undef PF_LAUNCH
  PF_CHECK_HIP(hipGetLastError());
  return PF_OK;
}

// Strip geometry: smallest columns-per-lane in {2,4,8} with S*NW*64*CPL >= W.
static int pick_cpl(int64_t W, int S, int nw) {
  for (int cpl = 2; cpl <= 8; cpl *= 2)
    if ((int64_t)S * nw * 64 * cpl >= W) return cpl;
  return 0;
}

__global__ void k_fill_i32(int *p, int v, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

}  // namespace pf

using namespace pf;

extern "C" int pf_stage1(int dtype, const void *d_C, const int64_t *d_tiles, int64_t ntiles,
                         int64_t Lmax, const int32_t *steps, const double *weights, int K,
                         int strips, int waves, int rows_per_block, int32_t *d_flags,
                         void *d_xch, uint32_t *d_Bpacked, double *d_edge_D, int32_t *d_edge_P,
                         double *d_edge_C, void *stream) {
  StepSet ss;
  int rc = make_stepset(steps, weights, K, &ss);
  if (rc) return rc;
  PF_REQUIRE(ntiles >= 0 && Lmax >= 2, "pf_stage1: bad tile count or tile length");
  PF_REQUIRE(strips >= 1 && waves >= 1 && waves <= 16, "pf_stage1: strips >= 1, 1 <= waves <= 16");
  PF_REQUIRE(rows_per_block >= 1 && rows_per_block <= kMaxG, "pf_stage1: 1 <= G <= 64");
  const int cpl = pick_cpl(Lmax, strips, waves);
  PF_REQUIRE(cpl > 0, "pf_stage1: strips*waves*64*8 must cover the tile length");
  PF_REQUIRE(waves * 64 <= (cpl == 2 ? 1024 : (cpl == 4 ? 512 : 256)),
             "pf_stage1: at most 16/8/4 waves for 2/4/8 columns per lane");
  PF_REQUIRE(Lmax <= 8192, "tile length above 8192 is not supported");
  if (ntiles == 0) return PF_OK;
  PF_REQUIRE(d_C && d_tiles && d_Bpacked && d_edge_D && d_edge_P && d_edge_C,
             "pf_stage1: null buffer");
  PF_REQUIRE(strips == 1 || (d_flags && d_xch), "pf_stage1: strips > 1 need flags and xch");
  const int64_t n_items = ntiles * strips;
  PF_REQUIRE(n_items < (1ll << 31), "pf_stage1: too many work items");
  hipStream_t st = as_stream(stream);
  const TileDesc *tiles = reinterpret_cast<const TileDesc *>(d_tiles);
  const TileDesc td0{};
  int *flags = nullptr;
  if (strips > 1) {
    // d_flags: [0] ticket counter, [1 .. n_items] per-strip published rows.
    hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)((n_items + 1 + 255) / 256)), dim3(256), 0,
                       st, d_flags, 0, n_items + 1);
    PF_CHECK_HIP(hipGetLastError());
    flags = d_flags + 1;
  }
  switch (dtype) {
    case PF_F64:
      return launch_dp<double, double>(cpl, waves, strips, rows_per_block, n_items, st,
                                       static_cast<const double *>(d_C), tiles, td0, ss, Lmax,
                                       flags, d_xch, d_Bpacked, d_edge_D, d_edge_P, d_edge_C,
                                       nullptr, nullptr, nullptr, false);
    case PF_F32:
      return launch_dp<float, float>(cpl, waves, strips, rows_per_block, n_items, st,
                                     static_cast<const float *>(d_C), tiles, td0, ss, Lmax, flags,
                                     d_xch, d_Bpacked, d_edge_D, d_edge_P, d_edge_C, nullptr,
                                     nullptr, nullptr, false);
    case PF_BF16:
      return launch_dp<__hip_bfloat16, float>(cpl, waves, strips, rows_per_block, n_items, st,
                                              static_cast<const __hip_bfloat16 *>(d_C), tiles,
                                              td0, ss, Lmax, flags, d_xch, d_Bpacked, d_edge_D,
                                              d_edge_P, d_edge_C, nullptr, nullptr, nullptr, false);
    default:
      PF_REQUIRE(false, "pf_stage1: unknown dtype");
  }
}

extern "C" int pf_stage1_features(const void *d_F1b, const void *d_F2b, int Kp,
                                  const int64_t *d_tiles, int64_t ntiles, int64_t Lmax,
                                  const int32_t *steps, const double *weights, int K, int strips,
                                  int waves, int rows_per_block, int32_t *d_flags, void *d_xch,
                                  uint32_t *d_Bpacked, double *d_edge_D, int32_t *d_edge_P,
                                  double *d_edge_C, void *stream) {
  StepSet ss;
  int rc = make_stepset(steps, weights, K, &ss);
  if (rc) return rc;
  PF_REQUIRE(ntiles >= 0 && Lmax >= 2 && Lmax <= 8192, "pf_stage1_features: bad tile length");
  PF_REQUIRE(Kp >= 16 && Kp % 16 == 0, "pf_stage1_features: Kp must be a positive multiple of 16");
  PF_REQUIRE(strips >= 1 && waves >= 1 && waves <= 16, "pf_stage1_features: 1 <= waves <= 16");
  PF_REQUIRE(rows_per_block >= 1 && rows_per_block <= kMaxG, "pf_stage1_features: 1 <= G <= 64");
  const int cpl = pick_cpl(Lmax, strips, waves);
  PF_REQUIRE(cpl == 2 || cpl == 4, "pf_stage1_features: strips*waves*64*4 must cover the tile");
  const bool std3 = ss.K == 3 && ss.dr[0] == 1 && ss.dc[0] == 1 && ss.dr[1] == 1 &&
                    ss.dc[1] == 2 && ss.dr[2] == 2 && ss.dc[2] == 1;
  // one strip of 4 columns per lane, standard steps, K <= 128: k_sweep4
  const bool sweep4 = std3 && (Kp == 16 || Kp == 32 || Kp == 64 || Kp == 128) &&
                      strips == 1 && cpl == 4;
  PF_REQUIRE(sweep4 || waves * 64 <= (cpl == 2 ? 1024 : 512), "pf_stage1_features: too many waves");
  const size_t lds = sweep4 ? (size_t)waves * kS4Rows * kS4Stride * sizeof(__hip_bfloat16)
                            : (size_t)waves * 32 * (64 * cpl + 4) * sizeof(__hip_bfloat16);
  PF_REQUIRE(lds <= 150 * 1024, "pf_stage1_features: C blocks exceed the LDS");
  if (ntiles == 0) return PF_OK;
  PF_REQUIRE(d_F1b && d_F2b && d_tiles && d_Bpacked && d_edge_D && d_edge_P && d_edge_C,
             "pf_stage1_features: null buffer");
  PF_REQUIRE(reinterpret_cast<uintptr_t>(d_F1b) % 16 == 0 &&
                 reinterpret_cast<uintptr_t>(d_F2b) % 16 == 0,
             "pf_stage1_features: features must be 16-byte aligned");
  PF_REQUIRE(strips == 1 || (d_flags && d_xch), "pf_stage1_features: strips > 1 need flags and xch");
  const int64_t n_items = ntiles * strips;
  PF_REQUIRE(n_items < (1ll << 31), "pf_stage1_features: too many work items");
  hipStream_t st = as_stream(stream);
  int *flags = nullptr;
  if (strips > 1) {
    hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)((n_items + 1 + 255) / 256)), dim3(256), 0,
                       st, d_flags, 0, n_items + 1);
    PF_CHECK_HIP(hipGetLastError());
    flags = d_flags + 1;
  }
  const TileDesc *tiles = reinterpret_cast<const TileDesc *>(d_tiles);
  const auto *Fa = static_cast<const __hip_bfloat16 *>(d_F1b);
  const auto *Fb = static_cast<const __hip_bfloat16 *>(d_F2b);
  const dim3 grid((unsigned)n_items), block((unsigned)(waves * 64));
  if (sweep4) {
    auto kern = Kp == 16 ? k_sweep4<1024, 1>
                         : (Kp == 32 ? k_sweep4<1024, 2>
                                     : (Kp == 64 ? k_sweep4<1024, 4> : k_sweep4<1024, 8>));
    if (lds > 64 * 1024)
      PF_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(kern, grid, block, lds, st, tiles, ss, Lmax, d_Bpacked, d_edge_D,
                       d_edge_P, d_edge_C, Fa, Fb, g_wait_limit);
    PF_CHECK_HIP(hipGetLastError());
    return PF_OK;
  }
#define PF_FEAT(STDV, CPLV)                                                                  \
  launch_feat_cpl<STDV, CPLV>(grid, block, lds, st, tiles, ss, Lmax, strips, rows_per_block, \
                              flags, d_xch, d_Bpacked, d_edge_D, d_edge_P, d_edge_C, Fa, Fb, Kp)
  if (std3) return cpl == 2 ? PF_FEAT(true, 2) : PF_FEAT(true, 4);
  return cpl == 2 ? PF_FEAT(false, 2) : PF_FEAT(false, 4);
#undef PF_FEAT
}

extern "C" int pf_set_wait_limit(int64_t polls, int64_t *previous) {
  PF_REQUIRE(polls >= 0 && polls <= 0x7fffffff, "pf_set_wait_limit: limit outside [0, 2^31)");
  if (previous) *previous = g_wait_limit;
  g_wait_limit = (int)polls;
  return PF_OK;
}

namespace pf {
int wait_limit() { return g_wait_limit; }
}  // namespace pf

extern "C" int pf_device_status(void *stream) {
  hipStream_t st = as_stream(stream);
  int word = 0, word2 = 0;
  PF_CHECK_HIP(hipMemcpyFromSymbolAsync(&word, HIP_SYMBOL(g_s1_wait_error), sizeof(int), 0,
                                        hipMemcpyDeviceToHost, st));
  PF_CHECK_HIP(hipStreamSynchronize(st));
  const int rc = take_s2_wait_error(st, &word2);
  if (rc != PF_OK) return rc;
  if (word2 != 0 && word == 0) {
    set_error("the stage-2 tile sweep gave up waiting for a tile (spin limit reached); its "
              "outputs are not valid");
    return PF_EHIP;
  }
  if (word != 0) {
    const int zero = 0;
    PF_CHECK_HIP(hipMemcpyToSymbolAsync(HIP_SYMBOL(g_s1_wait_error), &zero, sizeof(int), 0,
                                        hipMemcpyHostToDevice, st));
    PF_CHECK_HIP(hipStreamSynchronize(st));
    set_error("a stage-1 kernel gave up waiting for a neighbouring wave or strip "
              "(spin limit reached); its outputs are not valid");
    return PF_EHIP;
  }
  return PF_OK;
}

extern "C" int pf_stage1_workspace(int dtype, int64_t ntiles, int64_t Lmax, int strips,
                                   int64_t *flags_count, int64_t *xch_bytes) {
  PF_REQUIRE(flags_count && xch_bytes, "pf_stage1_workspace: null output");
  PF_REQUIRE(ntiles >= 0 && Lmax >= 2 && strips >= 1, "pf_stage1_workspace: bad arguments");
  const int64_t esz = dtype == PF_F64 ? (int64_t)sizeof(Xch<double>) : (int64_t)sizeof(Xch<float>);
  PF_REQUIRE(dtype == PF_F64 || dtype == PF_F32 || dtype == PF_BF16,
             "pf_stage1_workspace: unknown dtype");
  *flags_count = strips > 1 ? ntiles * strips + 1 : 0;
  *xch_bytes = strips > 1 ? ntiles * strips * Lmax * esz : 0;
  return PF_OK;
}

// Scratch of pf_stage1_noalloc: the strip-exchange counters and rows of one
// call, in device memory the library holds for the whole process, one copy
// per GPU.  Host code never uses the name `scratch` as an address: the
// device address of each GPU's copy comes from hipGetSymbolAddress
// (scratch_base) and only that is passed to kernels.
static __device__ __align__(256) unsigned char scratch[PF_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[kMaxDevices];

// Device address of the current GPU's scratch, looked up on the first call
// on that GPU (with it current) and kept.
static int scratch_base(void **out) {
  int dev = 0;
  PF_CHECK_HIP(hipGetDevice(&dev));
  PF_REQUIRE(dev >= 0 && dev < kMaxDevices, "pf_stage1_noalloc: device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void *p = nullptr;
    PF_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    PF_REQUIRE(p != nullptr, "pf_stage1_noalloc: no device address for the scratch array");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return PF_OK;
}

extern "C" int pf_stage1_noalloc(int dtype, const void *d_C, const int64_t *d_tiles,
                                 int64_t ntiles, int64_t Lmax, const int32_t *steps,
                                 const double *weights, int K, int strips, int waves,
                                 int rows_per_block, uint32_t *d_Bpacked, double *d_edge_D,
                                 int32_t *d_edge_P, double *d_edge_C, void *stream) {
  int32_t *flags = nullptr;
  void *xch = nullptr;
  if (strips > 1 && ntiles > 0) {
    int64_t nf = 0, nx = 0;
    int rc = pf_stage1_workspace(dtype, ntiles, Lmax, strips, &nf, &nx);
    if (rc) return rc;
    const int64_t fbytes = (nf * (int64_t)sizeof(int32_t) + 255) / 256 * 256;
    if (fbytes + nx > (int64_t)PF_NOALLOC_SCRATCH_BYTES) {
      set_error("pf_stage1_noalloc: the strip exchange needs " + std::to_string(fbytes + nx) +
                " bytes, above the " + std::to_string((int64_t)PF_NOALLOC_SCRATCH_BYTES) +
                "-byte scratch (use fewer strips or tiles per call, or pf_stage1)");
      return PF_EINVAL;
    }
    void *base = nullptr;
    rc = scratch_base(&base);
    if (rc) return rc;
    flags = static_cast<int32_t *>(base);
    xch = static_cast<unsigned char *>(base) + fbytes;
  }
  // The counters are cleared by pf_stage1's fill kernel (no memset).
  return pf_stage1(dtype, d_C, d_tiles, ntiles, Lmax, steps, weights, K, strips, waves,
                   rows_per_block, flags, xch, d_Bpacked, d_edge_D, d_edge_P, d_edge_C, stream);
}

extern "C" int pf_tile_full(const double *d_C, int64_t ldc, int64_t r0, int64_t r1, int64_t c0,
                            int64_t c1, const int32_t *steps, const double *weights, int K,
                            double *d_D, int32_t *d_P, int8_t *d_B, void *stream) {
  StepSet ss;
  int rc = make_stepset(steps, weights, K, &ss);
  if (rc) return rc;
  const int64_t R = r1 - r0, W = c1 - c0;
  PF_REQUIRE(R > 0 && W > 0, "Empty chunk");
  PF_REQUIRE(r0 >= 0 && c0 >= 0 && c1 <= ldc, "pf_tile_full: chunk outside C");
  PF_REQUIRE(W <= 8192, "tile width above 8192 is not supported");
  PF_REQUIRE(d_C && d_D && d_P && d_B, "pf_tile_full: null buffer");
  int cpl = W <= 2048 ? 2 : (W <= 4096 ? 4 : 8);
  const int nw = (int)(((W + cpl - 1) / cpl + 63) / 64);
  const TileDesc td{r0 * ldc + c0, ldc, R, W, 0, 0, 0, 0};
  return launch_dp<double, double>(cpl, nw, 1, 16, 1, as_stream(stream), d_C, nullptr, td, ss,
                                   W, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, d_D, d_P, d_B,
                                   true);
}
