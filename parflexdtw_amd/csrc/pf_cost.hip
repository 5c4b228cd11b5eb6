// Cost matrix C = 1 - L2norm(F1)^T L2norm(F2) on gfx950
// (FlexDTW.py:1207-1210 L2norm, Parflex.py:1470 / gpu_flexdtw.py:95-109).
//
// FP64 form: column norms sqrt(sum_k F[k]^2) + 1e-9 (epsilon after the
// square root, as in the reference), features divided by their norm once, then
// every cell sums the K products in increasing k with separately rounded
// multiplies and adds (the library is built with -ffp-contract=off), so the
// result is bit-identical to the fixed-order host restatement.  The kernel is
// HBM-bound on the C write for chroma (K = 12): 8 bytes stored per 2K flops.
#include "pf_common.h"

namespace pf {

// Fn[k*n + i] = F[k*n + i] / (sqrt(sum_k F[k*n+i]^2) + 1e-9)
__global__ void k_cost_normalise(const double *__restrict__ F, int64_t K, int64_t n,
                                 double *__restrict__ Fn) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int64_t k = 0; k < K; ++k) {
    const double v = F[k * n + i];
    s = s + v * v;
  }
  const double nrm = sqrt(s) + 1e-9;
  for (int64_t k = 0; k < K; ++k) Fn[k * n + i] = F[k * n + i] / nrm;
}

// One cost job: a (N x M) block of C from frames [f1_off, f1_off + N) of the
// first feature array and [f2_off, f2_off + M) of the second; job tables are
// int64[8] rows (f1_off, f2_off, N, M, c_off, ldc, blk0, 0), blk0 = first
// workgroup of the job in a batched launch.
struct CostJob {
  int64_t f1_off, f2_off, N, M, c_off, ldc, blk0, pad;
};
static_assert(sizeof(CostJob) == 64, "cost job rows are int64[8]");

// Workgroup -> (job, local workgroup).  Batched: the job is the number of
// jobs with blk0 <= blockIdx.x, minus one, counted 64 jobs per step with a
// wave ballot (the loads of all steps are independent, so they overlap);
// tables longer than 1024 jobs use a binary search.
template <bool BATCH>
__device__ __forceinline__ CostJob find_job(const CostJob &one, const CostJob *__restrict__ jobs,
                                            int64_t njobs, int64_t &blk) {
  blk = blockIdx.x;
  if constexpr (!BATCH) {
    return one;
  } else {
    int64_t p;
    if (njobs <= 1024) {
      const int lane = threadIdx.x & 63;
      int cnt = 0;
      for (int64_t base = 0; base < njobs; base += 64) {
        const int64_t jj = base + lane;
        const bool le = jj < njobs && jobs[jj].blk0 <= blk;
        cnt += __popcll(__ballot(le));
      }
      p = cnt - 1;
    } else {
      int64_t lo = 0, hi = njobs - 1;
      while (lo < hi) {
        const int64_t mid = (lo + hi + 1) >> 1;
        if (jobs[mid].blk0 <= blk) lo = mid; else hi = mid - 1;
      }
      p = lo;
    }
    const CostJob jb = jobs[p];
    blk -= jb.blk0;
    return jb;
  }
}

__device__ __forceinline__ void st_nt(double *p, double v) { __builtin_nontemporal_store(v, p); }
__device__ __forceinline__ void st_nt(float *p, float v) { __builtin_nontemporal_store(v, p); }
__device__ __forceinline__ void st_nt(__hip_bfloat16 *p, __hip_bfloat16 v) {
  __builtin_nontemporal_store(__builtin_bit_cast(unsigned short, v),
                              reinterpret_cast<unsigned short *>(p));
}

// FP64: a workgroup of 256 threads covers 64 rows x 256 columns, one column
// per thread, 16 rows at a time; F1n values are uniform across the
// workgroup (scalar loads), F2n loads and C stores are coalesced along j.
constexpr int kF64Rows = 64, kF64Cols = 256, kF64Chunk = 16;

template <bool BATCH>
__global__ __launch_bounds__(256) void k_cost_f64(const double *__restrict__ F1n, int64_t n1,
                                                  const double *__restrict__ F2n, int64_t n2,
                                                  int64_t K, CostJob one,
                                                  const CostJob *__restrict__ jobs, int64_t njobs,
                                                  double *__restrict__ C) {
  int64_t blk;
  const CostJob jb = find_job<BATCH>(one, jobs, njobs, blk);
  const int64_t nbc = (jb.M + kF64Cols - 1) / kF64Cols;
  const int64_t i0 = (blk / nbc) * kF64Rows;
  const int64_t j = (blk % nbc) * kF64Cols + threadIdx.x;
  const int64_t jc = j < jb.M ? j : jb.M - 1;
  const double *a_base = F1n + jb.f1_off;
  const double *b_base = F2n + jb.f2_off + jc;
  double *c_base = C + jb.c_off + j;
  for (int ch = 0; ch < kF64Rows; ch += kF64Chunk) {
    if (i0 + ch >= jb.N) break;
    double acc[kF64Chunk];
#pragma unroll
    for (int r = 0; r < kF64Chunk; ++r) acc[r] = 0.0;
    for (int64_t k = 0; k < K; ++k) {
      const double b = b_base[k * n2];
#pragma unroll
      for (int r = 0; r < kF64Chunk; ++r) {
        const int64_t i = i0 + ch + r < jb.N ? i0 + ch + r : jb.N - 1;
        const double p = a_base[k * n1 + i] * b;
        acc[r] = acc[r] + p;
      }
    }
    if (j < jb.M) {
#pragma unroll
      for (int r = 0; r < kF64Chunk; ++r)
        if (i0 + ch + r < jb.N) st_nt(c_base + (i0 + ch + r) * jb.ldc, 1.0 - acc[r]);
    }
  }
}

// ---------------------------------------------------------------------------
// BF16 MFMA form (precision form): features normalised in FP64, rounded to
// BF16 and stored frame-major with K padded to a multiple of 16; the product
// runs on v_mfma_f32_32x32x16_bf16 with FP32 accumulation, as C = 1 + sum_k
// (-a_k) b_k (accumulators started at 1, row features negated), so the MFMA
// chain yields C and the epilogue only converts it.  A workgroup of
// 4 waves computes a 128 x 256 block of C, the waves in a 2 x 2 arrangement,
// each 64 x 128 as 2 x 4 MFMA blocks of 32 x 32: every A fragment feeds four
// MFMAs and every B fragment two, and the workgroup reads (128 + 256) rows
// of K features for 32768 cells (about 0.023 K bytes per cell), so large K
// stays bound by the C write rather than by operand traffic.  Operand layout
// (per lane l of a wave): A row l%32, k-offset (l/32)*8, 8 consecutive k;
// B column l%32, same k; D column l%32, rows (v/4)*8 + (l/32)*4 + v%4 for
// the 16 accumulators v.  Every cell sees the same instruction sequence over
// k as in the stage-1 sweep that computes C itself, so both give the same
// bits.  Each half of the block (64 rows) is staged through LDS and written
// row by row, 16 bytes per lane where the row is 16-byte aligned in C.
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float floatx16_t __attribute__((ext_vector_type(16)));
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));
typedef float floatx4_t __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
typedef float floatx2_t __attribute__((ext_vector_type(2)));
// two floats as BF16 (round to nearest even, as __float2bfloat16) in one
// packed conversion, a in the low half
__device__ __forceinline__ uint32_t bf16x2_bits(float a, float b) {
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(floatx2_t{a, b}, bf16x2_t));
}

// -x of 8 bf16 values (sign bits flipped; exact)
__device__ __forceinline__ bf16x8_t neg8(bf16x8_t v) {
  const u32x4_t u = __builtin_bit_cast(u32x4_t, v) ^ 0x80008000u;
  return __builtin_bit_cast(bf16x8_t, u);
}

// Frame i's feature k, bf16(F[k*n + i] / (sqrt(sum_k F[k*n+i]^2) + 1e-9)),
// zero padded to Kp, goes to Fb[i*Kp + k] (frame-major, the layout stage 1
// reads) or, when `blocked`, to Fb[((k/16)*n + i)*16 + k%16]: chunks of 16
// features, all frames of a chunk together, so that an MFMA fragment of 32
// frames x 8 features is one contiguous 512-byte run per half wave.  For
// Kp = 16 the two layouts are the same.  A thread writes each chunk of 16
// features of its frame as two 16-byte stores (Fb is 16-byte aligned).
__global__ void k_cost_prep_bf16(const double *__restrict__ F, int64_t K, int64_t n, int64_t Kp,
                                 __hip_bfloat16 *__restrict__ Fb, int blocked) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  // the squares summed in feature order; loads eight at a time so that they
  // are in flight together
  double s = 0.0;
  int64_t k0 = 0;
  for (; k0 + 8 <= K; k0 += 8) {
    double v[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) v[q] = F[(k0 + q) * n + i];
#pragma unroll
    for (int q = 0; q < 8; ++q) s = s + v[q] * v[q];
  }
  for (; k0 < K; ++k0) {
    const double v = F[k0 * n + i];
    s = s + v * v;
  }
  const double nrm = sqrt(s) + 1e-9;
  for (int64_t kc = 0; kc < Kp; kc += 16) {
    // the chunk's 16 loads first (features past K read feature K - 1, unused)
    double v[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) v[q] = F[(kc + q < K ? kc + q : K - 1) * n + i];
    uint32_t w[8];
#pragma unroll
    for (int q = 0; q < 16; q += 2) {
      const int64_t k = kc + q;
      const __hip_bfloat16 lo = __float2bfloat16(k < K ? (float)(v[q] / nrm) : 0.0f);
      const __hip_bfloat16 hi = __float2bfloat16(k + 1 < K ? (float)(v[q + 1] / nrm) : 0.0f);
      w[q >> 1] = (uint32_t)__builtin_bit_cast(uint16_t, lo) |
                  ((uint32_t)__builtin_bit_cast(uint16_t, hi) << 16);
    }
    const int64_t at = blocked ? ((kc >> 4) * n + i) * 16 : i * Kp + kc;
    u32x4_t *dst = reinterpret_cast<u32x4_t *>(Fb + at);
    dst[0] = u32x4_t{w[0], w[1], w[2], w[3]};
    dst[1] = u32x4_t{w[4], w[5], w[6], w[7]};
  }
}

// four adjacent elements of a staged row (8 bytes as BF16, 16 as FP32)
__device__ __forceinline__ void store4(float *p, float a, float b, float c, float d) {
  *reinterpret_cast<floatx4_t *>(p) = floatx4_t{a, b, c, d};
}
__device__ __forceinline__ void store4(__hip_bfloat16 *p, float a, float b, float c, float d) {
  *reinterpret_cast<uint2 *>(p) = uint2{bf16x2_bits(a, b), bf16x2_bits(c, d)};
}
__device__ __forceinline__ float cvt_out(float v, float *) { return v; }
__device__ __forceinline__ __hip_bfloat16 cvt_out(float v, __hip_bfloat16 *) {
  return __float2bfloat16(v);
}

// Workgroup block of C: 128 rows x 256 columns, 4 waves of 64 x 128 (2 x 4
// sub-blocks of 32 x 32); kWJ sub-block columns per wave.  (128 x 128
// blocks, 4 waves a SIMD instead of 2, were not faster: the row-major
// stores of C bound the kernel.)
constexpr int kWJ = 4;
constexpr int kMfmaRows = 128, kMfmaCols = 2 * 32 * kWJ;
constexpr int kTilePad = 8;  // keeps staged rows 16-byte aligned, no bank conflicts

template <typename OT, bool BATCH, int WPE = 2>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WPE))) void k_cost_mfma_bf16(const __hip_bfloat16 *__restrict__ A0,
                                                        const __hip_bfloat16 *__restrict__ B0,
                                                        int64_t n1, int64_t n2, int64_t Kp,
                                                        CostJob one,
                                                        const CostJob *__restrict__ jobs,
                                                        int64_t njobs, OT *__restrict__ C,
                                                        int vec_ok) {
  constexpr int kHalf = kMfmaRows / 2;
  __shared__ __attribute__((aligned(16))) OT tile[kHalf][kMfmaCols + kTilePad];
  int64_t blk;
  const CostJob jb = find_job<BATCH>(one, jobs, njobs, blk);
  const int64_t N = jb.N, M = jb.M;
  const int64_t nbr = (N + kMfmaRows - 1) / kMfmaRows;
  const int64_t nbc = (M + kMfmaCols - 1) / kMfmaCols;
  // XCD-aware order.  Workgroups go to the 8 XCDs round robin, so the
  // workgroups with equal blk % 8 share one XCD (and its L2).  Each such
  // class takes one contiguous eighth of the sequence `lin` (the blocks past
  // the last multiple of 8 keep their place).  `lin` runs over strips of
  // kStrip column blocks, each strip row by row: consecutive workgroups of
  // an XCD reuse one A row block kStrip times, and the strip's B columns
  // (kStrip x kMfmaCols frames) stay in that XCD's L2.  A bijection on
  // [0, nbr*nbc).
  constexpr int64_t kStrip = 16;
  const int64_t nblk = nbr * nbc, q = nblk >> 3;
  const int64_t lin = blk < (q << 3) ? (blk & 7) * q + (blk >> 3) : blk;
  const int64_t wide = nbc / kStrip * kStrip, in_wide = nbr * wide;
  int64_t rb, cbk;
  if (lin < in_wide) {
    const int64_t t = lin % (nbr * kStrip);
    rb = t / kStrip;
    cbk = lin / (nbr * kStrip) * kStrip + t % kStrip;
  } else {
    const int64_t t = lin - in_wide, w = nbc - wide;
    rb = t / w;
    cbk = wide + t % w;
  }
  const int64_t row0 = rb * kMfmaRows, col0 = cbk * kMfmaCols;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int wr = wave >> 1, wc = (wave & 1) * 32 * kWJ;
  const int lr = lane & 31, kh = (lane >> 5) * 8;
  // Operands in the blocked layout of k_cost_prep_bf16: frame f, features
  // [kk + kh, kk + kh + 8) at ((kk/16)*n + f)*16 + kh.  Rows/columns past
  // the matrix read the last valid frame (results unused).  Element offsets
  // fit 32 bits: (n1 + n2) * Kp < 2^31 (host).
  uint32_t ra[2], cb[kWJ];
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const int64_t r = row0 + wr * 64 + i * 32 + lr;
    ra[i] = (uint32_t)((jb.f1_off + (r < N ? r : N - 1)) * 16 + kh);
  }
#pragma unroll
  for (int j = 0; j < kWJ; ++j) {
    const int64_t c = col0 + wc + j * 32 + lr;
    cb[j] = (uint32_t)((jb.f2_off + (c < M ? c : M - 1)) * 16 + kh);
  }
  const uint32_t sa = (uint32_t)(n1 * 16), sb = (uint32_t)(n2 * 16);
  // C = 1 - a.b as 1 + sum_k (-a_k) b_k: accumulators start at 1 and the row
  // features are negated, so the MFMA chain yields C itself
  floatx16_t acc[2][kWJ];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < kWJ; ++j)
#pragma unroll
      for (int v = 0; v < 16; ++v) acc[i][j][v] = 1.0f;
  // Operands of three k-chunks in flight: chunk k + 3 is loaded into the
  // registers of chunk k once its MFMAs are issued (the accumulation order
  // over k is unchanged).
  const uint32_t nk = (uint32_t)(Kp >> 4);
  bf16x8_t a[3][2], b[3][kWJ];
  auto load_chunk = [&](const uint32_t kc, bf16x8_t(&ak)[2], bf16x8_t(&bk)[kWJ])
                        __attribute__((always_inline)) {
    const __hip_bfloat16 *An = A0 + kc * sa;
    const __hip_bfloat16 *Bn = B0 + kc * sb;
#pragma unroll
    for (int i = 0; i < 2; ++i) ak[i] = neg8(*reinterpret_cast<const bf16x8_t *>(An + ra[i]));
#pragma unroll
    for (int j = 0; j < kWJ; ++j) bk[j] = *reinterpret_cast<const bf16x8_t *>(Bn + cb[j]);
  };
#pragma unroll
  for (int d = 0; d < 3; ++d)
    if (d < (int)nk) load_chunk(d, a[d], b[d]);
  for (uint32_t k0 = 0; k0 < nk; k0 += 3) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      if (k0 + d < nk) {
        // column features as the A operand: a lane's accumulator holds 4
        // adjacent columns of one row (see the epilogue)
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < kWJ; ++j)
            acc[i][j] =
                __builtin_amdgcn_mfma_f32_32x32x16_bf16(b[d][j], a[d][i], acc[i][j], 0, 0, 0);
        if (k0 + d + 3 < nk) load_chunk(k0 + d + 3, a[d], b[d]);
      }
    }
  }
  constexpr int E = 16 / (int)sizeof(OT);
  const int64_t cols = M - col0 < kMfmaCols ? M - col0 : kMfmaCols;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    if (h) __syncthreads();  // the previous half has been written out
    if (wr == h) {
      // accumulator element v of sub-block (i, j): row i*32 + lr, column
      // wc + j*32 + 8*(v/4) + 4*(lane/32) + v%4, so each group of four is one
      // 4-element store
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < kWJ; ++j)
#pragma unroll
          for (int g = 0; g < 4; ++g)
            store4(&tile[i * 32 + lr][wc + j * 32 + g * 8 + (lane >> 5) * 4], acc[i][j][4 * g],
                   acc[i][j][4 * g + 1], acc[i][j][4 * g + 2], acc[i][j][4 * g + 3]);
    }
    __syncthreads();
    // RPI rows per wave and store instruction (LPR lanes a row): 16-byte
    // chunks of E elements when the row's first element is 16-byte aligned
    // in C, else element by element.
    constexpr int LPR = kMfmaCols / E, RPI = 64 / LPR;
    const int sub = lane / LPR, c = (lane % LPR) * E;
    const int64_t r0 = row0 + h * kHalf;
    const int rows = (int)(N - r0 < kHalf ? N - r0 : kHalf);
    for (int rr0 = wave * RPI; rr0 < rows; rr0 += 4 * RPI) {
      const int rr = rr0 + sub;
      if (rr < rows) {
        const int64_t e0 = jb.c_off + (r0 + rr) * jb.ldc + col0;
        OT *dst = C + e0;
        if (vec_ok && (e0 % E) == 0 && c + E <= cols) {
          const u32x4_t v = *reinterpret_cast<const u32x4_t *>(&tile[rr][c]);
          __builtin_nontemporal_store(v, reinterpret_cast<u32x4_t *>(dst + c));
        } else {
#pragma unroll
          for (int e = 0; e < E; ++e)
            if (c + e < cols) st_nt(dst + c + e, tile[rr][c + e]);
        }
      }
    }
  }
}

}  // namespace pf

using namespace pf;

extern "C" int pf_cost_workspace(int out_dtype, int64_t K, int64_t N, int64_t M, int64_t *bytes) {
  PF_REQUIRE(bytes && K >= 1 && N >= 1 && M >= 1, "pf_cost_workspace: bad arguments");
  if (out_dtype == PF_F64) {
    *bytes = K * (N + M) * (int64_t)sizeof(double);
  } else {
    PF_REQUIRE(out_dtype == PF_F32 || out_dtype == PF_BF16, "pf_cost_workspace: unknown dtype");
    const int64_t Kp = (K + 15) / 16 * 16;
    *bytes = (N + M) * Kp * 2 + 16;
  }
  return PF_OK;
}

extern "C" int pf_cost_tile(int out_dtype, int64_t *rows, int64_t *cols) {
  PF_REQUIRE(rows && cols, "pf_cost_tile: null argument");
  if (out_dtype == PF_F64) {
    *rows = kF64Rows;
    *cols = kF64Cols;
  } else {
    PF_REQUIRE(out_dtype == PF_F32 || out_dtype == PF_BF16, "pf_cost_tile: unknown dtype");
    *rows = kMfmaRows;
    *cols = kMfmaCols;
  }
  return PF_OK;
}

namespace {

// Normalise (FP64) or normalise-and-round (BF16) both feature arrays into
// d_work, then launch the product over nblocks workgroups.
template <bool BATCH>
int cost_launch(int out_dtype, const double *d_F1, int64_t n1, const double *d_F2, int64_t n2,
                int64_t K, const CostJob &one, const CostJob *d_jobs, int64_t njobs,
                int64_t nblocks, void *d_C, double *d_work, hipStream_t st) {
  PF_REQUIRE(nblocks >= 1 && nblocks < (1ll << 31), "pf_cost: grid size out of range");
  if (out_dtype == PF_F32 || out_dtype == PF_BF16) {
    const int64_t Kp = (K + 15) / 16 * 16;
    PF_REQUIRE((n1 + n2) * Kp < (1ll << 31), "pf_cost: (frames of both sides) x K too large");
    PF_REQUIRE(reinterpret_cast<uintptr_t>(d_work) % 16 == 0,
               "pf_cost: d_work must be 16-byte aligned");
    __hip_bfloat16 *A = reinterpret_cast<__hip_bfloat16 *>(d_work);
    __hip_bfloat16 *B = A + n1 * Kp;
    hipLaunchKernelGGL(k_cost_prep_bf16, dim3((unsigned)((n1 + 255) / 256)), dim3(256), 0, st,
                       d_F1, K, n1, Kp, A, 1);
    PF_CHECK_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_cost_prep_bf16, dim3((unsigned)((n2 + 255) / 256)), dim3(256), 0, st,
                       d_F2, K, n2, Kp, B, 1);
    PF_CHECK_HIP(hipGetLastError());
    const int vec_ok = reinterpret_cast<uintptr_t>(d_C) % 16 == 0;
    if (out_dtype == PF_F32)
      hipLaunchKernelGGL((k_cost_mfma_bf16<float, BATCH>), dim3((unsigned)nblocks), dim3(256), 0,
                         st, A, B, n1, n2, Kp, one, d_jobs, njobs, static_cast<float *>(d_C),
                         vec_ok);
    else
      hipLaunchKernelGGL((k_cost_mfma_bf16<__hip_bfloat16, BATCH>), dim3((unsigned)nblocks),
                         dim3(256), 0, st, A, B, n1, n2, Kp, one, d_jobs, njobs,
                         static_cast<__hip_bfloat16 *>(d_C), vec_ok);
    PF_CHECK_HIP(hipGetLastError());
    return PF_OK;
  }
  PF_REQUIRE(out_dtype == PF_F64, "pf_cost: unknown output dtype");
  double *F1n = d_work, *F2n = d_work + K * n1;
  hipLaunchKernelGGL(k_cost_normalise, dim3((unsigned)((n1 + 255) / 256)), dim3(256), 0, st, d_F1,
                     K, n1, F1n);
  PF_CHECK_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_cost_normalise, dim3((unsigned)((n2 + 255) / 256)), dim3(256), 0, st, d_F2,
                     K, n2, F2n);
  PF_CHECK_HIP(hipGetLastError());
  hipLaunchKernelGGL((k_cost_f64<BATCH>), dim3((unsigned)nblocks), dim3(256), 0, st, F1n, n1, F2n,
                     n2, K, one, d_jobs, njobs, static_cast<double *>(d_C));
  PF_CHECK_HIP(hipGetLastError());
  return PF_OK;
}

}  // namespace

extern "C" int pf_cost(int out_dtype, const double *d_F1, const double *d_F2, int64_t K,
                       int64_t N, int64_t M, void *d_C, int64_t ldc, double *d_work,
                       void *stream) {
  PF_REQUIRE(K >= 1 && N >= 1 && M >= 1 && ldc >= M, "pf_cost: bad shape");
  PF_REQUIRE(d_F1 && d_F2 && d_C && d_work, "pf_cost: null buffer");
  int64_t tr, tc;
  if (int rc = pf_cost_tile(out_dtype, &tr, &tc)) return rc;
  const CostJob one{0, 0, N, M, 0, ldc, 0, 0};
  const int64_t nblocks = ((N + tr - 1) / tr) * ((M + tc - 1) / tc);
  return cost_launch<false>(out_dtype, d_F1, N, d_F2, M, K, one, nullptr, 1, nblocks, d_C,
                            d_work, as_stream(stream));
}

extern "C" int pf_cost_batch(int out_dtype, const double *d_F1, int64_t n1_total,
                             const double *d_F2, int64_t n2_total, int64_t K,
                             const int64_t *d_jobs, int64_t njobs, int64_t total_blocks,
                             void *d_C, double *d_work, void *stream) {
  PF_REQUIRE(K >= 1 && n1_total >= 1 && n2_total >= 1 && njobs >= 1, "pf_cost_batch: bad shape");
  PF_REQUIRE(d_F1 && d_F2 && d_jobs && d_C && d_work, "pf_cost_batch: null buffer");
  int64_t tr, tc;
  if (int rc = pf_cost_tile(out_dtype, &tr, &tc)) return rc;
  const CostJob none{};
  return cost_launch<true>(out_dtype, d_F1, n1_total, d_F2, n2_total, K, none,
                           reinterpret_cast<const CostJob *>(d_jobs), njobs, total_blocks, d_C,
                           d_work, as_stream(stream));
}

extern "C" int pf_features_bf16(const double *d_F, int64_t K, int64_t n, void *d_out,
                                void *stream) {
  PF_REQUIRE(d_F && d_out && K >= 1 && n >= 1, "pf_features_bf16: bad arguments");
  PF_REQUIRE(reinterpret_cast<uintptr_t>(d_out) % 16 == 0,
             "pf_features_bf16: output must be 16-byte aligned");
  const int64_t Kp = (K + 15) / 16 * 16;
  hipLaunchKernelGGL(k_cost_prep_bf16, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     as_stream(stream), d_F, K, n, Kp, static_cast<__hip_bfloat16 *>(d_out), 0);
  PF_CHECK_HIP(hipGetLastError());
  return PF_OK;
}
