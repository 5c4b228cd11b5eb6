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

// Each thread: ROWS consecutive rows of one column j; a block covers ROWS rows
// x 256 columns, so F2n loads and C stores are coalesced along j and F1n
// values are uniform across the block.
template <int ROWS>
__global__ __launch_bounds__(256) void k_cost_f64(const double *__restrict__ F1n,
                                                  const double *__restrict__ F2n, int64_t K,
                                                  int64_t N, int64_t M, double *__restrict__ C,
                                                  int64_t ldc) {
  const int64_t j = blockIdx.y * (int64_t)blockDim.x + threadIdx.x;
  const int64_t i0 = blockIdx.x * (int64_t)ROWS;
  if (j >= M) return;
  double acc[ROWS];
#pragma unroll
  for (int r = 0; r < ROWS; ++r) acc[r] = 0.0;
  for (int64_t k = 0; k < K; ++k) {
    const double b = F2n[k * M + j];
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
      const int64_t i = i0 + r < N ? i0 + r : N - 1;
      const double p = F1n[k * N + i] * b;
      acc[r] = acc[r] + p;
    }
  }
#pragma unroll
  for (int r = 0; r < ROWS; ++r)
    if (i0 + r < N) C[(i0 + r) * ldc + j] = 1.0 - acc[r];
}

// ---------------------------------------------------------------------------
// BF16 MFMA form (precision form): features normalised in FP64, rounded to
// BF16 and stored frame-major with K padded to a multiple of 16; the product
// runs on v_mfma_f32_32x32x16_bf16 with FP32 accumulation.  A workgroup of
// 4 waves computes a 128 x 128 block of C, each wave 2 x 2 MFMA blocks of
// 32 x 32.  Operand layout (per lane l of a wave): A row l%32, k-offset
// (l/32)*8, 8 consecutive k; B column l%32, same k; D column l%32, rows
// (v/4)*8 + (l/32)*4 + v%4 for the 16 accumulators v.

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float floatx16_t __attribute__((ext_vector_type(16)));

// Fb[i*Kp + k] = bf16(F[k*n + i] / (sqrt(sum_k F[k*n+i]^2) + 1e-9)), zero pad.
__global__ void k_cost_prep_bf16(const double *__restrict__ F, int64_t K, int64_t n, int64_t Kp,
                                 __hip_bfloat16 *__restrict__ Fb) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int64_t k = 0; k < K; ++k) {
    const double v = F[k * n + i];
    s = s + v * v;
  }
  const double nrm = sqrt(s) + 1e-9;
  for (int64_t k = 0; k < Kp; ++k)
    Fb[i * Kp + k] = __float2bfloat16(k < K ? (float)(F[k * n + i] / nrm) : 0.0f);
}

__device__ __forceinline__ void store_c(float *p, float v) { *p = v; }
__device__ __forceinline__ void store_c(__hip_bfloat16 *p, float v) { *p = __float2bfloat16(v); }

template <typename OT>
__global__ __launch_bounds__(256) void k_cost_mfma_bf16(const __hip_bfloat16 *__restrict__ A,
                                                        const __hip_bfloat16 *__restrict__ B,
                                                        int64_t Kp, int64_t N, int64_t M,
                                                        OT *__restrict__ C, int64_t ldc) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int64_t row0 = blockIdx.x * 128ll + (wave >> 1) * 64;
  const int64_t col0 = blockIdx.y * 128ll + (wave & 1) * 64;
  const int lr = lane & 31, kh = (lane >> 5) * 8;
  // rows/columns past the matrix read the last valid frame (results unused)
  int64_t ra[2], cb[2];
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    ra[i] = row0 + i * 32 + lr < N ? row0 + i * 32 + lr : N - 1;
    cb[i] = col0 + i * 32 + lr < M ? col0 + i * 32 + lr : M - 1;
  }
  floatx16_t acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = floatx16_t{};
  for (int64_t kk = 0; kk < Kp; kk += 16) {
    bf16x8_t a[2], b[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      a[i] = *reinterpret_cast<const bf16x8_t *>(A + ra[i] * Kp + kk + kh);
      b[i] = *reinterpret_cast<const bf16x8_t *>(B + cb[i] * Kp + kk + kh);
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int64_t col = col0 + j * 32 + lr;
      if (col >= M) continue;
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int64_t row = row0 + i * 32 + (v >> 2) * 8 + (lane >> 5) * 4 + (v & 3);
        if (row < N) store_c(C + row * ldc + col, 1.0f - acc[i][j][v]);
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

extern "C" int pf_cost(int out_dtype, const double *d_F1, const double *d_F2, int64_t K,
                       int64_t N, int64_t M, void *d_C, int64_t ldc, double *d_work,
                       void *stream) {
  PF_REQUIRE(K >= 1 && N >= 1 && M >= 1 && ldc >= M, "pf_cost: bad shape");
  PF_REQUIRE(d_F1 && d_F2 && d_C && d_work, "pf_cost: null buffer");
  hipStream_t st = as_stream(stream);
  if (out_dtype == PF_F32 || out_dtype == PF_BF16) {
    const int64_t Kp = (K + 15) / 16 * 16;
    __hip_bfloat16 *A = reinterpret_cast<__hip_bfloat16 *>(d_work);
    __hip_bfloat16 *B = A + N * Kp;
    PF_REQUIRE(reinterpret_cast<uintptr_t>(d_work) % 16 == 0, "pf_cost: d_work must be 16-byte aligned");
    hipLaunchKernelGGL(k_cost_prep_bf16, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, st, d_F1,
                       K, N, Kp, A);
    PF_CHECK_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_cost_prep_bf16, dim3((unsigned)((M + 255) / 256)), dim3(256), 0, st, d_F2,
                       K, M, Kp, B);
    PF_CHECK_HIP(hipGetLastError());
    PF_REQUIRE((N + 127) / 128 < (1ll << 31) && (M + 127) / 128 < 65536,
               "pf_cost: matrix too large for one launch");
    const dim3 grid((unsigned)((N + 127) / 128), (unsigned)((M + 127) / 128));
    if (out_dtype == PF_F32)
      hipLaunchKernelGGL(k_cost_mfma_bf16<float>, grid, dim3(256), 0, st, A, B, Kp, N, M,
                         static_cast<float *>(d_C), ldc);
    else
      hipLaunchKernelGGL(k_cost_mfma_bf16<__hip_bfloat16>, grid, dim3(256), 0, st, A, B, Kp, N, M,
                         static_cast<__hip_bfloat16 *>(d_C), ldc);
    PF_CHECK_HIP(hipGetLastError());
    return PF_OK;
  }
  PF_REQUIRE(out_dtype == PF_F64, "pf_cost: unknown output dtype");
  double *F1n = d_work, *F2n = d_work + K * N;
  hipLaunchKernelGGL(k_cost_normalise, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, st, d_F1,
                     K, N, F1n);
  PF_CHECK_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_cost_normalise, dim3((unsigned)((M + 255) / 256)), dim3(256), 0, st, d_F2,
                     K, M, F2n);
  PF_CHECK_HIP(hipGetLastError());
  constexpr int ROWS = 4;
  PF_REQUIRE((N + ROWS - 1) / ROWS < (1ll << 31) && (M + 255) / 256 < 65536,
             "pf_cost: matrix too large for one launch");
  hipLaunchKernelGGL(k_cost_f64<ROWS>, dim3((unsigned)((N + ROWS - 1) / ROWS),
                                            (unsigned)((M + 255) / 256)),
                     dim3(256), 0, st, F1n, F2n, K, N, M, static_cast<double *>(d_C), ldc);
  PF_CHECK_HIP(hipGetLastError());
  return PF_OK;
}
