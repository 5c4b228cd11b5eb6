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

}  // namespace pf

using namespace pf;

extern "C" int pf_cost(int out_dtype, const double *d_F1, const double *d_F2, int64_t K,
                       int64_t N, int64_t M, void *d_C, int64_t ldc, double *d_work,
                       void *stream) {
  PF_REQUIRE(K >= 1 && N >= 1 && M >= 1 && ldc >= M, "pf_cost: bad shape");
  PF_REQUIRE(d_F1 && d_F2 && d_C && d_work, "pf_cost: null buffer");
  PF_REQUIRE(out_dtype == PF_F64, "pf_cost: this build provides the FP64 cost form");
  hipStream_t st = as_stream(stream);
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
