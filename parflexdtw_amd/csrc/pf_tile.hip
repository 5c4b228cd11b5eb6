// Per-tile epilogue of FlexDTW on gfx950: best endpoint and backtrace on the
// full D/P/B of one tile (the 6-tuple API of
// gpu_flexdtw.flexdtw_chunk_from_global_C).
#include "pf_common.h"

namespace pf {

// find_best_endpoint (FlexDTW.py:1098-1148): candidate i < n_row is
// (R-1, buffer+i); later ones walk the last column upward from row R-2.
__global__ __launch_bounds__(256) void k_tile_endpoint(const double *__restrict__ D,
                                                       const int32_t *__restrict__ P,
                                                       int64_t R, int64_t W, int64_t buffer,
                                                       double *__restrict__ avg,
                                                       double *__restrict__ best) {
  __shared__ double sv[256];
  __shared__ int64_t si[256];
  const int64_t n_row = W > buffer ? W - buffer : 0;
  const int64_t n_col = R - 1 > buffer ? R - 1 - buffer : 0;
  const int64_t n = n_row + n_col;
  double bv = __builtin_huge_val();
  int64_t bi = -1;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
    int64_t r, c;
    if (i < n_row) {
      r = R - 1;
      c = buffer + i;
    } else {
      r = R - 2 - (i - n_row);
      c = W - 1;
    }
    const int64_t p = P[r * W + c];
    const int64_t rs = p >= 0 ? 0 : -p, cs = p >= 0 ? p : 0;
    const int64_t md = (r - rs) + (c - cs);
    const double a = D[r * W + c] / (double)md;
    avg[i] = a;
    if (a < bv) {  // i increases within a thread: first minimum kept
      bv = a;
      bi = i;
    }
  }
  sv[threadIdx.x] = bv;
  si[threadIdx.x] = bi;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      const double ov = sv[threadIdx.x + s];
      const int64_t oi = si[threadIdx.x + s];
      if (oi >= 0 && (si[threadIdx.x] < 0 || ov < sv[threadIdx.x] ||
                      (ov == sv[threadIdx.x] && oi < si[threadIdx.x]))) {
        sv[threadIdx.x] = ov;
        si[threadIdx.x] = oi;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const int64_t i = si[0];
    int64_t r = -1, c = -1;
    if (i >= 0) {
      if (i < n_row) {
        r = R - 1;
        c = buffer + i;
      } else {
        r = R - 2 - (i - n_row);
        c = W - 1;
      }
    }
    best[0] = i >= 0 ? sv[0] : __builtin_huge_val();
    best[1] = (double)r;
    best[2] = (double)c;
  }
}

// backtrace_flexdtw + reverse (FlexDTW.py:1155-1180, :1366-1368).
__global__ void k_tile_backtrace(const int8_t *__restrict__ B, int64_t R, int64_t W,
                                 StepSet ss, const double *__restrict__ best,
                                 int64_t *__restrict__ path, int64_t *__restrict__ len) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  int64_t r = (int64_t)best[1], c = (int64_t)best[2];
  if (r < 0 || c < 0) {
    *len = 0;
    return;
  }
  int64_t n = 0;
  path[0] = r;
  path[1] = c;
  n = 1;
  while (r != 0 && c != 0 && n < R + W) {
    if (r < 0 || c < 0 || r >= R || c >= W) break;  // B always points inside the tile
    const int k = B[r * W + c];
    if (k < 0 || k >= ss.K) break;
    r -= ss.dr[k];
    c -= ss.dc[k];
    path[2 * n] = r;
    path[2 * n + 1] = c;
    ++n;
  }
  for (int64_t a = 0, b = n - 1; a < b; ++a, --b) {
    const int64_t t0 = path[2 * a], t1 = path[2 * a + 1];
    path[2 * a] = path[2 * b];
    path[2 * a + 1] = path[2 * b + 1];
    path[2 * b] = t0;
    path[2 * b + 1] = t1;
  }
  *len = n;
}

}  // namespace pf

using namespace pf;

extern "C" int pf_tile_endpoint(const double *d_D, const int32_t *d_P, int64_t R, int64_t W,
                                int64_t buffer, double *d_avg, double *d_best, void *stream) {
  PF_REQUIRE(R > 0 && W > 0 && buffer >= 0, "pf_tile_endpoint: bad shape or buffer");
  PF_REQUIRE(d_D && d_P && d_avg && d_best, "pf_tile_endpoint: null buffer");
  hipLaunchKernelGGL(k_tile_endpoint, dim3(1), dim3(256), 0, as_stream(stream), d_D, d_P, R, W,
                     buffer, d_avg, d_best);
  PF_CHECK_HIP(hipGetLastError());
  return PF_OK;
}

extern "C" int pf_tile_backtrace(const int8_t *d_B, int64_t R, int64_t W, const int32_t *steps,
                                 int K, const double *d_best, int64_t *d_path, int64_t *d_len,
                                 void *stream) {
  StepSet ss;
  double w[4] = {1, 1, 1, 1};
  int rc = make_stepset(steps, w, K, &ss);
  if (rc) return rc;
  PF_REQUIRE(R > 0 && W > 0, "pf_tile_backtrace: bad shape");
  PF_REQUIRE(d_B && d_best && d_path && d_len, "pf_tile_backtrace: null buffer");
  hipLaunchKernelGGL(k_tile_backtrace, dim3(1), dim3(64), 0, as_stream(stream), d_B, R, W, ss,
                     d_best, d_path, d_len);
  PF_CHECK_HIP(hipGetLastError());
  return PF_OK;
}
