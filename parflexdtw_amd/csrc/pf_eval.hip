// Alignment evaluation on the device: for many warping paths at once, the
// predicted reference time of every ground-truth query time (np.interp along
// the path, as 05_Evaluate_Benchmarks_L.ipynb#cell8 and eval_tools.py:44-46
// compute it one file at a time), its error, and the histogram behind the
// error-rate-versus-tolerance curve (05_Evaluate_Benchmarks_L.ipynb#cell23).
//
// Work is tiny next to alignment (a few thousand beats per query, one
// bisection of ~17 steps each), so the kernel is latency bound: one thread
// per beat, one row of blocks per query, the thresholds and the block's
// histogram in LDS, one global atomic per non-empty bin per block.
#include "pf_common.h"

namespace pf {
namespace {

constexpr int kEvalThreads = 256;

// Path q read as np.interp's sample points: xp_j = rows[j*stride] * hop,
// fp_j = cols[j*stride] * hop (int64 frames to float64 seconds, one rounding
// each, as numpy's `wp[0, :] * hop_sec`).
struct PathView {
  const int64_t *rows, *cols;
  int64_t stride, n;
  double hop;
  __device__ double x(int64_t j) const { return (double)rows[j * stride] * hop; }
  __device__ double y(int64_t j) const { return (double)cols[j * stride] * hop; }
};

// np.interp(x, xp, fp) with its default left = fp[0], right = fp[-1], for
// non-decreasing xp (numpy/_core/src/multiarray/compiled_base.c, arr_interp):
//   one sample point: fp[0] for every x (NaN included);
//   NaN x: NaN;  x < xp[0]: fp[0];  x > xp[-1]: fp[-1];
//   otherwise j = last index with xp[j] <= x, and fp[j] when j is the last
//   point or xp[j] == x, else slope * (x - xp[j]) + fp[j] with
//   slope = (fp[j+1] - fp[j]) / (xp[j+1] - xp[j]), re-evaluated from the
//   right end point when that is NaN.
// Built with -ffp-contract=off, so every product and sum is rounded on its
// own as in numpy's loop.
__device__ double interp_path(double x, const PathView &v) {
  if (v.n == 1) return v.y(0);
  if (x != x) return x;
  if (x < v.x(0)) return v.y(0);
  const double xr = v.x(v.n - 1);
  if (x > xr) return v.y(v.n - 1);
  if (x == xr) return v.y(v.n - 1);
  // xp[lo] <= x < xp[hi] holds throughout.
  int64_t lo = 0, hi = v.n - 1;
  while (hi - lo > 1) {
    const int64_t mid = lo + ((hi - lo) >> 1);
    if (x >= v.x(mid)) lo = mid;
    else hi = mid;
  }
  const double xj = v.x(lo), yj = v.y(lo);
  if (xj == x) return yj;
  const double xj1 = v.x(lo + 1), yj1 = v.y(lo + 1);
  const double slope = (yj1 - yj) / (xj1 - xj);
  double r = slope * (x - xj) + yj;
  if (r != r) {
    r = slope * (x - xj1) + yj1;
    if (r != r && yj == yj1) r = yj;
  }
  return r;
}

// Beat b of query q (b in [gt_off[q], gt_off[q+1])):
//   err[b] = interp(gt_x[b]) - gt_y[b]   (+inf for a query without a path)
//   hist[k] += 1, k = #{i : |err[b]| > tol[i]}  (tol ascending; NaN -> 0)
// so that the error rate at tolerance i is sum_{k > i} hist[k] / #beats.
__global__ __launch_bounds__(kEvalThreads) void k_eval_paths(
    const int64_t *__restrict__ path, int64_t stride, const int64_t *__restrict__ offs,
    const int64_t *__restrict__ lens, const int64_t *__restrict__ gt_off,
    const double *__restrict__ gt_x, const double *__restrict__ gt_y, double hop,
    const double *__restrict__ tol, int ntol, double *__restrict__ err,
    unsigned long long *__restrict__ hist) {
  extern __shared__ double lds[];
  double *stol = lds;
  uint32_t *shist = reinterpret_cast<uint32_t *>(lds + ntol);

  const int64_t q = blockIdx.y;
  const int64_t b0 = gt_off[q], b1 = gt_off[q + 1];
  // Whole block leaves together, before any barrier.
  if (b0 + (int64_t)blockIdx.x * kEvalThreads >= b1) return;
  const int tid = threadIdx.x;
  if (hist) {
    for (int i = tid; i < ntol; i += kEvalThreads) stol[i] = tol[i];
    for (int i = tid; i <= ntol; i += kEvalThreads) shist[i] = 0;
    __syncthreads();
  }

  const int64_t b = b0 + (int64_t)blockIdx.x * kEvalThreads + tid;
  if (b < b1) {
    PathView v{path + offs[2 * q], path + offs[2 * q + 1], stride, lens[q], hop};
    const double pred = v.n > 0 ? interp_path(gt_x[b], v) : __builtin_inf();
    const double e = pred - gt_y[b];
    err[b] = e;
    if (hist) {
      const double a = fabs(e);
      int lo = 0, hi = ntol;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (a > stol[mid]) lo = mid + 1;
        else hi = mid;
      }
      atomicAdd(&shist[lo], 1u);
    }
  }
  if (hist) {
    __syncthreads();
    for (int i = tid; i <= ntol; i += kEvalThreads) {
      const uint32_t c = shist[i];
      if (c) atomicAdd(&hist[i], (unsigned long long)c);
    }
  }
}

}  // namespace
}  // namespace pf

extern "C" int pf_eval_paths(const int64_t *d_path, int64_t stride, const int64_t *d_offs,
                             const int64_t *d_lens, int64_t nq, const int64_t *d_gt_off,
                             int64_t max_beats, const double *d_gt_x, const double *d_gt_y,
                             double hop, const double *d_tol, int64_t ntol, double *d_err,
                             uint64_t *d_hist, void *stream) {
  PF_REQUIRE(d_path && d_offs && d_lens && d_gt_off && d_gt_x && d_gt_y && d_err,
             "pf_eval_paths: null pointer");
  PF_REQUIRE(stride >= 1 && nq >= 0 && max_beats >= 0, "pf_eval_paths: bad sizes");
  PF_REQUIRE(nq <= 65535, "pf_eval_paths: more than 65535 queries in one launch");
  PF_REQUIRE(ntol >= 0 && ntol <= PF_EVAL_MAX_TOL, "pf_eval_paths: ntol out of range");
  PF_REQUIRE(!d_hist || ntol >= 1, "pf_eval_paths: a histogram needs thresholds");
  PF_REQUIRE(!d_hist || d_tol, "pf_eval_paths: histogram without thresholds");
  PF_REQUIRE(hop > 0.0 && std::isfinite(hop), "pf_eval_paths: hop must be finite and > 0");
  if (nq == 0 || max_beats == 0) return PF_OK;
  const int nt = d_hist ? (int)ntol : 0;
  const size_t lds = d_hist ? (size_t)nt * sizeof(double) + (size_t)(nt + 1) * sizeof(uint32_t)
                            : 0;
  const int64_t bx = (max_beats + pf::kEvalThreads - 1) / pf::kEvalThreads;
  PF_REQUIRE(bx <= 0x7fffffff, "pf_eval_paths: too many beats in one query");
  hipLaunchKernelGGL(pf::k_eval_paths, dim3((unsigned)bx, (unsigned)nq), dim3(pf::kEvalThreads),
                     lds, pf::as_stream(stream), d_path, stride, d_offs, d_lens, d_gt_off, d_gt_x,
                     d_gt_y, hop, d_tol, nt, d_err,
                     reinterpret_cast<unsigned long long *>(d_hist));
  PF_CHECK_HIP(hipGetLastError());
  return PF_OK;
}
