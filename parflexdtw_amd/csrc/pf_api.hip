// Error reporting and argument validation shared by the C-ABI entry points.
#include <cstring>

#include "pf_common.h"

namespace pf {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

int make_stepset(const int32_t *steps, const double *weights, int K, StepSet *out) {
  PF_REQUIRE(steps != nullptr && weights != nullptr, "steps/weights must not be null");
  PF_REQUIRE(K >= 1 && K <= 4, "the step set must have 1 to 4 steps");
  StepSet ss;
  std::memset(&ss, 0, sizeof(ss));
  ss.K = K;
  bool has_unit = false;
  for (int k = 0; k < K; ++k) {
    const int dr = steps[2 * k], dc = steps[2 * k + 1];
    PF_REQUIRE(dr >= 1 && dr <= 2 && dc >= 0 && dc <= 2,
               "steps must satisfy 1 <= dr <= 2 and 0 <= dc <= 2");
    PF_REQUIRE(std::isfinite(weights[k]), "weights must be finite");
    ss.dr[k] = dr;
    ss.dc[k] = dc;
    ss.w[k] = weights[k];
    if (dr == 1 && dc <= 1) has_unit = true;
  }
  PF_REQUIRE(has_unit, "the step set needs a step (1,0) or (1,1) so every cell has a predecessor");
  for (int k = K; k < 4; ++k) {
    ss.dr[k] = 1;
    ss.dc[k] = 1;
    ss.w[k] = 0.0;
  }
  *out = ss;
  return PF_OK;
}

}  // namespace pf

extern "C" const char *pf_last_error(void) { return pf::g_last_error.c_str(); }

extern "C" int pf_version(void) { return 101; }

extern "C" int pf_check_steps(const int32_t *steps, const double *weights, int K) {
  pf::StepSet ss;
  return pf::make_stepset(steps, weights, K, &ss);
}

// VALU issue-rate probe (measurement only): each lane runs `iters` rounds
// of 16 independent FP32 fused multiply-adds (v_fma_f32 / v_fmac_f32, one
// VALU wave instruction each), then stores one sum so that nothing is
// removed.  `blocks` workgroups of 256 threads.
namespace pf {
__global__ __launch_bounds__(256) void k_valu_probe(float *__restrict__ out, int iters, float a,
                                                    float b) {
  float x[16];
#pragma unroll
  for (int j = 0; j < 16; ++j) x[j] = threadIdx.x * 0.001f + j;
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int j = 0; j < 16; ++j) x[j] = __builtin_fmaf(x[j], a, b);
    asm volatile("" ::: "memory");
  }
  float s = 0.0f;
#pragma unroll
  for (int j = 0; j < 16; ++j) s += x[j];
  out[blockIdx.x * 256 + threadIdx.x] = s;
}
}  // namespace pf

extern "C" int pf_valu_probe(float *d_out, int blocks, int iters, void *stream) {
  PF_REQUIRE(d_out && blocks >= 1 && iters >= 1, "pf_valu_probe: bad arguments");
  hipLaunchKernelGGL(pf::k_valu_probe, dim3((unsigned)blocks), dim3(256), 0,
                     pf::as_stream(stream), d_out, iters, 1.0001f, 0.5f);
  PF_CHECK_HIP(hipGetLastError());
  return PF_OK;
}
