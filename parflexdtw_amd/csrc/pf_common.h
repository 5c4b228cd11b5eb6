// Shared definitions of the gfx950 ParFlex kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <stdint.h>

#include <cmath>
#include <string>

#include "parflex_hip.h"

namespace pf {

// Thread-local last-error text behind pf_last_error().
void set_error(const std::string &msg);

#define PF_CHECK_HIP(expr)                                                        \
  do {                                                                            \
    hipError_t _e = (expr);                                                       \
    if (_e != hipSuccess) {                                                       \
      ::pf::set_error(std::string(#expr) + ": " + hipGetErrorString(_e));         \
      return PF_EHIP;                                                             \
    }                                                                             \
  } while (0)

#define PF_REQUIRE(cond, msg)                                                     \
  do {                                                                            \
    if (!(cond)) {                                                                \
      ::pf::set_error(msg);                                                       \
      return PF_EINVAL;                                                           \
    }                                                                             \
  } while (0)

// Step set passed by value in kernel arguments (uniform, lives in SGPRs).
struct StepSet {
  int K;
  int dr[4];
  int dc[4];
  double w[4];
};

int make_stepset(const int32_t *steps, const double *weights, int K, StepSet *out);

// Tile table row (int64[8]), see parflex_hip.h.
struct TileDesc {
  int64_t c_off, ldc, R, W, b_off, e_off, pair, ij;
};
static_assert(sizeof(TileDesc) == 64, "tile table rows are int64[8]");

// A value every lane of the wave holds, moved to scalar registers.
__device__ __forceinline__ int64_t wave_uniform(int64_t v) {
  const int lo = __builtin_amdgcn_readfirstlane((int)(v & 0xffffffffll));
  const int hi = __builtin_amdgcn_readfirstlane((int)(v >> 32));
  return ((int64_t)hi << 32) | (uint32_t)lo;
}
__device__ __forceinline__ TileDesc wave_uniform(const TileDesc &d) {
  return TileDesc{wave_uniform(d.c_off), wave_uniform(d.ldc),  wave_uniform(d.R),
                  wave_uniform(d.W),     wave_uniform(d.b_off), wave_uniform(d.e_off),
                  wave_uniform(d.pair),  wave_uniform(d.ij)};
}

// Pair table row (int64[16]), see parflex_hip.h.
struct PairDesc {
  int64_t L1, L2, n1, n2, tile_base, c_off, ldc, buf, lk_off, seg_off, max_segs,
      path_off, r12, r13, r14, r15;
};
static_assert(sizeof(PairDesc) == 128, "pair table rows are int64[16]");

// Poll limit of the device waits (pf_set_wait_limit), and the wait error
// word of the stage-2 tile DAG (read and cleared; pf_device_status).
int wait_limit();
int take_s2_wait_error(hipStream_t st, int *word);

// Launch of the stage-2 back-pointer walk (pf_walk.hip): grid (101 segments,
// requests), one wave each, on stream st.
void launch_s2_walk(dim3 grid, hipStream_t st, const TileDesc *tiles, int64_t L, StepSet ss,
                    const uint32_t *Bp, int64_t *chain, const int64_t *cnt, int64_t *pts);

__device__ __forceinline__ double to_f64(double v) { return v; }
__device__ __forceinline__ double to_f64(float v) { return (double)v; }
__device__ __forceinline__ double to_f64(__hip_bfloat16 v) { return (double)__bfloat162float(v); }
__device__ __forceinline__ float to_f32(float v) { return v; }
__device__ __forceinline__ float to_f32(__hip_bfloat16 v) { return __bfloat162float(v); }

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

}  // namespace pf
