// Learned tile proposal on gfx950: pooled cost, a 4-layer 3x3 convolution
// network (32 channels) trained and run with BF16 MFMA operands and FP32
// accumulation / master weights, BCE loss, Adam, and per-tile scores.
//
// Activations are NHWC with 32 channels, bf16: X[(y*W + x)*32 + c].  Every
// layer is a 32 -> 32 channel 3x3 convolution (stride 1, zero padding 1); the
// network input carries the pooled cost in channel 0 and zeros elsewhere, and
// only channel 0 of the last layer is used (its logit).  A layer is an
// implicit GEMM  Y (pixels x 32) = im2col(X) (pixels x 288) * W^T, with the
// reduction index k = (dy*3 + dx)*32 + cin, so that 8 consecutive k are 8
// consecutive channels of one neighbour pixel (one 16-byte load).  Weights are
// stored [cout][288].  The gradient with respect to the layer input is the
// same convolution of dY with the flipped weights Wf[cin][(8-dydx)*32 + cout]
// = W[cout][dydx*32 + cin]; the weight gradient is the GEMM
// dW (32 x 288) = dY^T (32 x pixels) * im2col(X) (pixels x 288).
//
// MFMA v_mfma_f32_32x32x16_bf16 lane layout (lane l of a wave): A row l%32,
// k = (l/32)*8 .. +7; B column l%32, same k; D column l%32, rows
// (v/4)*8 + (l/32)*4 + v%4 for accumulator v.
#include "pf_common.h"

#include <type_traits>

namespace pf {

typedef __bf16 lbf16x8_t __attribute__((ext_vector_type(8)));
typedef float lfloatx16_t __attribute__((ext_vector_type(16)));
typedef unsigned int lu32x4_t __attribute__((ext_vector_type(4)));

constexpr int kCh = 32;           // channels of every layer
constexpr int kKdim = 9 * kCh;    // im2col width (288)

__device__ __forceinline__ float bf2f(__hip_bfloat16 v) { return __bfloat162float(v); }

// A run of frames pooled into consecutive blocks: frames [f_off, f_off + n)
// of the feature array into pooled blocks [b_off, b_off + ceil(n / 32)).
struct PoolSeg {
  int64_t f_off, n, b_off, pad;
};

// Fp[k*ldp + g] = mean over the frames i of pooled block g (32 frames, the
// last block of a segment ragged) of F[k*ldf + i] / (sqrt(sum_k F[k*ldf+i]^2)
// + 1e-9).  A workgroup pools 8 consecutive blocks: thread t takes frame t % 32
// of block t / 32 (coalesced loads), normalises it and writes its scaled
// features to LDS, 16 at a time; then thread k*8 + b sums the 32 values of
// feature k of block b in the tree of a 32-lane xor butterfly (offsets 16, 8,
// 4, 2, 1), so the means equal a shuffle reduction's bit for bit.  The
// segment of block g is the last one with b_off <= g (binary search over
// `segs`, or `one` when segs is null), so one launch pools every pair of a
// batch.  With FB, the thread of frame i also writes the frame's BF16
// features Fb[i*Kp + k] = bf16(F[k*ldf + i] / (sqrt(sum_k F^2) + 1e-9)),
// zero padded to Kp (a multiple of 16), from the same loads: the values and
// layout of k_cost_prep_bf16's frame-major form, bit for bit.  With REG (K
// <= 16) a frame's features are read once, into registers, for both the
// norm and the values; otherwise they are read twice.
constexpr int kPoolBlocks = 8;                 // pooled blocks per workgroup
constexpr int kPoolStride = 36;                // floats per block row in LDS (32 + pad)
template <bool FB, bool REG>
__global__ __launch_bounds__(256) void k_pool_features(const double *__restrict__ F, int64_t K,
                                                       int64_t ldf,
                                                       const PoolSeg *__restrict__ segs,
                                                       int64_t nseg, PoolSeg one,
                                                       float *__restrict__ Fp, int64_t ldp,
                                                       int64_t nb,
                                                       __hip_bfloat16 *__restrict__ Fb,
                                                       int64_t Kp) {
  __shared__ __attribute__((aligned(16))) float v[16][kPoolBlocks * kPoolStride];
  const int t = threadIdx.x;
  const int64_t g0 = blockIdx.x * (int64_t)kPoolBlocks;
  const int64_t g = g0 + (t >> 5);
  PoolSeg sg = one;
  if (segs) {
    int64_t lo = 0, hi = nseg - 1;
    while (lo < hi) {
      const int64_t mid = (lo + hi + 1) >> 1;
      if (segs[mid].b_off <= g) lo = mid; else hi = mid - 1;
    }
    sg = segs[lo];
  }
  const int64_t b = g - sg.b_off;
  const int64_t fi = b * 32 + (t & 31);
  const bool in = g < nb && fi < sg.n;
  const int64_t i = sg.f_off + fi;
  const int64_t cnt = (b * 32 + 32 < sg.n) ? 32 : sg.n - b * 32;
  double s = 0.0;
  double xr[REG ? 16 : 1];
  if constexpr (REG) {
#pragma unroll
    for (int k = 0; k < 16; ++k) xr[k] = (in && k < K) ? F[k * ldf + i] : 0.0;
#pragma unroll
    for (int k = 0; k < 16; ++k)
      if (k < K) s = s + xr[k] * xr[k];
  } else {
    for (int64_t k = 0; k < K; ++k) {
      const double x = in ? F[k * ldf + i] : 0.0;
      s = s + x * x;
    }
  }
  const double nrm = sqrt(s) + 1e-9;
  const double inv = 1.0 / (nrm * (double)(cnt > 0 ? cnt : 1));
  const int slot = (t >> 5) * kPoolStride + (t & 31);
  const int rk = t >> 3, rb = t & 7;   // reducing thread: feature rk of block rb
  for (int64_t k0 = 0; k0 < K; k0 += 16) {
    const int kn = (int)((K - k0) < 16 ? (K - k0) : 16);
    if (REG && !FB) {
#pragma unroll
      for (int q = 0; q < 16; ++q)
        if (q < kn) v[q][slot] = in ? (float)(xr[q] * inv) : 0.0f;
    } else if (FB) {
      uint32_t w[8] = {0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u};
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        if (q < kn) {
          double x;
          if constexpr (REG) x = xr[q];
          else x = in ? F[(k0 + q) * ldf + i] : 0.0;
          v[q][slot] = in ? (float)(x * inv) : 0.0f;
          const uint32_t h = __builtin_bit_cast(uint16_t, __float2bfloat16((float)(x / nrm)));
          w[q >> 1] |= h << (16 * (q & 1));
        }
      }
      if (in) {
        lu32x4_t *dst = reinterpret_cast<lu32x4_t *>(Fb + i * Kp + k0);
        dst[0] = lu32x4_t{w[0], w[1], w[2], w[3]};
        dst[1] = lu32x4_t{w[4], w[5], w[6], w[7]};
      }
    } else {
      for (int q = 0; q < kn; ++q)
        v[q][slot] = in ? (float)(F[(k0 + q) * ldf + i] * inv) : 0.0f;
    }
    __syncthreads();
    if (rk < kn) {
      float a[32];
      const float4 *src = reinterpret_cast<const float4 *>(&v[rk][rb * kPoolStride]);
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const float4 w = src[q];
        a[4 * q] = w.x;
        a[4 * q + 1] = w.y;
        a[4 * q + 2] = w.z;
        a[4 * q + 3] = w.w;
      }
#pragma unroll
      for (int o = 16; o > 0; o >>= 1)
#pragma unroll
        for (int j = 0; j < o; ++j) a[j] = a[j] + a[j + o];
      if (g0 + rb < nb) Fp[(k0 + rk) * ldp + g0 + rb] = a[0];
    }
    __syncthreads();
  }
}

// Pooled cost P[y][x] = 1 - sum_k A[k*H + y] * B[k*W + x] (the mean of C over
// the block, C being bilinear in the normalised features), f32.
__global__ void k_pooled_cost(const float *__restrict__ A, int64_t H, const float *__restrict__ B,
                              int64_t W, int64_t K, float *__restrict__ P) {
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= H * W) return;
  const int64_t y = p / W, x = p - y * W;
  float s = 0.0f;
  for (int64_t k = 0; k < K; ++k) s += A[k * H + y] * B[k * W + x];
  P[p] = 1.0f - s;
}

// Network input: the pooled cost double-centred (row and column means
// removed, the global mean added back) and divided by its standard
// deviation, so that a path block stands out by its local contrast.  P being
// 1 - A^T B, the double-centred value is -(A_y - abar).(B_x - bbar) with
// abar, bbar the means of the pooled features over the frames, and its mean
// square is sum_{k,l} GA[k][l] GB[k][l] / (H W) with GA, GB the Gram
// matrices of the centred features: O(K^2 (H + W)) work, no H x W pass.
// One workgroup: st[0..K) = abar, st[K..2K) = bbar, st[2K] = 1 / std.
__global__ __launch_bounds__(256) void k_input_stats_any(const float *__restrict__ A, int64_t H,
                                                     const float *__restrict__ B, int64_t W,
                                                     int64_t K, float *__restrict__ st) {
  __shared__ float red[256];
  const int t = threadIdx.x;
  for (int64_t k = 0; k < 2 * K; ++k) {
    const float *src = k < K ? A + k * H : B + (k - K) * W;
    const int64_t n = k < K ? H : W;
    float s = 0.0f;
    for (int64_t y = t; y < n; y += 256) s += src[y];
    red[t] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
      if (t < h) red[t] += red[t + h];
      __syncthreads();
    }
    if (t == 0) st[k] = red[0] / (float)n;
    __syncthreads();
  }
  // entries (k, l), k <= l, of both Gram matrices, strided over the threads
  float q = 0.0f;
  const int64_t ne = K * (K + 1) / 2;
  for (int64_t e = t; e < ne; e += 256) {
    int64_t k = 0, rem = e;
    while (rem >= K - k) {
      rem -= K - k;
      ++k;
    }
    const int64_t l = k + rem;
    float ga = 0.0f, gb = 0.0f;
    for (int64_t y = 0; y < H; ++y) ga += (A[k * H + y] - st[k]) * (A[l * H + y] - st[l]);
    for (int64_t x = 0; x < W; ++x) gb += (B[k * W + x] - st[K + k]) * (B[l * W + x] - st[K + l]);
    q += (k == l ? 1.0f : 2.0f) * ga * gb;
  }
  red[t] = q;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (t < h) red[t] += red[t + h];
    __syncthreads();
  }
  if (t == 0) st[2 * K] = 1.0f / sqrtf(red[0] / ((float)H * (float)W) + 1e-12f);
}

// Same statistics for K <= KM, one workgroup per pair of a batch: pair p's
// pooled features are A[k*lda + a_off + y] (y < H) and B[k*ldb + b_off + x]
// (x < W), its statistics st + p*(2K+1) (`one` when `pairs` is null).  Wave w
// sums the means of k = w, w+4, ...; the centred features are then staged in
// LDS in chunks of kStChunk frames and wave w accumulates the Gram entries
// e = w, w+4, ... (lanes over frames).  Every sum has a fixed order, so the
// result does not depend on the launch or on other pairs.
struct StatPair {
  int64_t a_off, H, b_off, W;
};
constexpr int kStChunk = 512;

template <int KM>
__global__ __launch_bounds__(256) void k_input_stats_seg(const float *__restrict__ A, int64_t lda,
                                                         const float *__restrict__ B, int64_t ldb,
                                                         int K, const StatPair *__restrict__ pairs,
                                                         StatPair one, float *__restrict__ st_all) {
  constexpr int E = KM * (KM + 1) / 2;
  constexpr int EW = (E + 3) / 4;
  __shared__ float cs[KM * kStChunk];
  __shared__ float mean[KM];
  __shared__ float G[2][E];
  __shared__ unsigned char ek[E], el[E];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const StatPair pr = pairs ? pairs[blockIdx.x] : one;
  float *st = st_all + (int64_t)blockIdx.x * (2 * K + 1);
  const int Ek = K * (K + 1) / 2;
  if (t == 0) {
    int e = 0;
    for (int k = 0; k < K; ++k)
      for (int l = k; l < K; ++l, ++e) {
        ek[e] = (unsigned char)k;
        el[e] = (unsigned char)l;
      }
  }
  for (int side = 0; side < 2; ++side) {
    const float *S = side ? B + pr.b_off : A + pr.a_off;
    const int64_t ld = side ? ldb : lda;
    const int64_t n = side ? pr.W : pr.H;
    for (int k = wave; k < K; k += 4) {
      float v = 0.0f;
      for (int64_t f = lane; f < n; f += 64) v += S[k * ld + f];
      for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
      if (lane == 0) {
        mean[k] = v / (float)n;
        st[side * K + k] = v / (float)n;
      }
    }
    __syncthreads();
    float g[EW];
#pragma unroll
    for (int j = 0; j < EW; ++j) g[j] = 0.0f;
    for (int64_t c0 = 0; c0 < n; c0 += kStChunk) {
      const int m = (int)(n - c0 < kStChunk ? n - c0 : kStChunk);
      for (int q = t; q < K * kStChunk; q += 256) {
        const int k = q / kStChunk, f = q - k * kStChunk;
        cs[q] = f < m ? S[k * ld + c0 + f] - mean[k] : 0.0f;
      }
      __syncthreads();
#pragma unroll
      for (int j = 0; j < EW; ++j) {
        const int e = wave + 4 * j;
        if (e < Ek) {
          const float *ck = cs + ek[e] * kStChunk, *cl = cs + el[e] * kStChunk;
          float v = g[j];
          for (int f = lane; f < m; f += 64) v += ck[f] * cl[f];
          g[j] = v;
        }
      }
      __syncthreads();
    }
#pragma unroll
    for (int j = 0; j < EW; ++j) {
      const int e = wave + 4 * j;
      float v = g[j];
      for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
      if (lane == 0 && e < Ek) G[side][e] = v;
    }
    __syncthreads();
  }
  if (t == 0) {
    float q = 0.0f;
    for (int e = 0; e < Ek; ++e) q += (ek[e] == el[e] ? 1.0f : 2.0f) * G[0][e] * G[1][e];
    st[2 * K] = 1.0f / sqrtf(q / ((float)pr.H * (float)pr.W) + 1e-12f);
  }
}

__global__ void k_write_input(const float *__restrict__ A, int64_t H, const float *__restrict__ B,
                              int64_t W, int64_t K, const float *__restrict__ st,
                              __hip_bfloat16 *__restrict__ X, __hip_bfloat16 *__restrict__ Xc) {
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= H * W) return;
  const int64_t y = p / W, x = p - y * W;
  float s = 0.0f;
  for (int64_t k = 0; k < K; ++k) s += (A[k * H + y] - st[k]) * (B[k * W + x] - st[K + k]);
  const __hip_bfloat16 v = __float2bfloat16(-s * st[2 * K]);
  if (Xc) Xc[p] = v;
  if (X) {
    __hip_bfloat16 *o = X + p * kCh;
    o[0] = v;
    for (int q = 1; q < kCh; ++q) o[q] = __float2bfloat16(0.0f);
  }
}

// 3x3 convolution, 32 -> 32 channels.  A workgroup of 4 waves computes a
// block of 16 rows x 32 columns of output pixels; wave w takes rows
// 4w .. 4w+3 (four MFMA tiles of 32 channels x 32 pixels).  The 18 x 34
// input halo of the block is staged in LDS once; each wave keeps its 18
// weight fragments in registers and reads every input row of its window
// once per (dx, channel half), feeding the up to three output rows that use
// it (dy = input row - output row), so an LDS read serves 2 MFMAs on
// average instead of 1 and no weight is read from LDS.
// The MFMA computes the transposed tile D[cout][pixel] = W * im2col(X)^T, so
// lane l holds pixel l%32 and output channels (v/4)*8 + (l/32)*4 + v%4: four
// consecutive channels per accumulator group, stored straight to HBM as one
// 8-byte write (the two half-waves together write 16 contiguous bytes of
// every pixel), without staging the tile through LDS.
// MODE 0: Y = relu(acc + bias) as bf16 (hidden layer).
// MODE 1: the logit map (output channel 0 + bias, f32, H*W) from per-row
//         partial sums (logit_frags).
// MODE 2: Y = (M > 0) ? acc : 0 as bf16 (gradient through the ReLU of the
//         layer input M, which is that layer's output activation).
constexpr int kRowsPerWave = 4;
constexpr int kCvRows = 4 * kRowsPerWave, kCvCols = 32;
constexpr int kPixPad = kCh + 8;      // bf16 per staged pixel (16-byte aligned rows)
typedef unsigned int lu32x2_t __attribute__((ext_vector_type(2)));
typedef float lfloatx4_t __attribute__((ext_vector_type(4)));

typedef __bf16 lbf16x2_t __attribute__((ext_vector_type(2)));
typedef float lfloatx2_t __attribute__((ext_vector_type(2)));
// Two floats rounded to bf16 (nearest even, as __float2bfloat16) in one
// packed conversion; lo in the low half.
__device__ __forceinline__ unsigned int pack_bf16x2(float lo, float hi) {
  return __builtin_bit_cast(unsigned int, __builtin_convertvector(lfloatx2_t{lo, hi}, lbf16x2_t));
}

// Epilogue of the transposed tile acc (channels x 32 pixels) of output row y,
// pixels x0 .. x0+31 (lane l holds pixel x0 + l%32).  MODE 1 writes the logit
// straight to HBM (32 consecutive floats per wave); the other modes write the
// row's bf16 outputs into the LDS staging row ys (pixel stride kPixPad, four
// channels = 8 bytes per lane and accumulator group), which flush_block then
// copies to HBM in 16-byte chunks.
// Bias of the output channels lane `lane` holds: b4[g][i] = bias[g*8 + (lane/32)*4 + i].
template <int MODE>
__device__ __forceinline__ void load_bias4(const float *bias, int lane, lfloatx4_t (&b4)[4]) {
#pragma unroll
  for (int g = 0; g < 4; ++g)
    b4[g] = (MODE == 2) ? lfloatx4_t{0.0f, 0.0f, 0.0f, 0.0f}
                        : *reinterpret_cast<const lfloatx4_t *>(&bias[g * 8 + (lane >> 5) * 4]);
}

template <int MODE>
__device__ __forceinline__ void conv_epilogue(const lfloatx16_t &acc, const lfloatx4_t (&b4)[4],
                                              int64_t y, int64_t x0, int64_t H, int64_t W, int lane,
                                              const __hip_bfloat16 *__restrict__ Mk, void *Yv,
                                              __hip_bfloat16 *ys) {
  const int64_t x = x0 + (lane & 31);
  const bool in = y < H && x < W;
  const uint32_t pix = in ? (uint32_t)(y * W + x) : 0u;
  if constexpr (MODE == 1) {
    if (in && lane < 32) static_cast<float *>(Yv)[pix] = acc[0] + b4[0][0];   // the logit map
  } else {
    const int cq = (lane >> 5) * 4;
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      float r[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        r[i] = acc[g * 4 + i] + b4[g][i];
        if constexpr (MODE == 0) r[i] = r[i] > 0.0f ? r[i] : 0.0f;
      }
      if constexpr (MODE == 2) {
        // the four mask values (bf16) as floats: their bits in the high half
        const lu32x2_t mv = in ? *reinterpret_cast<const lu32x2_t *>(&Mk[pix * (uint32_t)kCh + g * 8 + cq])
                               : lu32x2_t{0u, 0u};
        const float m[4] = {__uint_as_float(mv[0] << 16), __uint_as_float(mv[0] & 0xffff0000u),
                            __uint_as_float(mv[1] << 16), __uint_as_float(mv[1] & 0xffff0000u)};
#pragma unroll
        for (int i = 0; i < 4; ++i) r[i] = m[i] > 0.0f ? r[i] : 0.0f;
      }
      *reinterpret_cast<lu32x2_t *>(&ys[(lane & 31) * kPixPad + g * 8 + cq]) =
          lu32x2_t{pack_bf16x2(r[0], r[1]), pack_bf16x2(r[2], r[3])};
    }
  }
}

// Copy a staged block of `rows` (even) x 32 pixels (bf16 NHWC 32, pixel
// stride kPixPad in LDS) to Y in 16-byte chunks: a block row is 2 KB
// contiguous.  Thread tid copies chunk tid % 4 of pixel column (tid / 4) % 32
// on rows tid / 128 + 2 i.  32-bit element offsets (H * W * 32 < 2^31).
__device__ __forceinline__ void flush_block(const __hip_bfloat16 *ys, __hip_bfloat16 *__restrict__ Y,
                                            int64_t y0, int64_t x0, int64_t H, int64_t W, int rows,
                                            int tid) {
  const int c8 = tid & 3, xl = (tid >> 2) & 31, rr0 = tid >> 7;
  const int x = (int)x0 + xl;
  if (x >= (int)W) return;
  for (int i = 0; i < rows / 2; ++i) {
    const int rr = rr0 + 2 * i;
    const int y = (int)y0 + rr;
    if (y < (int)H)
      *reinterpret_cast<lu32x4_t *>(&Y[(uint32_t)(y * (int)W + x) * (uint32_t)kCh + c8 * 8]) =
          *reinterpret_cast<const lu32x4_t *>(&ys[(rr * kCvCols + xl) * kPixPad + c8 * 8]);
  }
}

constexpr int kHaloChunks = (kCvRows + 2) * (kCvCols + 2) * 4;   // 16-byte chunks of a halo
constexpr int kHaloPerThread = (kHaloChunks + 255) / 256;

// The logit layer (one output channel) as per-row partial sums: for input
// row y' the MFMA with A[m][k] = W[0][(m*3 + dx)*32 + k] (rows m = dy < 3,
// zero rows above) gives, in accumulator v = dy of lanes 0..31,
//   P_y'[dy][x] = sum over (channel half, dx) of W[0][dy, dx, c] X[y'][x + dx - 1][c]
// (k-steps in the order channel half, dx), and the logit of row y is
//   ((bias + P_{y-1}[0][x]) + P_y[1][x]) + P_{y+1}[2][x].
// Each input row is read once per (dx, channel half): 6 MFMAs a row instead
// of 18 per output row.  Fragment q = ch*3 + dx of lane l: k = ch*16 + kh.
__device__ __forceinline__ void logit_frags(const __hip_bfloat16 *__restrict__ Wt, int lane,
                                            lbf16x8_t (&wl)[6]) {
  const int m = lane & 31, kh = (lane >> 5) * 8;
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    const int ch = q / 3, dx = q - ch * 3;
#pragma unroll
    for (int i = 0; i < 8; ++i)
      wl[q][i] = m < 3 ? (__bf16)bf2f(Wt[(m * 3 + dx) * kCh + ch * 16 + kh + i]) : (__bf16)0.0f;
  }
}

// A workgroup computes `seg` consecutive blocks of one 32-column strip,
// top to bottom.  The next block's input halo is loaded into registers
// (clamped addresses, no branch around the loads) while the current block
// is computed and written, and is moved into LDS after it: the HBM latency
// of a halo overlaps the MFMA work and the stores of the previous block.
template <int MODE>
__global__ __launch_bounds__(256) void k_conv3x3(const __hip_bfloat16 *__restrict__ X, int64_t H,
                                                 int64_t W, const __hip_bfloat16 *__restrict__ Wt,
                                                 const float *__restrict__ bias,
                                                 const __hip_bfloat16 *__restrict__ Mk,
                                                 void *__restrict__ Yv, int seg) {
  constexpr int HR = kCvRows + 2, HC = kCvCols + 2;
  __shared__ __attribute__((aligned(16))) __hip_bfloat16 xs[HR * HC * kPixPad];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int64_t x0 = (int64_t)blockIdx.x * kCvCols;
  const int64_t ya = (int64_t)blockIdx.y * seg * kCvRows;
  const int64_t left = (H - ya + kCvRows - 1) / kCvRows;
  const int nb = (int)(left < seg ? left : seg);
  const int px = lane & 31;
  const int kh = (lane >> 5) * 8;
  // weight fragments of this lane: row px (output channel), k = s*16 + kh .. +7
  lbf16x8_t wb[18];
#pragma unroll
  for (int s = 0; s < 18; ++s)
    wb[s] = *reinterpret_cast<const lbf16x8_t *>(&Wt[px * kKdim + s * 16 + kh]);
  // bias in LDS (read per block: no registers held across the sweep)
  __shared__ __attribute__((aligned(16))) float sbias[kCh];
  if (tid < kCh) sbias[tid] = (MODE == 2) ? 0.0f : bias[tid];
  // input halo of the block at row y0: HR x HC pixels x 4 chunks of 8
  // channels; thread tid owns chunk c8 = tid % 4 of halo pixels
  // (tid / 4) + 64 i = r_i * HC + c_i, stepped without division.  Element
  // offsets fit in 32 bits (the host checks H * W * 32 < 2^31).
  lu32x4_t P[kHaloPerThread];
  const int c8 = tid & 3, pix0 = tid >> 2;
  const int rh0 = pix0 / HC, ch0 = pix0 - rh0 * HC;
  const int Hi = (int)H, Wi = (int)W, x0i = (int)x0;
  // (r, c) restart from an opaque copy in every block, so the compiler steps
  // them again instead of keeping kHaloPerThread addresses alive across the
  // sweep (the halo registers need the room)
  auto halo_load = [&](const int y0) {
    int r = rh0, c = ch0;
    asm volatile("" : "+v"(r), "+v"(c));
#pragma unroll
    for (int i = 0; i < kHaloPerThread; ++i) {
      int yy = y0 - 1 + r, xx = x0i - 1 + c;   // clamped: every load is in bounds
      yy = yy < 0 ? 0 : (yy >= Hi ? Hi - 1 : yy);
      xx = xx < 0 ? 0 : (xx >= Wi ? Wi - 1 : xx);
      P[i] = *reinterpret_cast<const lu32x4_t *>(
          &X[(uint32_t)(yy * Wi + xx) * (uint32_t)kCh + (uint32_t)(c8 * 8)]);
      c += 64 - HC;   // next pixel: 64 = HC + 30
      r += 1;
      if (c >= HC) {
        c -= HC;
        r += 1;
      }
    }
  };
  auto halo_store = [&](const int y0) {
    int r = rh0, c = ch0;
    asm volatile("" : "+v"(r), "+v"(c));
#pragma unroll
    for (int i = 0; i < kHaloPerThread; ++i) {
      const int pix = pix0 + 64 * i;
      if (pix < HR * HC) {
        const int yy = y0 - 1 + r, xx = x0i - 1 + c;
        const bool ok = (uint32_t)yy < (uint32_t)Hi && (uint32_t)xx < (uint32_t)Wi;   // zero padding
        *reinterpret_cast<lu32x4_t *>(&xs[pix * kPixPad + c8 * 8]) =
            ok ? P[i] : lu32x4_t{0u, 0u, 0u, 0u};
      }
      c += 64 - HC;
      r += 1;
      if (c >= HC) {
        c -= HC;
        r += 1;
      }
    }
  };
  halo_load((int)ya);
  halo_store((int)ya);
  __syncthreads();
  const int r0 = kRowsPerWave * wave;   // first output row of the wave, halo row r0 + dy
  lbf16x8_t wl[6];
  if constexpr (MODE == 1) logit_frags(Wt, lane, wl);
  for (int b = 0; b < nb; ++b) {
    const int64_t y0 = ya + (int64_t)b * kCvRows;
    const bool more = b + 1 < nb;
    if (more) halo_load((int)y0 + kCvRows);
    if constexpr (MODE == 1) {
      // partial sums of the wave's 6 halo rows, then its 4 logit rows
      float p0[kRowsPerWave + 2], p1[kRowsPerWave + 2], p2[kRowsPerWave + 2];
#pragma unroll
      for (int xr = 0; xr < kRowsPerWave + 2; ++xr) {
        lfloatx16_t d = lfloatx16_t{};
#pragma unroll
        for (int q = 0; q < 6; ++q) {
          const int ch = q / 3, dx = q - ch * 3;
          const lbf16x8_t a = *reinterpret_cast<const lbf16x8_t *>(
              &xs[((r0 + xr) * HC + px + dx) * kPixPad + ch * 16 + kh]);
          d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wl[q], a, d, 0, 0, 0);
        }
        p0[xr] = d[0];
        p1[xr] = d[1];
        p2[xr] = d[2];
      }
#pragma unroll
      for (int t = 0; t < kRowsPerWave; ++t) {
        const int64_t y = y0 + r0 + t, x = x0 + px;
        if (lane < 32 && y < H && x < W)
          static_cast<float *>(Yv)[(uint32_t)(y * W + x)] =
              ((sbias[0] + p0[t]) + p1[t + 1]) + p2[t + 2];
      }
    } else {
    lfloatx16_t acc[kRowsPerWave];
#pragma unroll
    for (int t = 0; t < kRowsPerWave; ++t) acc[t] = lfloatx16_t{};
#pragma unroll
    for (int ch = 0; ch < 2; ++ch)
#pragma unroll
      for (int xr = 0; xr < kRowsPerWave + 2; ++xr)
#pragma unroll
        for (int dx = 0; dx < 3; ++dx) {
          const lbf16x8_t a = *reinterpret_cast<const lbf16x8_t *>(
              &xs[((r0 + xr) * HC + px + dx) * kPixPad + ch * 16 + kh]);
#pragma unroll
          for (int t = 0; t < kRowsPerWave; ++t) {
            const int dy = xr - t;
            if (dy >= 0 && dy <= 2)
              acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wb[(dy * 3 + dx) * 2 + ch], a,
                                                               acc[t], 0, 0, 0);
          }
        }
    if constexpr (MODE != 1) __syncthreads();   // every wave is done reading the halo
    lfloatx4_t b4[4];
    load_bias4<MODE>(sbias, lane, b4);
#pragma unroll
    for (int t = 0; t < kRowsPerWave; ++t)
      conv_epilogue<MODE>(acc[t], b4, y0 + r0 + t, x0, H, W, lane, Mk, Yv,
                          xs + (r0 + t) * kCvCols * kPixPad);
    if constexpr (MODE != 1) {
      __syncthreads();
      flush_block(xs, static_cast<__hip_bfloat16 *>(Yv), y0, x0, H, W, kCvRows, tid);
    }
    }
    __syncthreads();   // the halo (or the staged block) has been read
    if (more) {
      halo_store((int)y0 + kCvRows);
      __syncthreads();
    }
  }
}

// First layer from a compact one-channel input Xc (bf16, H*W): the 9 taps of
// channel 0, padded to 16, are one MFMA step (k = dy*3 + dx; the other input
// channels of the layer are zero).  Output as MODE 0: relu(acc + bias), bf16
// NHWC 32.  A workgroup of 4 waves covers 4 rows x 32 columns.
constexpr int kC1Rows = 4, kC1Cols = 32;
__global__ __launch_bounds__(256) void k_conv3x3_first(const __hip_bfloat16 *__restrict__ Xc,
                                                       int64_t H, int64_t W,
                                                       const __hip_bfloat16 *__restrict__ Wt,
                                                       const float *__restrict__ bias,
                                                       __hip_bfloat16 *__restrict__ Y) {
  __shared__ __hip_bfloat16 xs[(kC1Rows + 2) * (kC1Cols + 2)];
  __shared__ __attribute__((aligned(16))) __hip_bfloat16 ys[kC1Rows * kCvCols * kPixPad];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int64_t y0 = (int64_t)blockIdx.y * kC1Rows, x0 = (int64_t)blockIdx.x * kC1Cols;
  for (int q = tid; q < (kC1Rows + 2) * (kC1Cols + 2); q += 256) {
    const int r = q / (kC1Cols + 2), c = q - r * (kC1Cols + 2);
    const int64_t yy = y0 - 1 + r, xx = x0 - 1 + c;
    const bool ok = yy >= 0 && yy < H && xx >= 0 && xx < W;
    const __hip_bfloat16 v = Xc[ok ? yy * W + xx : 0];
    xs[q] = ok ? v : __float2bfloat16(0.0f);
  }
  __syncthreads();
  const int px = lane & 31, kh = (lane >> 5) * 8;
  lbf16x8_t a, b;
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    const int k = kh + t;                       // tap dy*3 + dx, or padding
    const int dy = k / 3, dx = k - (k / 3) * 3;
    const __hip_bfloat16 xv = xs[(k < 9 ? (wave + dy) * (kC1Cols + 2) + px + dx : 0)];
    const __hip_bfloat16 wv = Wt[px * kKdim + (k < 9 ? k : 0) * kCh];
    a[t] = k < 9 ? (__bf16)bf2f(xv) : (__bf16)0.0f;
    b[t] = k < 9 ? (__bf16)bf2f(wv) : (__bf16)0.0f;
  }
  // transposed tile: channels x pixels (see k_conv3x3)
  const lfloatx16_t acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(b, a, lfloatx16_t{}, 0, 0, 0);
  lfloatx4_t b4[4];
  load_bias4<0>(bias, lane, b4);
  conv_epilogue<0>(acc, b4, y0 + wave, x0, H, W, lane, nullptr, Y, ys + wave * kCvCols * kPixPad);
  __syncthreads();
  flush_block(ys, Y, y0, x0, H, W, kC1Rows, tid);
}

// First layer straight from a pair's pooled features and input statistics
// (no input map): the halo value at (y, x) is the network input
// bf16(-sum_k (A[k][y] - abar_k) (B[k][x] - bbar_k) * inv_std), computed in the
// same order as k_write_input, so the output equals k_write_input followed by
// k_conv3x3_first bit for bit.  A workgroup sweeps `seg` blocks of 16 rows x
// 32 columns down one column strip (wave w rows 4w .. 4w+3 of a block, one
// MFMA each): the strip's column features are staged once, the next block's
// row features are loaded into registers while the current block is
// computed and written; K <= 16.
__global__ __launch_bounds__(256) void k_conv3x3_first_pooled(
    const float *__restrict__ A, int64_t lda, int64_t a_off, int64_t H,
    const float *__restrict__ B, int64_t ldb, int64_t b_off, int64_t W, int K,
    const float *__restrict__ st, const __hip_bfloat16 *__restrict__ Wt,
    const float *__restrict__ bias, __hip_bfloat16 *__restrict__ Y, int seg) {
  constexpr int HR = kCvRows + 2, HC = kCvCols + 2;
  constexpr int KS = 17;   // row stride of ca / cb (odd: no bank conflicts)
  __shared__ float ca[HR * KS], cb[HC * KS];
  __shared__ __hip_bfloat16 xs[HR * HC];
  __shared__ __attribute__((aligned(16))) __hip_bfloat16 ys[kCvRows * kCvCols * kPixPad];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int64_t x0 = (int64_t)blockIdx.x * kCvCols;
  const int64_t ya = (int64_t)blockIdx.y * seg * kCvRows;
  const int64_t left = (H - ya + kCvRows - 1) / kCvRows;
  const int nb = (int)(left < seg ? left : seg);
  // centred features of the strip's halo columns; entries k >= K of ca and
  // cb are zero, so the dot products below run over 16 terms unrolled (the
  // extra +0.0 terms leave every sum bit-identical)
  for (int q = tid; q < HC * 16; q += 256) {
    const int c = q >> 4, k = q & 15;
    const int64_t xx = x0 - 1 + c;
    cb[c * KS + k] = (k < K && xx >= 0 && xx < W) ? B[k * ldb + b_off + xx] - st[K + k] : 0.0f;
  }
  for (int q = tid; q < HR * 16; q += 256)
    if ((q & 15) >= K) ca[(q >> 4) * KS + (q & 15)] = 0.0f;
  // row features of a block's halo rows: entry q < HR*K, (r, k) = (q / K,
  // q % K), is held in a register by thread q % 256 (slot q / 256); clamped
  // loads, masked when stored
  constexpr int kRowSlots = (HR * 16 + 255) / 256;
  int ra[kRowSlots], ka[kRowSlots];
#pragma unroll
  for (int j = 0; j < kRowSlots; ++j) {
    const int q = tid + 256 * j < HR * K ? tid + 256 * j : 0;
    ra[j] = q / K;
    ka[j] = q - ra[j] * K;
  }
  auto row_load = [&](const int64_t y0, const int j) {
    int64_t yy = y0 - 1 + ra[j];
    yy = yy < 0 ? 0 : (yy >= H ? H - 1 : yy);
    return A[ka[j] * lda + a_off + yy];
  };
  float pa[kRowSlots];
#pragma unroll
  for (int j = 0; j < kRowSlots; ++j) pa[j] = row_load(ya, j);
  const int px = lane & 31, kh = (lane >> 5) * 8;
  lbf16x8_t b;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const int k = kh + i;                       // tap dy*3 + dx, or padding
    b[i] = k < 9 ? (__bf16)bf2f(Wt[px * kKdim + k * kCh]) : (__bf16)0.0f;
  }
  lfloatx4_t b4[4];
  load_bias4<0>(bias, lane, b4);
  const float inv = st[2 * K];
  for (int blk = 0; blk < nb; ++blk) {
    const int64_t y0 = ya + (int64_t)blk * kCvRows;
#pragma unroll
    for (int j = 0; j < kRowSlots; ++j) {
      if (tid + 256 * j < HR * K) {
        const int64_t yy = y0 - 1 + ra[j];
        ca[ra[j] * KS + ka[j]] = (yy >= 0 && yy < H) ? pa[j] - st[ka[j]] : 0.0f;
      }
    }
    __syncthreads();
    if (blk + 1 < nb) {   // in flight during this block
#pragma unroll
      for (int j = 0; j < kRowSlots; ++j) pa[j] = row_load(y0 + kCvRows, j);
    }
    for (int q = tid; q < HR * HC; q += 256) {
      const int r = q / HC, c = q - r * HC;
      const int64_t yy = y0 - 1 + r, xx = x0 - 1 + c;
      const bool ok = yy >= 0 && yy < H && xx >= 0 && xx < W;
      float s = 0.0f;
#pragma unroll
      for (int k = 0; k < 16; ++k) s += ca[r * KS + k] * cb[c * KS + k];
      xs[q] = ok ? __float2bfloat16(-s * inv) : __float2bfloat16(0.0f);
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < kRowsPerWave; ++t) {
      const int orow = kRowsPerWave * wave + t;
      lbf16x8_t a;
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const int k = kh + i;
        const int dy = k / 3, dx = k - (k / 3) * 3;
        const __bf16 xv = reinterpret_cast<const __bf16 *>(xs)[k < 9 ? (orow + dy) * HC + px + dx : 0];
        a[i] = k < 9 ? xv : (__bf16)0.0f;
      }
      // transposed tile: channels x pixels (see k_conv3x3)
      const lfloatx16_t acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(b, a, lfloatx16_t{}, 0, 0, 0);
      conv_epilogue<0>(acc, b4, y0 + orow, x0, H, W, lane, nullptr, Y,
                       ys + orow * kCvCols * kPixPad);
    }
    __syncthreads();
    flush_block(ys, Y, y0, x0, H, W, kCvRows, tid);
    __syncthreads();   // ys, xs and ca are rewritten by the next block
  }
}

// Weight gradient partials: one wave per chunk of kWgChunk pixels;
// part[blk][o*288 + k] = sum over the chunk of dY[p][o] * im2col(X)[p][k],
// part[blk][32*288 + o] = sum over the chunk of dY[p][o].
constexpr int kWgChunk = 256;
constexpr int kWgSize = kCh * kKdim + kCh;

__global__ __launch_bounds__(64) void k_conv3x3_wgrad(const __hip_bfloat16 *__restrict__ dY,
                                                      const __hip_bfloat16 *__restrict__ X,
                                                      int64_t H, int64_t W,
                                                      float *__restrict__ part) {
  const int lane = threadIdx.x;
  const int64_t npix = H * W;
  const int64_t c0 = blockIdx.x * (int64_t)kWgChunk;
  const int kh = (lane >> 5) * 8;
  const int col = lane & 31;
  lfloatx16_t acc[9];
#pragma unroll
  for (int nb = 0; nb < 9; ++nb) acc[nb] = lfloatx16_t{};
  float bsum = 0.0f;
  for (int step = 0; step < kWgChunk; step += 16) {
    // A: dY^T, row o = lane%32, pixels c0+step+kh .. +7
    lbf16x8_t a;
    int64_t py[8], px[8];
    bool pin[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const int64_t q = c0 + step + kh + t;
      pin[t] = q < npix;
      const int64_t qq = pin[t] ? q : 0;
      py[t] = qq / W;
      px[t] = qq - py[t] * W;
      const __hip_bfloat16 g = dY[qq * kCh + col];
      const float gf = pin[t] ? bf2f(g) : 0.0f;
      bsum += gf;
      a[t] = (__bf16)gf;
    }
#pragma unroll
    for (int nb = 0; nb < 9; ++nb) {
      const int dy = nb / 3, dx = nb - dy * 3;
      lbf16x8_t b;
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        const int64_t yy = py[t] + dy - 1, xx = px[t] + dx - 1;
        const bool ok = pin[t] && yy >= 0 && yy < H && xx >= 0 && xx < W;
        const __hip_bfloat16 v = X[(ok ? yy * W + xx : 0) * kCh + col];
        b[t] = ok ? (__bf16)bf2f(v) : (__bf16)0.0f;
      }
      acc[nb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc[nb], 0, 0, 0);
    }
  }
  float *dst = part + blockIdx.x * (int64_t)kWgSize;
#pragma unroll
  for (int nb = 0; nb < 9; ++nb)
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      const int o = (v >> 2) * 8 + (lane >> 5) * 4 + (v & 3);
      dst[o * kKdim + nb * kCh + col] = acc[nb][v];
    }
  // lanes l and l+32 hold the same output channel
  bsum += __shfl_xor(bsum, 32);
  if (lane < 32) dst[kCh * kKdim + lane] = bsum;
}

// g[i] += sum over nblk partials of part[b*n + i]
__global__ void k_reduce_add(const float *__restrict__ part, int64_t nblk, int64_t n,
                             float *__restrict__ g) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  float s = 0.0f;
  for (int64_t b = 0; b < nblk; ++b) s += part[b * n + i];
  g[i] += s;
}

// Weighted BCE on the logit map Z (f32, H*W): dZ channel 0 =
// scale * w * (sigmoid(z) - y), other channels 0; loss_acc += scale * sum w*bce.
__global__ __launch_bounds__(256) void k_bce(const float *__restrict__ Z,
                                             const uint8_t *__restrict__ label, int64_t npix,
                                             float pos_weight, float scale,
                                             __hip_bfloat16 *__restrict__ dZ,
                                             float *__restrict__ loss_acc) {
  __shared__ float red[256];
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  float l = 0.0f;
  if (p < npix) {
    const float z = Z[p];
    const float y = label[p] ? 1.0f : 0.0f;
    const float w = label[p] ? pos_weight : 1.0f;
    l = w * (fmaxf(z, 0.0f) - z * y + log1pf(expf(-fabsf(z)))) * scale;
    const float s = 1.0f / (1.0f + expf(-z));
    __hip_bfloat16 *o = dZ + p * kCh;
    o[0] = __float2bfloat16(w * (s - y) * scale);
    for (int q = 1; q < kCh; ++q) o[q] = __float2bfloat16(0.0f);
  }
  red[threadIdx.x] = l;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(loss_acc, red[0]);
}

// Adam on FP32 master weights; the gradient is cleared after use.
__global__ void k_adam(float *__restrict__ w, float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, const uint8_t *__restrict__ frozen, int64_t n,
                       float lr, float b1, float b2, float eps, float c1, float c2) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  g[i] = 0.0f;
  if (frozen && frozen[i]) return;
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// BF16 copies of one layer's master weights: Wt[o][k] and the flipped
// Wf[cin][(8 - dydx)*32 + cout] = W[cout][dydx*32 + cin].
__global__ void k_pack_weights(const float *__restrict__ w, __hip_bfloat16 *__restrict__ Wt,
                               __hip_bfloat16 *__restrict__ Wf) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kCh * kKdim) return;
  const int o = i / kKdim, k = i - o * kKdim;
  const int dydx = k / kCh, cin = k - dydx * kCh;
  const __hip_bfloat16 b = __float2bfloat16(w[i]);
  Wt[i] = b;
  if (Wf) Wf[cin * kKdim + (8 - dydx) * kCh + o] = b;
}

// Per-tile score: the largest logit (map Z, f32 H*W) over the pooled cells a tile
// covers, the region extended by `ext` cells on every side.  Tile (i, j) of
// tile length L covers rows [i*hop, min(i*hop+L, L1)) and columns [j*hop,
// min(j*hop+L, L2)), hop = L-1; pooled cell (y, x) covers frames
// [y*pool, y*pool+pool) x [x*pool, x*pool+pool).
// A wave walks rows wave, wave+4, ... of the region with its lanes along
// the row (coalesced), four rows in flight per step; max is exact in any
// order.  Element offsets fit 32 bits (the host checks H * W < 2^31).
__device__ __forceinline__ float tile_score_of(const float *__restrict__ Z, int64_t H, int64_t W,
                                               int pool, int64_t L, int64_t L1, int64_t L2,
                                               int64_t i, int64_t j, int ext, float *red) {
  const int64_t hop = L - 1;
  const int64_t r0 = i * hop, r1 = (r0 + L < L1 ? r0 + L : L1) - 1;
  const int64_t c0 = j * hop, c1 = (c0 + L < L2 ? c0 + L : L2) - 1;
  const int y0 = (int)(r0 / pool - ext > 0 ? r0 / pool - ext : 0);
  const int y1 = (int)(r1 / pool + ext < H - 1 ? r1 / pool + ext : H - 1);
  const int x0 = (int)(c0 / pool - ext > 0 ? c0 / pool - ext : 0);
  const int x1 = (int)(c1 / pool + ext < W - 1 ? c1 / pool + ext : W - 1);
  const int Wi = (int)W;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const float ninf = -__builtin_huge_valf();
  float m0 = ninf, m1 = ninf, m2 = ninf, m3 = ninf;
  for (int x = x0 + lane; x <= x1; x += 64) {
    int y = y0 + wave;
    for (; y + 12 <= y1; y += 16) {
      const uint32_t o = (uint32_t)(y * Wi + x), s4 = (uint32_t)(4 * Wi);
      const float z0 = Z[o], z1 = Z[o + s4], z2 = Z[o + 2 * s4], z3 = Z[o + 3 * s4];
      m0 = fmaxf(m0, z0);
      m1 = fmaxf(m1, z1);
      m2 = fmaxf(m2, z2);
      m3 = fmaxf(m3, z3);
    }
    for (; y <= y1; y += 4) m0 = fmaxf(m0, Z[(uint32_t)(y * Wi + x)]);
  }
  float m = fmaxf(fmaxf(m0, m1), fmaxf(m2, m3));
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
  if (lane == 0) red[wave] = m;
  __syncthreads();
  return fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
}

__global__ __launch_bounds__(256) void k_tile_scores(const float *__restrict__ Z, int64_t H,
                                                     int64_t W, int pool, int64_t L, int64_t L1,
                                                     int64_t L2, int64_t n2, int ext,
                                                     float *__restrict__ score) {
  __shared__ float red[4];
  const int64_t t = blockIdx.x;
  const int64_t i = t / n2, j = t - i * n2;
  const float m = tile_score_of(Z, H, W, pool, L, L1, L2, i, j, ext, red);
  if (threadIdx.x == 0) score[t] = m;
}

// Tile scores of every pair of a batch in one launch: tile t (of the
// batch's i-major tile order) belongs to pair tpair[t], whose row of
// `pairs` (int64[8]) is {H, W, z_off, L1, L2, n2, tile_base, 0}: its logit
// map is Z + z_off (H x W), its tiles tile_base .. tile_base + n1*n2 - 1.
struct ScorePair {
  int64_t H, W, z_off, L1, L2, n2, tile_base, r7;
};
__global__ __launch_bounds__(256) void k_tile_scores_batch(const float *__restrict__ Z,
                                                           const ScorePair *__restrict__ pairs,
                                                           const int32_t *__restrict__ tpair,
                                                           int64_t ntiles, int pool, int64_t L,
                                                           int ext, float *__restrict__ score) {
  __shared__ float red[4];
  const int64_t t = blockIdx.x + (int64_t)blockIdx.y * 65536;
  if (t >= ntiles) return;   // the whole workgroup: the last grid row is partial
  const ScorePair pd = pairs[tpair[t]];
  const int64_t q = t - pd.tile_base;
  const int64_t i = q / pd.n2, j = q - i * pd.n2;
  const float m = tile_score_of(Z + pd.z_off, pd.H, pd.W, pool, L, pd.L1, pd.L2, i, j, ext, red);
  if (threadIdx.x == 0) score[t] = m;
}

// The four layers of the tile network in one kernel (inference).  A job is
// a strip of kNtOut output columns x up to kNtSeg output rows of one pair's
// logit map.  The kernel streams down the rows: every step each layer
// computes one row of kNtRow pixels (two MFMA tiles of 32), the first layer
// from the network input (built from the pooled features as
// k_conv3x3_first_pooled does), the others from the last three rows of the
// layer below, which LDS holds as rings of four rows.  Layer l's row runs
// 2l rows behind the first layer's, so every read is of a row written in one
// of the three steps before, and each step writes the fourth slot: one
// barrier per step.  Each layer's columns start one pixel right of those
// of the layer below (a 3x3 tap reads pixels c .. c+2 of the row below), so
// of kNtRow pixels 2 per layer are lost and kNtRow - 6 logits remain.
// Waves 0, 1 run layers 2, 3, each with its layer's weight fragments in
// registers; wave 2 runs layer 4 from per-row partial sums (logit_frags: the
// newest layer-3 row once, 6 MFMAs a tile instead of 18); wave 3 builds the
// input row and runs layer 1 (one MFMA per tile), off the path of the others.
// The roles rotate with the workgroup index.
// Pixels and rows outside the map are stored as zeros (the zero padding of
// the per-layer kernels), and every sum runs in the order of those kernels,
// so the logits equal those of k_conv3x3_first_pooled, k_conv3x3<0> twice
// and k_conv3x3<1> bit for bit.
constexpr int kNtT = 2;                  // MFMA tiles per layer row
constexpr int kNtRow = 32 * kNtT;        // pixels a layer computes per row
constexpr int kNtOut = kNtRow - 6;       // logits per job row
constexpr int kNtRing = kNtRow + 2;      // input slots of a row
constexpr int kNtSeg = 256;              // output rows per job
constexpr int kNtDepth = 4;              // rows per ring: 3 read, 1 written per step
constexpr int kNtMaxK = 16;

struct NetPair {
  int64_t a_off, H, b_off, W, z_off, st_off, r6, r7;
};
struct NetJob {
  int64_t pair, x0, y0, y1;
};

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_tile_net(
    const float *__restrict__ A, int64_t lda, const float *__restrict__ B, int64_t ldb, int K,
    const float *__restrict__ st_all, const NetPair *__restrict__ pairs,
    const NetJob *__restrict__ jobs, const __hip_bfloat16 *__restrict__ W1,
    const float *__restrict__ b1, const __hip_bfloat16 *__restrict__ W2,
    const float *__restrict__ b2, const __hip_bfloat16 *__restrict__ W3,
    const float *__restrict__ b3, const __hip_bfloat16 *__restrict__ W4,
    const float *__restrict__ b4, float *__restrict__ Z_all) {
  // rings of layers 1-3: [layer][slot][pixel][32 channels], kNtRow pixels
  // a row. The four 16-byte chunks (8 channels) of pixel p are stored in
  // the order c ^ ((p >> 2) & 3): 16 consecutive pixels of one chunk (one
  // 16-lane group of a 16-byte read) then fill all 64 banks, from any
  // starting pixel. Reads of pixels kNtRow and kNtRow + 1 wrap to pixels 0
  // and 1 of the same row: they feed only the last two pixels of a layer
  // row (p >= kNtRow - 2), and those reach only layer-4 pixels at or beyond
  // kNtOut, which are not written (each pixel is a column of its own in the
  // MFMA). 48 KB of rings + 4.2 KB of column features: three workgroups a
  // CU.
  __shared__ __attribute__((aligned(16))) __hip_bfloat16 ring[3][kNtDepth][kNtRow * kCh];
  __shared__ __hip_bfloat16 xin[kNtDepth][kNtRing];   // network input rows (one channel)
  // colf[k][c]: centred column features of input slot c (column x0 - 4 + c),
  // zero for k >= K and outside the map; lanes read consecutive slots
  __shared__ __attribute__((aligned(16))) float colf[kNtMaxK * kNtRing];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  // the role of this wave (0, 1, 2: layers 2, 3, 4; 3: input and layer 1),
  // rotated with the workgroup index so that the heavy roles of the
  // workgroups sharing a CU fall on different SIMDs
  const int wave = __builtin_amdgcn_readfirstlane(((tid >> 6) + (int)blockIdx.x) & 3);
  const NetJob jb = jobs[blockIdx.x];
  const NetPair pd = pairs[jb.pair];
  const int64_t H = pd.H, W = pd.W;
  const int x0 = (int)jb.x0, y0 = (int)jb.y0, y1 = (int)jb.y1;
  const float *st = st_all + pd.st_off;
  float *Z = Z_all + pd.z_off;
  const int px = lane & 31, kh = (lane >> 5) * 8;
  const int Hi = (int)H, Wi = (int)W;
  // element offset of chunk c (channels 8c .. 8c+7) of ring pixel p
  auto chunk_at = [](const int p, const int c) {
    const int q = p & (kNtRow - 1);
    return q * kCh + ((c ^ ((q >> 2) & 3)) << 3);
  };

  for (int q = tid; q < kNtRing * kNtMaxK; q += 256) {
    const int k = q / kNtRing, c = q - k * kNtRing;
    const int x = x0 - 4 + c;
    colf[q] = (k < K && x >= 0 && x < Wi) ? B[k * ldb + pd.b_off + x] - st[K + k] : 0.0f;
  }
  // zero rings and input rows
  for (int q = tid; q < 3 * kNtDepth * kNtRow * kCh / 8; q += 256)
    reinterpret_cast<lu32x4_t *>(&ring[0][0][0])[q] = lu32x4_t{0u, 0u, 0u, 0u};
  for (int q = tid; q < kNtDepth * kNtRing; q += 256) (&xin[0][0])[q] = __float2bfloat16(0.0f);

  // the layer this wave runs (waves 0, 1, 2: layers 2, 3, 4; wave 3 layer 1)
  const __hip_bfloat16 *Wl = wave == 0 ? W2 : W3;   // layers 2, 3 (layer 4: logit_frags)
  lbf16x8_t wb[18];
#pragma unroll
  for (int q = 0; q < 18; ++q)
    wb[q] = *reinterpret_cast<const lbf16x8_t *>(&Wl[px * kKdim + q * 16 + kh]);
  // first-layer taps (wave 3): k = dy*3 + dx of input channel 0
  lbf16x8_t w1;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const int k = kh + i;
    w1[i] = k < 9 ? (__bf16)bf2f(W1[px * kKdim + k * kCh]) : (__bf16)0.0f;
  }
  // biases of layers 2 and 3 in LDS, read where a row is stored (no
  // registers held across their MFMA chains): bias_s[role][channel]; the
  // other roles keep theirs in registers (256 bytes keep three workgroups
  // a CU within the LDS)
  __shared__ __attribute__((aligned(16))) float bias_s[2][kCh];
  if (wave < 2 && lane < kCh) bias_s[wave][lane] = (wave == 0 ? b2 : b3)[lane];
  const float inv = st[2 * K];
  // lane k (k < 16) of feat_of(y): centred row feature k of input row y,
  // zero for k >= K and outside the map
  auto feat_of = [&](const int y) {
    const int k = lane & (kNtMaxK - 1);
    return (k < K && y >= 0 && y < Hi) ? A[k * lda + pd.a_off + y] - st[k] : 0.0f;
  };
  // input row y into slot `slot` (wave 3) from its features f (lane k
  // holds feature k): bf16(-s * inv), s summed over 16 terms in order, zero
  // outside the map; lane l writes slots l and l + 64
  auto input_row = [&](const int y, const int slot, const float f) {
    const bool yok = y >= 0 && y < Hi;
    const int l1 = lane + 64 < kNtRing ? lane + 64 : 0;
    float s0 = 0.0f, s1 = 0.0f;
#pragma unroll
    for (int k = 0; k < kNtMaxK; ++k) {
      const float ck = __builtin_bit_cast(
          float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, f), k));
      s0 += ck * colf[k * kNtRing + lane];
      s1 += ck * colf[k * kNtRing + l1];
    }
    const int xa = x0 - 4 + lane, xb = xa + 64;
    xin[slot][lane] = (yok && xa >= 0 && xa < Wi) ? __float2bfloat16(-s0 * inv)
                                                  : __float2bfloat16(0.0f);
    if (lane + 64 < kNtRing)
      xin[slot][lane + 64] = (yok && xb >= 0 && xb < Wi) ? __float2bfloat16(-s1 * inv)
                                                         : __float2bfloat16(0.0f);
  };
  __syncthreads();
  // features of the next input row (wave 3), loaded a step ahead
  float f_next = 0.0f;
  if (wave == 3) {
    input_row(y0 - 4, 0, feat_of(y0 - 4));
    input_row(y0 - 3, 1, feat_of(y0 - 3));
    f_next = feat_of(y0 - 2);
  }
  // one 3x3 layer row from ring `src`: two tiles, k-steps in the order
  // (channel half, dy, dx) of k_conv3x3
  auto layer_row = [&](const __hip_bfloat16 *src, const int s, lfloatx16_t (&acc)[kNtT]) {
    // MFMA i = k-step q = i / 2 ((ch * 3 + dy) * 3 + dx) of tile t = i % 2;
    // its operand is loaded three MFMAs ahead, so the LDS latency overlaps
    // the chains
    // Tile 1 reads at tile 0's address plus 32 pixels (same bank swizzle),
    // so both tiles share their address registers; its pixels 64 and 65 then
    // lie past the ring row (the next slot's first pixels, or past the last
    // ring, inside the workgroup's LDS) instead of wrapping to pixels 0 and
    // 1: they feed only logits at or beyond kNtOut, which are not written.
    auto opnd = [&](const int i) {
      const int q = i >> 1, t = i & 1;
      const int ch = q / 9, dy = (q / 3) % 3, dx = q % 3;
      const __hip_bfloat16 *rowp = src + ((s + 1 + dy) % kNtDepth) * kNtRow * kCh;
      return *reinterpret_cast<const lbf16x8_t *>(
          &rowp[chunk_at(px + dx, ch * 2 + (kh >> 3)) + t * 32 * kCh]);
    };
#pragma unroll
    for (int t = 0; t < kNtT; ++t) acc[t] = lfloatx16_t{};
    lbf16x8_t b0 = opnd(0), b1 = opnd(1), b2 = opnd(2);
#pragma unroll
    for (int i = 0; i < 36; ++i) {
      const lbf16x8_t b3 = (i + 3 < 36) ? opnd(i + 3) : b2;
      // keep the load above this MFMA (the scheduler would sink it)
      __builtin_amdgcn_sched_barrier(0);
      const int q = i >> 1, t = i & 1;
      const int ch = q / 9, dy = (q / 3) % 3, dx = q % 3;
      acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wb[(dy * 3 + dx) * 2 + ch], b0, acc[t], 0,
                                                       0, 0);
      b0 = b1;
      b1 = b2;
      b2 = b3;
    }
  };
  // relu(acc + bias) of row y (columns xs0 + slot) into ring row `dst`,
  // zero outside the map
  auto store_row = [&](const lfloatx16_t (&acc)[kNtT], const lfloatx4_t (&bias)[4], const int y,
                       const int xs0, __hip_bfloat16 *dst) {
    const bool yok = y >= 0 && y < Hi;
    const int cq = (lane >> 5) * 4;
#pragma unroll
    for (int t = 0; t < kNtT; ++t) {
      const int x = xs0 + t * 32 + px;
      const bool ok = yok && x >= 0 && x < Wi;
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        float r[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          r[i] = acc[t][g * 4 + i] + bias[g][i];
          r[i] = r[i] > 0.0f ? r[i] : 0.0f;
        }
        *reinterpret_cast<lu32x2_t *>(&dst[chunk_at(t * 32 + px, g) + cq]) =
            ok ? lu32x2_t{pack_bf16x2(r[0], r[1]), pack_bf16x2(r[2], r[3])}
               : lu32x2_t{0u, 0u};
      }
    }
  };
  auto zero_row = [&](__hip_bfloat16 *dst) {
    for (int q = lane; q < kNtRow * kCh / 8; q += 64)
      reinterpret_cast<lu32x4_t *>(dst)[q] = lu32x4_t{0u, 0u, 0u, 0u};
  };
  // Step s: layer l (1..4) computes its row a - 2(l-1), a = y0 - 3 + s;
  // the rows layer l needs are [y0 - 4 + l, y1 + 4 - l). Every layer reads
  // rows written in the three previous steps (slots (s+1 .. s+3) % 4) and
  // writes its row into slot s % 4, which nobody reads in this step: one
  // barrier per step. Each role runs its own copy of the step loop (wave
  // uniform, the same barrier count), so each holds only its own registers.
  const int nsteps = (y1 - y0) + 9;
  if (wave == 3) {
    lfloatx4_t bb[4];
    load_bias4<0>(b1, lane, bb);
    auto step = [&](const int s, auto J) {
      constexpr int j = decltype(J)::value;
      const int a = y0 - 3 + s;          // first-layer row of this step
      // first layer, row a, from input rows a-1 .. a+1 (input row a+1 now,
      // into slot (s+2) % 4, whose row a-3 the first layer no longer reads)
      if (a + 1 < y1 + 4) {
        const float f = f_next;
        f_next = feat_of(a + 2);   // in flight until the next step
        input_row(a + 1, (j + 2) % kNtDepth, f);
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      if (a >= y0 - 3 && a < y1 + 3) {
        lfloatx16_t acc1[kNtT];
#pragma unroll
        for (int t = 0; t < kNtT; ++t) {
          lbf16x8_t in;
#pragma unroll
          for (int i = 0; i < 8; ++i) {
            const int k = kh + i;
            const int dy = k / 3, dx = k - (k / 3) * 3;
            const __bf16 xv = reinterpret_cast<const __bf16 *>(&xin[0][0])
                [k < 9 ? ((j + dy) % kNtDepth) * kNtRing + t * 32 + px + dx : 0];
            in[i] = k < 9 ? xv : (__bf16)0.0f;
          }
          acc1[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w1, in, lfloatx16_t{}, 0, 0, 0);
        }
        store_row(acc1, bb, a, x0 - 3, &ring[0][j][0]);
      }
      __syncthreads();
    };
    // steps four at a time: every slot a constant
    for (int s = 0; s < nsteps; s += 4) {
      step(s, std::integral_constant<int, 0>{});
      if (s + 1 < nsteps) step(s + 1, std::integral_constant<int, 1>{});
      if (s + 2 < nsteps) step(s + 2, std::integral_constant<int, 2>{});
      if (s + 3 < nsteps) step(s + 3, std::integral_constant<int, 3>{});
    }
  } else if (wave == 2) {
    // layer 4, the logits, from per-row partial sums (logit_frags): each
    // step takes the layer-3 row written in the step before, row yp = a - 5,
    // and completes logit row yp - 1 = a - 6 from the partials of rows
    // yp - 2, yp - 1 and yp (kept per tile in registers)
    lbf16x8_t wl[6];
    logit_frags(W4, lane, wl);
    const float bias0 = b4[0];
    float q02[kNtT], q01[kNtT], q11[kNtT];
#pragma unroll
    for (int t = 0; t < kNtT; ++t) q02[t] = q01[t] = q11[t] = 0.0f;
    const __hip_bfloat16 *src = &ring[2][0][0];
    auto step = [&](const int s, auto J) {
      constexpr int j = decltype(J)::value;
      const int yp = y0 - 8 + s;
      const int row = yp - 1;
      if (yp >= y0 - 1 && yp < y1 + 1) {
        const __hip_bfloat16 *rowp = src + ((j + 3) % kNtDepth) * kNtRow * kCh;
        float p0[kNtT], p1[kNtT], p2[kNtT];
#pragma unroll
        for (int t = 0; t < kNtT; ++t) {
          lfloatx16_t d = lfloatx16_t{};
#pragma unroll
          for (int q = 0; q < 6; ++q) {
            const int ch = q / 3, dx = q - ch * 3;
            const lbf16x8_t a = *reinterpret_cast<const lbf16x8_t *>(
                &rowp[chunk_at(t * 32 + px + dx, ch * 2 + (kh >> 3))]);
            d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wl[q], a, d, 0, 0, 0);
          }
          p0[t] = d[0];
          p1[t] = d[1];
          p2[t] = d[2];
        }
        if (row >= y0 && row < y1 && row >= 0 && row < Hi && lane < 32) {
          // logits of row `row`, columns x0 .. x0 + kNtOut - 1
#pragma unroll
          for (int t = 0; t < kNtT; ++t) {
            const int c = t * 32 + px, x = x0 + c;
            if (c < kNtOut && x < Wi)
              Z[(int64_t)row * W + x] = ((bias0 + q02[t]) + q11[t]) + p2[t];
          }
        }
#pragma unroll
        for (int t = 0; t < kNtT; ++t) {
          q02[t] = q01[t];
          q01[t] = p0[t];
          q11[t] = p1[t];
        }
      }
      __syncthreads();
    };
    for (int s = 0; s < nsteps; s += 4) {
      step(s, std::integral_constant<int, 0>{});
      if (s + 1 < nsteps) step(s + 1, std::integral_constant<int, 1>{});
      if (s + 2 < nsteps) step(s + 2, std::integral_constant<int, 2>{});
      if (s + 3 < nsteps) step(s + 3, std::integral_constant<int, 3>{});
    }
  } else {
    // wave w runs layer w + 2: row a - 2(w + 1), rows [y0 - 2 + w, y1 + 2 - w).
    // Steps run four at a time, so every ring slot is a constant offset.
    const __hip_bfloat16 *src = &ring[wave][0][0];
    __hip_bfloat16 *dst0 = &ring[wave + 1][0][0];
    auto step = [&](const int s, auto J) {
      constexpr int j = decltype(J)::value;
      const int row = y0 - 5 - 2 * wave + s;
      const bool run = row >= y0 - 2 + wave && row < y1 + 2 - wave;
      const bool inmap = row >= 0 && row < Hi;
      lfloatx16_t acc[kNtT];
      if (run && inmap) layer_row(src, j, acc);
      __hip_bfloat16 *dst = dst0 + j * kNtRow * kCh;
      if (run) {
        if (inmap) {
          lfloatx4_t bl[4];
#pragma unroll
          for (int g = 0; g < 4; ++g)
            bl[g] = *reinterpret_cast<const lfloatx4_t *>(&bias_s[wave][g * 8 + (lane >> 5) * 4]);
          store_row(acc, bl, row, x0 - 2 + wave, dst);
        }
        else zero_row(dst);
      }
      __syncthreads();
    };
    for (int s = 0; s < nsteps; s += 4) {
      step(s, std::integral_constant<int, 0>{});
      if (s + 1 < nsteps) step(s + 1, std::integral_constant<int, 1>{});
      if (s + 2 < nsteps) step(s + 2, std::integral_constant<int, 2>{});
      if (s + 3 < nsteps) step(s + 3, std::integral_constant<int, 3>{});
    }
  }
}

// Tile selection of the learned form on the device (the rule of
// tile_net.select_tiles with margin 0): tile t is selected when
// score[t] >= threshold (compared in FP64). Outputs:
//  * sel[t] (0/1);
//  * full: the plan's tile table with b_off remapped (selected tiles packed
//    back to back in plan order after the zero region of zero_words words,
//    pruned tiles at offset 0, the zero region) and c_off = -1 for pruned
//    tiles (no stage 1; stage 2 takes their edges as +inf);
//  * work: the selected tiles in plan order, c_off / ldc replaced by the
//    feature frames of the tile's first row and column;
//  * list: every tile once, diagonal by diagonal (diagonal i + j, first slot
//    diag_off[d]), the selected tiles of a diagonal first (in any order:
//    tiles of one diagonal are independent);
//  * counts: nsel, b_words (zero region included), selected cells, then
//    (first slot, selected, pruned) of every diagonal.
// Three launches over blocks of 1024 tiles: counters cleared; per block
// the selected count, words and cells, and the selected tiles of every
// diagonal (atomics); then every block its prefix over the blocks before
// it, a block scan and the writes.
constexpr int kSelThreads = 1024;
constexpr int kSelMaxDiag = 4096;

__device__ __forceinline__ int sel_diag(const TileDesc &td) {
  return (int)((td.ij >> 32) + (td.ij & 0xffffffffll));
}
__device__ __forceinline__ int64_t sel_words(const TileDesc &td) {
  return ((td.R + 15) / 16) * td.W;
}

__global__ void k_select_clear(int32_t *__restrict__ ctr, int64_t n) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) ctr[i] = 0;
}

// inclusive scan of v over the 1024 threads of the block (16 waves)
__device__ __forceinline__ int64_t sel_block_scan(int64_t v, int64_t *wsum, int64_t &total) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int64_t u = __shfl_up(v, d);
    if (lane >= d) v += u;
  }
  if (lane == 63) wsum[wv] = v;
  __syncthreads();
  int64_t before = 0;
  total = 0;
  for (int k = 0; k < kSelThreads / 64; ++k) {
    const int64_t x = wsum[k];
    before += k < wv ? x : 0;
    total += x;
  }
  __syncthreads();
  return v + before;
}

__global__ __launch_bounds__(kSelThreads) void k_select_count(
    const float *__restrict__ score, double threshold, const TileDesc *__restrict__ plan,
    int64_t ntiles, int64_t *__restrict__ bsum, int32_t *__restrict__ nlive) {
  __shared__ int64_t wsum[kSelThreads / 64];
  const int64_t t = blockIdx.x * (int64_t)kSelThreads + threadIdx.x;
  int64_t n = 0, w = 0, c = 0;
  if (t < ntiles && (double)score[t] >= threshold) {
    const TileDesc td = plan[t];
    n = 1;
    w = sel_words(td);
    c = td.R * td.W;
    atomicAdd(&nlive[sel_diag(td)], 1);
  }
  int64_t tn, tw, tc;
  sel_block_scan(n, wsum, tn);
  sel_block_scan(w, wsum, tw);
  sel_block_scan(c, wsum, tc);
  if (threadIdx.x == 0) {
    bsum[blockIdx.x * 3] = tn;
    bsum[blockIdx.x * 3 + 1] = tw;
    bsum[blockIdx.x * 3 + 2] = tc;
  }
}

__global__ __launch_bounds__(kSelThreads) void k_select_write(
    const float *__restrict__ score, double threshold, const TileDesc *__restrict__ plan,
    int64_t ntiles, const int64_t *__restrict__ frames, int64_t L,
    const int64_t *__restrict__ diag_off, int64_t max_diag, const int64_t *__restrict__ bsum,
    const int32_t *__restrict__ nlive, int32_t *__restrict__ lpos, int32_t *__restrict__ ppos,
    uint8_t *__restrict__ sel, TileDesc *__restrict__ full, TileDesc *__restrict__ work,
    int32_t *__restrict__ list, int64_t *__restrict__ counts) {
  __shared__ int64_t wsum[kSelThreads / 64];
  const int64_t nb = gridDim.x;
  // selected tiles and words of the blocks before this one, and the totals
  int64_t bn = 0, bw = 0, an = 0, aw = 0, ac = 0;
  for (int64_t b = 0; b < nb; ++b) {
    const int64_t x = bsum[b * 3], y = bsum[b * 3 + 1];
    if (b < blockIdx.x) {
      bn += x;
      bw += y;
    }
    an += x;
    aw += y;
    ac += bsum[b * 3 + 2];
  }
  const int64_t t = blockIdx.x * (int64_t)kSelThreads + threadIdx.x;
  const bool in = t < ntiles;
  TileDesc td{};
  bool on = false;
  if (in) {
    td = plan[t];
    on = (double)score[t] >= threshold;
  }
  const int64_t w = on ? sel_words(td) : 0;
  int64_t tn, tw;
  const int64_t pn = sel_block_scan(on ? 1 : 0, wsum, tn) - (on ? 1 : 0);
  const int64_t pw = sel_block_scan(w, wsum, tw) - w;
  const int64_t zero_words = ((L + 15) / 16) * L;
  if (in) {
    const int d = sel_diag(td);
    sel[t] = on ? 1 : 0;
    int32_t slot;
    if (on) {
      td.b_off = zero_words + bw + pw;
      full[t] = td;
      TileDesc wk = td;
      const int64_t hop = L - 1, pr = td.pair;
      wk.c_off = frames[pr * 2] + (td.ij >> 32) * hop;
      wk.ldc = frames[pr * 2 + 1] + (td.ij & 0xffffffffll) * hop;
      work[bn + pn] = wk;
      slot = (int32_t)diag_off[d] + atomicAdd(&lpos[d], 1);
    } else {
      td.b_off = 0;
      td.c_off = -1;
      full[t] = td;
      slot = (int32_t)diag_off[d] + nlive[d] + atomicAdd(&ppos[d], 1);
    }
    list[slot] = (int32_t)t;
  }
  if (blockIdx.x == 0) {
    if (threadIdx.x == 0) {
      counts[0] = an;
      counts[1] = zero_words + aw;
      counts[2] = ac;
    }
    for (int64_t d = threadIdx.x; d < max_diag; d += kSelThreads) {
      counts[3 + 3 * d] = diag_off[d];
      counts[3 + 3 * d + 1] = nlive[d];
      counts[3 + 3 * d + 2] = (diag_off[d + 1] - diag_off[d]) - nlive[d];
    }
  }
}

// Selection check of the learned form: flag[p] = 1 when a tile that holds a
// cell within `guard` frames (in row and column) of a point of pair p's
// stitched path is not selected.  The tiles holding frames [x0, x1] are
// max(x0-1, 0)/hop .. min(x1/hop, n-1) (a frame on an overlap belongs to two
// tiles).  Grid: (point blocks, pairs).
__global__ __launch_bounds__(256) void k_selection_check(const PairDesc *__restrict__ pairs,
                                                         const int64_t *__restrict__ path,
                                                         const int64_t *__restrict__ path_len,
                                                         const uint8_t *__restrict__ sel, int64_t L,
                                                         int64_t guard, int32_t *__restrict__ flag) {
  const int64_t p = blockIdx.y;
  const PairDesc pd = pairs[p];
  const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (q >= path_len[p]) return;
  const int64_t hop = L - 1;
  const int64_t r = path[(pd.path_off + q) * 2], c = path[(pd.path_off + q) * 2 + 1];
  const int64_t ra = r - guard - 1 > 0 ? r - guard - 1 : 0;
  const int64_t rb = r + guard < pd.L1 - 1 ? r + guard : pd.L1 - 1;
  const int64_t ca = c - guard - 1 > 0 ? c - guard - 1 : 0;
  const int64_t cb = c + guard < pd.L2 - 1 ? c + guard : pd.L2 - 1;
  const int64_t i0 = ra / hop < pd.n1 - 1 ? ra / hop : pd.n1 - 1;
  const int64_t i1 = rb / hop < pd.n1 - 1 ? rb / hop : pd.n1 - 1;
  const int64_t j0 = ca / hop < pd.n2 - 1 ? ca / hop : pd.n2 - 1;
  const int64_t j1 = cb / hop < pd.n2 - 1 ? cb / hop : pd.n2 - 1;
  bool bad = false;
  for (int64_t i = i0; i <= i1; ++i)
    for (int64_t j = j0; j <= j1; ++j) bad = bad || !sel[pd.tile_base + i * pd.n2 + j];
  if (bad) atomicOr(&flag[p], 1);
}

inline unsigned nb256(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace pf

using namespace pf;

extern "C" int pf_pool_features(const double *d_F, int64_t K, int64_t n, int pool, float *d_Fp,
                                void *stream) {
  PF_REQUIRE(d_F && d_Fp && K >= 1 && n >= 1, "pf_pool_features: bad arguments");
  PF_REQUIRE(pool == 32, "pf_pool_features: the pooling factor is 32");
  const int64_t nb = (n + pool - 1) / pool;
  const PoolSeg one{0, n, 0, 0};
  auto kern = K <= 16 ? k_pool_features<false, true> : k_pool_features<false, false>;
  hipLaunchKernelGGL(kern, dim3((unsigned)((nb + kPoolBlocks - 1) / kPoolBlocks)), dim3(256), 0, as_stream(stream), d_F,
                     K, n, (const PoolSeg *)nullptr, (int64_t)1, one, d_Fp, nb, nb,
                     (__hip_bfloat16 *)nullptr, (int64_t)0);
  PF_CHECK_HIP(hipGetLastError());
  return PF_OK;
}

extern "C" int pf_pool_features_batch(const double *d_F, int64_t K, int64_t n,
                                      const int64_t *d_seg, int64_t nseg, int pool, float *d_Fp,
                                      int64_t nb, void *stream) {
  PF_REQUIRE(d_F && d_seg && d_Fp && K >= 1 && n >= 1 && nseg >= 1 && nb >= 1,
             "pf_pool_features_batch: bad arguments");
  PF_REQUIRE(pool == 32, "pf_pool_features_batch: the pooling factor is 32");
  auto kern = K <= 16 ? k_pool_features<false, true> : k_pool_features<false, false>;
  hipLaunchKernelGGL(kern, dim3((unsigned)((nb + kPoolBlocks - 1) / kPoolBlocks)), dim3(256), 0, as_stream(stream), d_F,
                     K, n, reinterpret_cast<const PoolSeg *>(d_seg), nseg, PoolSeg{0, 0, 0, 0},
                     d_Fp, nb, nb, (__hip_bfloat16 *)nullptr, (int64_t)0);
  PF_CHECK_HIP(hipGetLastError());
  return PF_OK;
}

extern "C" int pf_pool_features_bf16_batch(const double *d_F, int64_t K, int64_t n,
                                           const int64_t *d_seg, int64_t nseg, int pool,
                                           float *d_Fp, int64_t nb, void *d_Fb, void *stream) {
  PF_REQUIRE(d_F && d_seg && d_Fp && d_Fb && K >= 1 && n >= 1 && nseg >= 1 && nb >= 1,
             "pf_pool_features_bf16_batch: bad arguments");
  PF_REQUIRE(pool == 32, "pf_pool_features_bf16_batch: the pooling factor is 32");
  PF_REQUIRE(reinterpret_cast<uintptr_t>(d_Fb) % 16 == 0,
             "pf_pool_features_bf16_batch: Fb must be 16-byte aligned");
  const int64_t Kp = (K + 15) / 16 * 16;
  auto kern = K <= 16 ? k_pool_features<true, true> : k_pool_features<true, false>;
  hipLaunchKernelGGL(kern, dim3((unsigned)((nb + kPoolBlocks - 1) / kPoolBlocks)), dim3(256), 0, as_stream(stream), d_F,
                     K, n, reinterpret_cast<const PoolSeg *>(d_seg), nseg, PoolSeg{0, 0, 0, 0},
                     d_Fp, nb, nb, static_cast<__hip_bfloat16 *>(d_Fb), Kp);
  PF_CHECK_HIP(hipGetLastError());
  return PF_OK;
}

extern "C" int pf_input_stats_batch(const float *d_A, int64_t lda, const float *d_B, int64_t ldb,
                                    int64_t K, const int64_t *d_pairs, int64_t npairs,
                                    float *d_st, void *stream) {
  PF_REQUIRE(d_A && d_B && d_pairs && d_st && K >= 1 && npairs >= 1,
             "pf_input_stats_batch: bad arguments");
  PF_REQUIRE(K <= 16, "pf_input_stats_batch: at most 16 features per frame");
  PF_REQUIRE(npairs < (1ll << 31), "pf_input_stats_batch: too many pairs");
  hipLaunchKernelGGL(k_input_stats_seg<16>, dim3((unsigned)npairs), dim3(256), 0,
                     as_stream(stream), d_A, lda, d_B, ldb, (int)K,
                     reinterpret_cast<const StatPair *>(d_pairs), StatPair{0, 0, 0, 0}, d_st);
  PF_CHECK_HIP(hipGetLastError());
  return PF_OK;
}

extern "C" int pf_conv3x3_first_pooled(const float *d_A, int64_t lda, int64_t a_off, int64_t H,
                                       const float *d_B, int64_t ldb, int64_t b_off, int64_t W,
                                       int64_t K, const float *d_st, const void *d_Wt,
                                       const float *d_bias, void *d_Y, void *stream) {
  PF_REQUIRE(d_A && d_B && d_st && d_Wt && d_bias && d_Y && H >= 1 && W >= 1 && K >= 1 &&
                 a_off >= 0 && b_off >= 0 && a_off + H <= lda && b_off + W <= ldb,
             "pf_conv3x3_first_pooled: bad arguments");
  PF_REQUIRE(K <= 16, "pf_conv3x3_first_pooled: at most 16 features per frame");
  PF_REQUIRE(H * W * kCh < (1ll << 31),
             "pf_conv3x3_first_pooled: image too large (H * W * 32 must be below 2^31)");
  PF_REQUIRE(reinterpret_cast<uintptr_t>(d_Y) % 16 == 0 && reinterpret_cast<uintptr_t>(d_bias) % 16 == 0,
             "pf_conv3x3_first_pooled: Y and bias must be 16-byte aligned");
  const int64_t ncol = (W + kCvCols - 1) / kCvCols, nrow = (H + kCvRows - 1) / kCvRows;
  int64_t segl = (ncol * nrow) / 768;
  segl = segl < 1 ? 1 : (segl > 8 ? 8 : segl);
  const int seg = (int)segl;
  PF_REQUIRE((nrow + seg - 1) / seg < 65536, "pf_conv3x3_first_pooled: image too large");
  const dim3 grid((unsigned)ncol, (unsigned)((nrow + seg - 1) / seg));
  hipLaunchKernelGGL(k_conv3x3_first_pooled, grid, dim3(256), 0, as_stream(stream), d_A, lda,
                     a_off, H, d_B, ldb, b_off, W, (int)K, d_st,
                     static_cast<const __hip_bfloat16 *>(d_Wt), d_bias,
                     static_cast<__hip_bfloat16 *>(d_Y), seg);
  PF_CHECK_HIP(hipGetLastError());
  return PF_OK;
}

extern "C" int pf_pooled_cost_workspace(int64_t K, int64_t *floats) {
  PF_REQUIRE(floats && K >= 1, "pf_pooled_cost_workspace: bad arguments");
  *floats = 2 * K + 2;
  return PF_OK;
}

extern "C" int pf_pooled_cost(const float *d_A, int64_t H, const float *d_B, int64_t W, int64_t K,
                              void *d_X, void *d_Xc, float *d_P, float *d_work, void *stream) {
  PF_REQUIRE(d_A && d_B && (d_X || d_Xc) && d_work && H >= 1 && W >= 1 && K >= 1,
             "pf_pooled_cost: bad arguments");
  hipStream_t st = as_stream(stream);
  if (d_P) {
    hipLaunchKernelGGL(k_pooled_cost, dim3(nb256(H * W)), dim3(256), 0, st, d_A, H, d_B, W, K,
                       d_P);
    PF_CHECK_HIP(hipGetLastError());
  }
  if (K <= 16)
    hipLaunchKernelGGL(k_input_stats_seg<16>, dim3(1), dim3(256), 0, st, d_A, H, d_B, W, (int)K,
                       (const StatPair *)nullptr, StatPair{0, H, 0, W}, d_work);
  else
    hipLaunchKernelGGL(k_input_stats_any, dim3(1), dim3(256), 0, st, d_A, H, d_B, W, K, d_work);
  PF_CHECK_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_write_input, dim3(nb256(H * W)), dim3(256), 0, st, d_A, H, d_B, W, K,
                     d_work, static_cast<__hip_bfloat16 *>(d_X),
                     static_cast<__hip_bfloat16 *>(d_Xc));
  PF_CHECK_HIP(hipGetLastError());
  return PF_OK;
}

extern "C" int pf_conv3x3(int mode, const void *d_X, int64_t H, int64_t W, const void *d_Wt,
                          const float *d_bias, const void *d_M, void *d_Y, void *stream) {
  PF_REQUIRE(d_X && d_Wt && d_Y && H >= 1 && W >= 1, "pf_conv3x3: bad arguments");
  PF_REQUIRE(mode == 2 ? d_M != nullptr : d_bias != nullptr, "pf_conv3x3: missing bias or mask");
  PF_REQUIRE(reinterpret_cast<uintptr_t>(d_X) % 16 == 0 && reinterpret_cast<uintptr_t>(d_Wt) % 16 == 0 &&
                 reinterpret_cast<uintptr_t>(d_bias) % 16 == 0 &&
                 reinterpret_cast<uintptr_t>(d_Y) % 8 == 0 && reinterpret_cast<uintptr_t>(d_M) % 8 == 0,
             "pf_conv3x3: X, weights and bias must be 16-byte aligned, Y and M 8-byte aligned");
  PF_REQUIRE(H * W * kCh < (1ll << 31) && (H + kC1Rows - 1) / kC1Rows < 65536,
             "pf_conv3x3: image too large (H * W * 32 must be below 2^31)");
  // blocks of one column strip per workgroup: about 768 workgroups or more
  const int64_t ncol = (W + kCvCols - 1) / kCvCols, nrow = (H + kCvRows - 1) / kCvRows;
  int64_t segl = (ncol * nrow) / 768;
  segl = segl < 1 ? 1 : (segl > 8 ? 8 : segl);
  const int seg = (int)segl;
  const dim3 grid((unsigned)ncol, (unsigned)((nrow + seg - 1) / seg));
  const dim3 grid1((unsigned)((W + kC1Cols - 1) / kC1Cols),
                   (unsigned)((H + kC1Rows - 1) / kC1Rows));
  const auto *X = static_cast<const __hip_bfloat16 *>(d_X);
  const auto *Wt = static_cast<const __hip_bfloat16 *>(d_Wt);
  const auto *M = static_cast<const __hip_bfloat16 *>(d_M);
  hipStream_t st = as_stream(stream);
  if (mode == 0)
    hipLaunchKernelGGL(k_conv3x3<0>, grid, dim3(256), 0, st, X, H, W, Wt, d_bias, M, d_Y, seg);
  else if (mode == 1)
    hipLaunchKernelGGL(k_conv3x3<1>, grid, dim3(256), 0, st, X, H, W, Wt, d_bias, M, d_Y, seg);
  else if (mode == 2)
    hipLaunchKernelGGL(k_conv3x3<2>, grid, dim3(256), 0, st, X, H, W, Wt, d_bias, M, d_Y, seg);
  else if (mode == 3)
    hipLaunchKernelGGL(k_conv3x3_first, grid1, dim3(256), 0, st, X, H, W, Wt, d_bias,
                       static_cast<__hip_bfloat16 *>(d_Y));
  else
    PF_REQUIRE(false, "pf_conv3x3: mode must be 0, 1, 2 or 3");
  PF_CHECK_HIP(hipGetLastError());
  return PF_OK;
}

extern "C" int pf_conv3x3_wgrad_workspace(int64_t H, int64_t W, int64_t *floats) {
  PF_REQUIRE(floats && H >= 1 && W >= 1, "pf_conv3x3_wgrad_workspace: bad arguments");
  *floats = ((H * W + kWgChunk - 1) / kWgChunk) * (int64_t)kWgSize;
  return PF_OK;
}

extern "C" int pf_conv3x3_wgrad(const void *d_dY, const void *d_X, int64_t H, int64_t W,
                                float *d_grad, float *d_part, void *stream) {
  PF_REQUIRE(d_dY && d_X && d_grad && d_part && H >= 1 && W >= 1, "pf_conv3x3_wgrad: bad arguments");
  const int64_t nblk = (H * W + kWgChunk - 1) / kWgChunk;
  hipStream_t st = as_stream(stream);
  hipLaunchKernelGGL(k_conv3x3_wgrad, dim3((unsigned)nblk), dim3(64), 0, st,
                     static_cast<const __hip_bfloat16 *>(d_dY),
                     static_cast<const __hip_bfloat16 *>(d_X), H, W, d_part);
  PF_CHECK_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_reduce_add, dim3(nb256(kWgSize)), dim3(256), 0, st, d_part, nblk,
                     (int64_t)kWgSize, d_grad);
  PF_CHECK_HIP(hipGetLastError());
  return PF_OK;
}

extern "C" int pf_bce(const float *d_Z, const uint8_t *d_label, int64_t npix, float pos_weight,
                      float scale, void *d_dZ, float *d_loss, void *stream) {
  PF_REQUIRE(d_Z && d_label && d_dZ && d_loss && npix >= 1, "pf_bce: bad arguments");
  hipLaunchKernelGGL(k_bce, dim3(nb256(npix)), dim3(256), 0, as_stream(stream), d_Z, d_label, npix,
                     pos_weight, scale, static_cast<__hip_bfloat16 *>(d_dZ), d_loss);
  PF_CHECK_HIP(hipGetLastError());
  return PF_OK;
}

extern "C" int pf_adam(float *d_w, float *d_g, float *d_m, float *d_v, const uint8_t *d_frozen,
                       int64_t n, float lr, float b1, float b2, float eps, int64_t step,
                       void *stream) {
  PF_REQUIRE(d_w && d_g && d_m && d_v && n >= 1 && step >= 1, "pf_adam: bad arguments");
  const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3(nb256(n)), dim3(256), 0, as_stream(stream), d_w, d_g, d_m, d_v,
                     d_frozen, n, lr, b1, b2, eps, c1, c2);
  PF_CHECK_HIP(hipGetLastError());
  return PF_OK;
}

extern "C" int pf_pack_weights(const float *d_w, void *d_Wt, void *d_Wf, void *stream) {
  PF_REQUIRE(d_w && d_Wt, "pf_pack_weights: bad arguments");
  hipLaunchKernelGGL(k_pack_weights, dim3(nb256(kCh * kKdim)), dim3(256), 0, as_stream(stream), d_w,
                     static_cast<__hip_bfloat16 *>(d_Wt), static_cast<__hip_bfloat16 *>(d_Wf));
  PF_CHECK_HIP(hipGetLastError());
  return PF_OK;
}

extern "C" int pf_tile_scores(const float *d_Z, int64_t H, int64_t W, int pool, int64_t L,
                              int64_t L1, int64_t L2, int ext, float *d_score, void *stream) {
  PF_REQUIRE(d_Z && d_score && H >= 1 && W >= 1 && pool >= 1 && L >= 2 && L1 >= 2 && L2 >= 2 &&
                 ext >= 0,
             "pf_tile_scores: bad arguments");
  PF_REQUIRE((L1 + pool - 1) / pool == H && (L2 + pool - 1) / pool == W,
             "pf_tile_scores: pooled map does not match the pair");
  PF_REQUIRE(H * W < (1ll << 31), "pf_tile_scores: pooled map too large");
  const int64_t hop = L - 1;
  const int64_t n1 = (L1 - 1 + hop - 1) / hop, n2 = (L2 - 1 + hop - 1) / hop;
  hipLaunchKernelGGL(k_tile_scores, dim3((unsigned)(n1 * n2)), dim3(256), 0, as_stream(stream), d_Z,
                     H, W, pool, L, L1, L2, n2, ext, d_score);
  PF_CHECK_HIP(hipGetLastError());
  return PF_OK;
}

extern "C" int pf_tile_net_geometry(int64_t *cols_per_job, int64_t *rows_per_job) {
  PF_REQUIRE(cols_per_job && rows_per_job, "pf_tile_net_geometry: null output");
  *cols_per_job = kNtOut;
  *rows_per_job = kNtSeg;
  return PF_OK;
}

extern "C" int pf_tile_net(const float *d_A, int64_t lda, const float *d_B, int64_t ldb, int64_t K,
                           const float *d_st, const int64_t *d_pairs, int64_t npairs,
                           const int64_t *d_jobs, int64_t njobs, const void *d_W1,
                           const float *d_b1, const void *d_W2, const float *d_b2,
                           const void *d_W3, const float *d_b3, const void *d_W4,
                           const float *d_b4, float *d_Z, void *stream) {
  PF_REQUIRE(d_A && d_B && d_st && d_pairs && d_jobs && d_W1 && d_b1 && d_W2 && d_b2 && d_W3 &&
                 d_b3 && d_W4 && d_b4 && d_Z && npairs >= 1 && njobs >= 0 && K >= 1,
             "pf_tile_net: bad arguments");
  PF_REQUIRE(K <= kNtMaxK, "pf_tile_net: at most 16 features per frame");
  PF_REQUIRE(njobs < (1ll << 31), "pf_tile_net: too many jobs");
  const uintptr_t al = reinterpret_cast<uintptr_t>(d_W1) | reinterpret_cast<uintptr_t>(d_W2) |
                       reinterpret_cast<uintptr_t>(d_W3) | reinterpret_cast<uintptr_t>(d_W4) |
                       reinterpret_cast<uintptr_t>(d_b1) | reinterpret_cast<uintptr_t>(d_b2) |
                       reinterpret_cast<uintptr_t>(d_b3) | reinterpret_cast<uintptr_t>(d_b4);
  PF_REQUIRE(al % 16 == 0, "pf_tile_net: weights and biases must be 16-byte aligned");
  if (njobs == 0) return PF_OK;
  hipLaunchKernelGGL(k_tile_net, dim3((unsigned)njobs), dim3(256), 0, as_stream(stream), d_A, lda,
                     d_B, ldb, (int)K, d_st, reinterpret_cast<const NetPair *>(d_pairs),
                     reinterpret_cast<const NetJob *>(d_jobs),
                     static_cast<const __hip_bfloat16 *>(d_W1), d_b1,
                     static_cast<const __hip_bfloat16 *>(d_W2), d_b2,
                     static_cast<const __hip_bfloat16 *>(d_W3), d_b3,
                     static_cast<const __hip_bfloat16 *>(d_W4), d_b4, d_Z);
  PF_CHECK_HIP(hipGetLastError());
  return PF_OK;
}

extern "C" int pf_tile_scores_batch(const float *d_Z, const int64_t *d_pairs,
                                    const int32_t *d_tpair, int64_t ntiles, int pool, int64_t L,
                                    int ext, float *d_score, void *stream) {
  PF_REQUIRE(d_Z && d_pairs && d_tpair && d_score && ntiles >= 0 && pool >= 1 && L >= 2 &&
                 ext >= 0,
             "pf_tile_scores_batch: bad arguments");
  if (ntiles == 0) return PF_OK;
  const int64_t gy = (ntiles + 65535) / 65536;
  PF_REQUIRE(gy < 65536, "pf_tile_scores_batch: too many tiles");
  hipLaunchKernelGGL(k_tile_scores_batch, dim3((unsigned)(ntiles < 65536 ? ntiles : 65536),
                                               (unsigned)gy),
                     dim3(256), 0, as_stream(stream), d_Z,
                     reinterpret_cast<const ScorePair *>(d_pairs), d_tpair, ntiles, pool, L, ext,
                     d_score);
  PF_CHECK_HIP(hipGetLastError());
  return PF_OK;
}

extern "C" int pf_select_tiles(const float *d_score, double threshold, const int64_t *d_plan,
                               int64_t ntiles, const int64_t *d_frames, int64_t L,
                               const int64_t *d_diag_off, int64_t max_diag, uint8_t *d_sel,
                               int64_t *d_full, int64_t *d_work, int32_t *d_list,
                               int64_t *d_counts, int64_t *d_tmp, void *stream) {
  PF_REQUIRE(d_score && d_plan && d_frames && d_diag_off && d_sel && d_full && d_work &&
                 d_list && d_counts && d_tmp && ntiles >= 1 && L >= 2 && max_diag >= 1,
             "pf_select_tiles: bad arguments");
  PF_REQUIRE(max_diag <= kSelMaxDiag, "pf_select_tiles: more than 4096 tile diagonals");
  PF_REQUIRE(ntiles < (1ll << 31), "pf_select_tiles: too many tiles");
  const int64_t nb = (ntiles + kSelThreads - 1) / kSelThreads;
  PF_REQUIRE(nb <= 65535, "pf_select_tiles: too many tiles for one launch");
  const hipStream_t st = as_stream(stream);
  int64_t *bsum = d_tmp;
  int32_t *ctr = reinterpret_cast<int32_t *>(d_tmp + 3 * nb);   // nlive, lpos, ppos
  hipLaunchKernelGGL(k_select_clear, dim3((unsigned)((3 * max_diag + 255) / 256)), dim3(256), 0,
                     st, ctr, 3 * max_diag);
  PF_CHECK_HIP(hipGetLastError());
  const TileDesc *plan = reinterpret_cast<const TileDesc *>(d_plan);
  hipLaunchKernelGGL(k_select_count, dim3((unsigned)nb), dim3(kSelThreads), 0, st, d_score,
                     threshold, plan, ntiles, bsum, ctr);
  PF_CHECK_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_select_write, dim3((unsigned)nb), dim3(kSelThreads), 0, st, d_score,
                     threshold, plan, ntiles, d_frames, L, d_diag_off, max_diag, bsum, ctr,
                     ctr + max_diag, ctr + 2 * max_diag, d_sel, reinterpret_cast<TileDesc *>(d_full),
                     reinterpret_cast<TileDesc *>(d_work), d_list, d_counts);
  PF_CHECK_HIP(hipGetLastError());
  return PF_OK;
}

extern "C" int pf_selection_check(const int64_t *d_pairs, int64_t npairs, int64_t max_path,
                                  const int64_t *d_path, const int64_t *d_path_len,
                                  const uint8_t *d_sel, int64_t L, int64_t guard, int32_t *d_flag,
                                  void *stream) {
  PF_REQUIRE(d_pairs && d_path && d_path_len && d_sel && d_flag && npairs >= 1 && L >= 2 &&
                 guard >= 0 && 2 * guard < L - 1 && max_path >= 0,
             "pf_selection_check: bad arguments");
  PF_REQUIRE(npairs <= 65535, "pf_selection_check: too many pairs for one launch");
  if (max_path == 0) return PF_OK;
  hipLaunchKernelGGL(k_selection_check, dim3(nb256(max_path), (unsigned)npairs), dim3(256), 0,
                     as_stream(stream), reinterpret_cast<const PairDesc *>(d_pairs), d_path,
                     d_path_len, d_sel, L, guard, d_flag);
  PF_CHECK_HIP(hipGetLastError());
  return PF_OK;
}
