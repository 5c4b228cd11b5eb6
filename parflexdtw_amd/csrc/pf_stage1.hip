// Stage 1 of ParFlex on gfx950: per-tile FlexDTW dynamic programme.
//
// Semantics follow FlexDTW.flexdtw (FlexDTW.py:1320-1364) exactly, in IEEE
// arithmetic with every multiply and add rounded on its own (the library is
// compiled with -ffp-contract=off):
//   boundary: D[0,:] = C[0,:], D[:,0] = C[:,0], P = B = 0 there;
//   interior: for step k in order, pathcost = D[pred] + C*w_k,
//             mdist = row + col - |P[pred]|, cpm = pathcost / mdist,
//             strict '<' keeps the first minimum;
//             D = pathcost of the winner, B = k, P = start column (pred on
//             row 0), -start row (pred on column 0) or P[pred].
//
// Mapping: one workgroup per tile; the tile's columns are split over all
// lanes (CPL consecutive columns per lane) and the workgroup sweeps the rows
// top to bottom.  A cell depends only on rows r-1 and r-2 (every step has
// dr >= 1), so a whole row is computed in parallel; each lane keeps its
// columns of rows r-1 and r-2 in registers, takes the two columns to its left
// from lane-1 by a lane shuffle, and lane 0 of each wave reads them from the
// previous wave through a 3-slot LDS ring.  One barrier per row.
//
// Outputs (packed mode): 2-bit back-pointers packed 16 rows per uint32 per
// column, and D/P on the tile's last row and last column only -- everything
// stage 2 needs (Parflex.py:509-564, :1063-1064).  Full mode writes D/P/B for
// every cell (the 6-tuple of gpu_flexdtw.flexdtw_chunk_from_global_C).
#include "pf_common.h"

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

// One tile.  FULL selects full D/P/B outputs instead of packed B + edges.
template <int CPL, typename CT, typename DT, bool FULL>
__device__ __forceinline__ void tile_dp(const TileDesc td, const CT *__restrict__ Cb,
                                        const StepSet &ss, int64_t Lmax,
                                        uint32_t *__restrict__ Bp, double *__restrict__ eD,
                                        int32_t *__restrict__ eP, double *__restrict__ Df,
                                        int32_t *__restrict__ Pf, int8_t *__restrict__ Bf) {
  static_assert(CPL >= 2, "a lane needs two columns of left context");
  __shared__ DT shD[3][16][2];
  __shared__ int shP[3][16][2];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int R = (int)td.R, W = (int)td.W;
  const CT *__restrict__ C = Cb + td.c_off;
  const int64_t ldc = td.ldc;
  const int c0 = tid * CPL;
  const DT INF = dp_inf<DT>();

  DT wk[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) wk[k] = (DT)ss.w[k];

  DT D1[CPL], D2[CPL], Cn[CPL];
  int P1[CPL], P2[CPL];
  uint32_t bacc[CPL];

#pragma unroll
  for (int j = 0; j < CPL; ++j) {
    const int c = c0 + j;
    D1[j] = (c < W) ? load_c<DT>(C + c) : INF;
    P1[j] = 0;
    D2[j] = INF;
    P2[j] = 0;
    bacc[j] = 0u;
    if (c < W) {
      if constexpr (FULL) {
        Df[c] = (double)D1[j];
        Pf[c] = 0;
        Bf[c] = 0;
      } else {
        if (c == W - 1) {
          eD[td.e_off + Lmax] = (double)D1[j];
          eP[td.e_off + Lmax] = 0;
        }
        if (R == 1) {
          eD[td.e_off + c] = (double)D1[j];
          eP[td.e_off + c] = 0;
          Bp[td.b_off + c] = 0u;
        }
      }
    }
  }
  if (R > 1) {
#pragma unroll
    for (int j = 0; j < CPL; ++j) {
      const int c = c0 + j;
      Cn[j] = (c < W) ? load_c<DT>(C + ldc + c) : (DT)0;
    }
  }
  if (lane == 63) {
    shD[0][wave][0] = D1[CPL - 2];
    shD[0][wave][1] = D1[CPL - 1];
    shP[0][wave][0] = P1[CPL - 2];
    shP[0][wave][1] = P1[CPL - 1];
  }
  __syncthreads();

  for (int r = 1; r < R; ++r) {
    DT Cc[CPL];
#pragma unroll
    for (int j = 0; j < CPL; ++j) Cc[j] = Cn[j];
    if (r + 1 < R) {
      const CT *row = C + (int64_t)(r + 1) * ldc;
#pragma unroll
      for (int j = 0; j < CPL; ++j) {
        const int c = c0 + j;
        Cn[j] = (c < W) ? load_c<DT>(row + c) : (DT)0;
      }
    }

    // Left context: columns c0-2, c0-1 of rows r-1 (E1) and r-2 (E2).
    DT a1 = __shfl_up(D1[CPL - 2], 1), b1 = __shfl_up(D1[CPL - 1], 1);
    DT a2 = __shfl_up(D2[CPL - 2], 1), b2 = __shfl_up(D2[CPL - 1], 1);
    int pa1 = __shfl_up(P1[CPL - 2], 1), pb1 = __shfl_up(P1[CPL - 1], 1);
    int pa2 = __shfl_up(P2[CPL - 2], 1), pb2 = __shfl_up(P2[CPL - 1], 1);
    if (lane == 0) {
      if (wave > 0) {
        const int s1 = (r + 2) % 3, s2 = (r + 1) % 3;  // rows r-1, r-2
        a1 = shD[s1][wave - 1][0];
        b1 = shD[s1][wave - 1][1];
        pa1 = shP[s1][wave - 1][0];
        pb1 = shP[s1][wave - 1][1];
        a2 = shD[s2][wave - 1][0];
        b2 = shD[s2][wave - 1][1];
        pa2 = shP[s2][wave - 1][0];
        pb2 = shP[s2][wave - 1][1];
      } else {
        // Columns -2, -1 do not exist; every step reaching them is invalid.
        a1 = b1 = a2 = b2 = INF;
        pa1 = pb1 = pa2 = pb2 = 0;
      }
    }
    DT E1[CPL + 2], E2[CPL + 2];
    int Q1[CPL + 2], Q2[CPL + 2];
    E1[0] = a1;
    E1[1] = b1;
    E2[0] = a2;
    E2[1] = b2;
    Q1[0] = pa1;
    Q1[1] = pb1;
    Q2[0] = pa2;
    Q2[1] = pb2;
#pragma unroll
    for (int j = 0; j < CPL; ++j) {
      E1[j + 2] = D1[j];
      E2[j + 2] = D2[j];
      Q1[j + 2] = P1[j];
      Q2[j + 2] = P2[j];
    }

    DT ND[CPL];
    int NP[CPL];
    int code[CPL];
#pragma unroll
    for (int j = 0; j < CPL; ++j) {
      const int c = c0 + j;
      const DT cval = Cc[j];
      if (c == 0) {
        ND[j] = cval;
        NP[j] = 0;
        code[j] = 0;
        continue;
      }
      DT best = INF, bestD = INF;
      int kb = 0, bestPp = 0;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        if (k < ss.K) {
          const int dr = ss.dr[k], dc = ss.dc[k];
          if (r - dr >= 0 && c - dc >= 0) {
            const DT v1 = (dc == 0) ? E1[j + 2] : ((dc == 1) ? E1[j + 1] : E1[j]);
            const DT v2 = (dc == 0) ? E2[j + 2] : ((dc == 1) ? E2[j + 1] : E2[j]);
            const int q1 = (dc == 0) ? Q1[j + 2] : ((dc == 1) ? Q1[j + 1] : Q1[j]);
            const int q2 = (dc == 0) ? Q2[j + 2] : ((dc == 1) ? Q2[j + 1] : Q2[j]);
            const DT Dp = (dr == 1) ? v1 : v2;
            const int Pp = (dr == 1) ? q1 : q2;
            const DT prod = cval * wk[k];
            const DT pathcost = Dp + prod;
            const int mdist = r + c - (Pp >= 0 ? Pp : -Pp);
            const DT cpm = pathcost / (DT)mdist;
            if (cpm < best) {
              best = cpm;
              bestD = pathcost;
              kb = k;
              bestPp = Pp;
            }
          }
        }
      }
      const int bpr = r - ss.dr[kb], bpc = c - ss.dc[kb];
      ND[j] = bestD;
      NP[j] = (bpr == 0) ? bpc : ((bpc == 0) ? -bpr : bestPp);
      code[j] = kb;
    }

#pragma unroll
    for (int j = 0; j < CPL; ++j) {
      D2[j] = D1[j];
      P2[j] = P1[j];
      D1[j] = ND[j];
      P1[j] = NP[j];
    }

    if constexpr (FULL) {
      const int64_t base = (int64_t)r * W;
#pragma unroll
      for (int j = 0; j < CPL; ++j) {
        const int c = c0 + j;
        if (c < W) {
          Df[base + c] = (double)ND[j];
          Pf[base + c] = NP[j];
          Bf[base + c] = (int8_t)code[j];
        }
      }
    } else {
      const int sh = 2 * (r & 15);
#pragma unroll
      for (int j = 0; j < CPL; ++j) bacc[j] |= (uint32_t)code[j] << sh;
      if ((r & 15) == 15 || r == R - 1) {
        uint32_t *dst = Bp + td.b_off + (int64_t)(r >> 4) * W;
#pragma unroll
        for (int j = 0; j < CPL; ++j) {
          const int c = c0 + j;
          if (c < W) dst[c] = bacc[j];
          bacc[j] = 0u;
        }
      }
#pragma unroll
      for (int j = 0; j < CPL; ++j) {
        const int c = c0 + j;
        if (c == W - 1) {
          eD[td.e_off + Lmax + r] = (double)ND[j];
          eP[td.e_off + Lmax + r] = NP[j];
        }
        if (r == R - 1 && c < W) {
          eD[td.e_off + c] = (double)ND[j];
          eP[td.e_off + c] = NP[j];
        }
      }
    }

    if (lane == 63) {
      const int s0 = r % 3;
      shD[s0][wave][0] = D1[CPL - 2];
      shD[s0][wave][1] = D1[CPL - 1];
      shP[s0][wave][0] = P1[CPL - 2];
      shP[s0][wave][1] = P1[CPL - 1];
    }
    __syncthreads();
  }
}

template <int CPL, typename CT, typename DT>
__global__ __launch_bounds__(1024) void k_stage1_packed(const CT *__restrict__ Cb,
                                                        const TileDesc *__restrict__ tiles,
                                                        StepSet ss, int64_t Lmax,
                                                        uint32_t *__restrict__ Bp,
                                                        double *__restrict__ eD,
                                                        int32_t *__restrict__ eP) {
  const TileDesc td = tiles[blockIdx.x];
  tile_dp<CPL, CT, DT, false>(td, Cb, ss, Lmax, Bp, eD, eP, nullptr, nullptr, nullptr);
}

template <int CPL>
__global__ __launch_bounds__(1024) void k_tile_full(const double *__restrict__ Cb, TileDesc td,
                                                    StepSet ss, double *__restrict__ Df,
                                                    int32_t *__restrict__ Pf,
                                                    int8_t *__restrict__ Bf) {
  tile_dp<CPL, double, double, true>(td, Cb, ss, 0, nullptr, nullptr, nullptr, Df, Pf, Bf);
}

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

// Columns per lane and waves per workgroup for tiles up to Wmax columns.
static int pick_cpl(int64_t Wmax) {
  if (Wmax <= 2048) return 2;
  if (Wmax <= 4096) return 4;
  if (Wmax <= 8192) return 8;
  return 0;
}

template <typename CT, typename DT>
static int launch_stage1(const void *C, const int64_t *tiles, int64_t ntiles, int64_t Lmax,
                         const StepSet &ss, uint32_t *Bp, double *eD, int32_t *eP,
                         hipStream_t st) {
  const int cpl = pick_cpl(Lmax);
  PF_REQUIRE(cpl > 0, "tile length above 8192 is not supported");
  const int nthreads = (int)(((Lmax + cpl - 1) / cpl + 63) / 64) * 64;
  const dim3 grid((unsigned)ntiles), block((unsigned)nthreads);
  const CT *Cp = static_cast<const CT *>(C);
  const TileDesc *td = reinterpret_cast<const TileDesc *>(tiles);
  if (cpl == 2)
    hipLaunchKernelGGL((k_stage1_packed<2, CT, DT>), grid, block, 0, st, Cp, td, ss, Lmax, Bp,
                       eD, eP);
  else if (cpl == 4)
    hipLaunchKernelGGL((k_stage1_packed<4, CT, DT>), grid, block, 0, st, Cp, td, ss, Lmax, Bp,
                       eD, eP);
  else
    hipLaunchKernelGGL((k_stage1_packed<8, CT, DT>), grid, block, 0, st, Cp, td, ss, Lmax, Bp,
                       eD, eP);
  PF_CHECK_HIP(hipGetLastError());
  return PF_OK;
}

}  // namespace pf

using namespace pf;

extern "C" int pf_stage1(int dtype, const void *d_C, const int64_t *d_tiles, int64_t ntiles,
                         int64_t Lmax, const int32_t *steps, const double *weights, int K,
                         uint32_t *d_Bpacked, double *d_edge_D, int32_t *d_edge_P,
                         void *stream) {
  StepSet ss;
  int rc = make_stepset(steps, weights, K, &ss);
  if (rc) return rc;
  PF_REQUIRE(ntiles >= 0 && Lmax >= 2, "pf_stage1: bad tile count or tile length");
  if (ntiles == 0) return PF_OK;
  PF_REQUIRE(d_C && d_tiles && d_Bpacked && d_edge_D && d_edge_P, "pf_stage1: null buffer");
  hipStream_t st = as_stream(stream);
  switch (dtype) {
    case PF_F64:
      return launch_stage1<double, double>(d_C, d_tiles, ntiles, Lmax, ss, d_Bpacked, d_edge_D,
                                           d_edge_P, st);
    case PF_F32:
      return launch_stage1<float, float>(d_C, d_tiles, ntiles, Lmax, ss, d_Bpacked, d_edge_D,
                                         d_edge_P, st);
    case PF_BF16:
      return launch_stage1<__hip_bfloat16, float>(d_C, d_tiles, ntiles, Lmax, ss, d_Bpacked,
                                                  d_edge_D, d_edge_P, st);
    default:
      PF_REQUIRE(false, "pf_stage1: unknown dtype");
  }
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
  PF_REQUIRE(d_C && d_D && d_P && d_B, "pf_tile_full: null buffer");
  const int cpl = pick_cpl(W);
  PF_REQUIRE(cpl > 0, "tile width above 8192 is not supported");
  TileDesc td{r0 * ldc + c0, ldc, R, W, 0, 0, 0, 0};
  const int nthreads = (int)(((W + cpl - 1) / cpl + 63) / 64) * 64;
  hipStream_t st = as_stream(stream);
  if (cpl == 2)
    hipLaunchKernelGGL(k_tile_full<2>, dim3(1), dim3(nthreads), 0, st, d_C, td, ss, d_D, d_P, d_B);
  else if (cpl == 4)
    hipLaunchKernelGGL(k_tile_full<4>, dim3(1), dim3(nthreads), 0, st, d_C, td, ss, d_D, d_P, d_B);
  else
    hipLaunchKernelGGL(k_tile_full<8>, dim3(1), dim3(nthreads), 0, st, d_C, td, ss, d_D, d_P, d_B);
  PF_CHECK_HIP(hipGetLastError());
  return PF_OK;
}

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
