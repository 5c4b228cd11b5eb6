/*
 * CPU oracle for the ParFlex hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference algorithm, compiled with
 * -O2 -ffp-contract=off so that every multiply and add is rounded on its own,
 * as in the reference's Numba code (fastmath off for the DP kernels).  Used by
 * tests/ as the parity checker and by bench.py as the timed CPU baseline.
 * The product path (parflexdtw_amd) never loads it.
 *
 * Each function names the reference lines it follows.
 */
#include "parflex_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

static inline int64_t i64abs(int64_t v) { return v < 0 ? -v : v; }

/* FlexDTW.py:1320-1364 (flexdtw, DP part). */
void or_flexdtw_dp(const double *C, int64_t ldc, int64_t R, int64_t W,
                   const int64_t *steps, const double *weights, int K,
                   double *D, int32_t *P, int8_t *B) {
  for (int64_t r = 0; r < R; ++r)
    for (int64_t c = 0; c < W; ++c) {
      D[r * W + c] = 0.0;
      P[r * W + c] = 0;
      B[r * W + c] = 0;
    }
  for (int64_t c = 0; c < W; ++c) D[c] = C[c];
  for (int64_t r = 0; r < R; ++r) D[r * W] = C[r * ldc];

  for (int64_t row = 1; row < R; ++row) {
    for (int64_t col = 1; col < W; ++col) {
      const double cval = C[row * ldc + col];
      double mincost = INFINITY;
      int minidx = -1;
      int64_t bpr = -1, bpc = -1;
      for (int k = 0; k < K; ++k) {
        const int64_t pr = row - steps[2 * k];
        const int64_t pc = col - steps[2 * k + 1];
        if (pr >= 0 && pc >= 0) {
          const double prod = cval * weights[k];
          const double pathcost = D[pr * W + pc] + prod;
          const int64_t p = P[pr * W + pc];
          const int64_t mdist = (p >= 0) ? row + (col - p) : (row + p) + col;
          const double cpm = pathcost / (double)mdist;
          if (cpm < mincost) {
            mincost = cpm;
            minidx = k;
            bpr = pr;
            bpc = pc;
          }
        }
      }
      if (minidx < 0) {
        /* Unreachable for step sets containing a step with dr == 1 and
         * dc <= 1 and finite C; the reference would index weights[-1]. */
        D[row * W + col] = NAN;
        B[row * W + col] = -1;
        P[row * W + col] = 0;
        continue;
      }
      const double prod = cval * weights[minidx];
      D[row * W + col] = D[bpr * W + bpc] + prod;
      B[row * W + col] = (int8_t)minidx;
      if (bpr == 0)
        P[row * W + col] = (int32_t)bpc;
      else if (bpc == 0)
        P[row * W + col] = (int32_t)(-bpr);
      else
        P[row * W + col] = P[bpr * W + bpc];
    }
  }
}

/* FlexDTW.py:1098-1148 (find_best_endpoint). Candidate order: last row
 * ascending in column from `buffer`, then last column descending in row from
 * R-2 down to `buffer`. */
int64_t or_find_best_endpoint(const double *D, const int32_t *P, int64_t R,
                              int64_t W, int64_t buffer, double *avg,
                              double *best_cost, int64_t *best_r,
                              int64_t *best_c) {
  double best = INFINITY;
  int64_t br = -1, bc = -1, n = 0;
  const int64_t n_row = (W > buffer) ? W - buffer : 0;
  const int64_t n_col = (R - 1 > buffer) ? R - 1 - buffer : 0;
  for (int64_t i = 0; i < n_row + n_col; ++i) {
    int64_t r, c;
    if (i < n_row) {
      r = R - 1;
      c = buffer + i;
    } else {
      r = R - 2 - (i - n_row);
      c = W - 1;
    }
    const int64_t p = P[r * W + c];
    int64_t rs, cs;
    if (p >= 0) {
      rs = 0;
      cs = p;
    } else {
      rs = -p;
      cs = 0;
    }
    const int64_t mdist = (r - rs) + (c - cs);
    const double a = D[r * W + c] / (double)mdist;
    if (avg) avg[n] = a;
    ++n;
    if (a < best) {
      best = a;
      br = r;
      bc = c;
    }
  }
  *best_cost = best;
  *best_r = br;
  *best_c = bc;
  return n;
}

/* FlexDTW.py:1155-1180 (backtrace_flexdtw) followed by path.reverse()
 * (FlexDTW.py:1368). */
int64_t or_backtrace(const int8_t *B, int64_t R, int64_t W,
                     const int64_t *steps, int64_t r, int64_t c,
                     int64_t *out_pts) {
  (void)R;
  int64_t n = 0;
  out_pts[2 * n] = r;
  out_pts[2 * n + 1] = c;
  ++n;
  while (r != 0 && c != 0) {
    const int k = B[r * W + c];
    r -= steps[2 * k];
    c -= steps[2 * k + 1];
    out_pts[2 * n] = r;
    out_pts[2 * n + 1] = c;
    ++n;
  }
  for (int64_t a = 0, b = n - 1; a < b; ++a, --b) {
    int64_t t0 = out_pts[2 * a], t1 = out_pts[2 * a + 1];
    out_pts[2 * a] = out_pts[2 * b];
    out_pts[2 * a + 1] = out_pts[2 * b + 1];
    out_pts[2 * b] = t0;
    out_pts[2 * b + 1] = t1;
  }
  return n;
}

/* Parflex.py:388-391: hop = L-1, n = ceil((Lx-1)/hop). */
void or_tile_grid(int64_t L1, int64_t L2, int64_t L, int64_t *n1, int64_t *n2) {
  const int64_t hop = L - 1;
  *n1 = (L1 - 1 + hop - 1) / hop;
  *n2 = (L2 - 1 + hop - 1) / hop;
}

typedef struct {
  int64_t s1, e1, s2, e2, rows, cols;
} tile_geo;

/* Parflex.py:424-427 tile bounds. */
static tile_geo geo(int64_t i, int64_t j, int64_t L1, int64_t L2, int64_t L) {
  tile_geo g;
  const int64_t hop = L - 1;
  g.s1 = i * hop;
  g.s2 = j * hop;
  g.e1 = g.s1 + L < L1 ? g.s1 + L : L1;
  g.e2 = g.s2 + L < L2 ? g.s2 + L : L2;
  g.rows = g.e1 - g.s1;
  g.cols = g.e2 - g.s2;
  return g;
}

typedef struct {
  const double *C;
  int64_t L1, L2, L, n1, n2;
  const int64_t *steps;
  const double *weights;
  int K;
  double *edge_D;
  int32_t *edge_P;
  int8_t *B_all;
  double *best;
  int64_t next;
  pthread_mutex_t mu;
  int fail;
} stage1_job;

static void *stage1_worker(void *arg) {
  stage1_job *job = (stage1_job *)arg;
  const int64_t L = job->L;
  double *D = (double *)malloc(sizeof(double) * L * L);
  int32_t *P = (int32_t *)malloc(sizeof(int32_t) * L * L);
  int8_t *Bs = (int8_t *)malloc((size_t)L * L);
  if (!D || !P || !Bs) {
    pthread_mutex_lock(&job->mu);
    job->fail = 1;
    pthread_mutex_unlock(&job->mu);
    free(D);
    free(P);
    free(Bs);
    return NULL;
  }
  for (;;) {
    pthread_mutex_lock(&job->mu);
    const int64_t t = job->next++;
    pthread_mutex_unlock(&job->mu);
    if (t >= job->n1 * job->n2) break;
    const int64_t i = t / job->n2, j = t % job->n2;
    const tile_geo g = geo(i, j, job->L1, job->L2, L);
    int8_t *B = job->B_all ? job->B_all + (size_t)t * L * L : Bs;
    or_flexdtw_dp(job->C + g.s1 * job->L2 + g.s2, job->L2, g.rows, g.cols,
                  job->steps, job->weights, job->K, D, P, B);
    double *eD = job->edge_D + (size_t)t * 2 * L;
    int32_t *eP = job->edge_P + (size_t)t * 2 * L;
    for (int64_t p = 0; p < 2 * L; ++p) {
      eD[p] = INFINITY;
      eP[p] = 0;
    }
    for (int64_t c = 0; c < g.cols; ++c) {
      eD[c] = D[(g.rows - 1) * g.cols + c];
      eP[c] = P[(g.rows - 1) * g.cols + c];
    }
    for (int64_t r = 0; r < g.rows; ++r) {
      eD[L + r] = D[r * g.cols + g.cols - 1];
      eP[L + r] = P[r * g.cols + g.cols - 1];
    }
    if (job->best) {
      double bc;
      int64_t br, bcc;
      or_find_best_endpoint(D, P, g.rows, g.cols, 1, NULL, &bc, &br, &bcc);
      job->best[t] = bc;
    }
  }
  free(D);
  free(P);
  free(Bs);
  return NULL;
}

/* Stage 1 (Parflex.py:376-469) on the cpu_parflex.py:380-428 schedule:
 * a pool of worker threads, one whole tile per task. */
int or_stage1(const double *C, int64_t L1, int64_t L2, int64_t L,
              const int64_t *steps, const double *weights, int K,
              int nthreads, double *edge_D, int32_t *edge_P, int8_t *B_all,
              double *best) {
  stage1_job job;
  memset(&job, 0, sizeof(job));
  job.C = C;
  job.L1 = L1;
  job.L2 = L2;
  job.L = L;
  or_tile_grid(L1, L2, L, &job.n1, &job.n2);
  job.steps = steps;
  job.weights = weights;
  job.K = K;
  job.edge_D = edge_D;
  job.edge_P = edge_P;
  job.B_all = B_all;
  job.best = best;
  pthread_mutex_init(&job.mu, NULL);
  const int64_t ntiles = job.n1 * job.n2;
  int nt = nthreads < 1 ? 1 : nthreads;
  if (nt > ntiles) nt = (int)(ntiles > 0 ? ntiles : 1);
  if (nt == 1) {
    stage1_worker(&job);
  } else {
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * nt);
    for (int k = 0; k < nt; ++k) pthread_create(&th[k], NULL, stage1_worker, &job);
    for (int k = 0; k < nt; ++k) pthread_join(th[k], NULL);
    free(th);
  }
  pthread_mutex_destroy(&job.mu);
  return job.fail ? -1 : 0;
}

/* Start decoding of Parflex.py:543-547: s>0 -> (0,s); s<0 -> (-s,0); 0 -> (0,0). */
static inline void decode_start(int64_t s, int64_t *sr, int64_t *sc) {
  if (s > 0) {
    *sr = 0;
    *sc = s;
  } else if (s < 0) {
    *sr = -s;
    *sc = 0;
  } else {
    *sr = 0;
    *sc = 0;
  }
}

#define IDX4(i, j, e, p) ((((i) * n2 + (j)) * 2 + (e)) * L + (p))

/* Parflex.py:976-1012 (_nb_backtrace_within_chunk). Writes global points
 * (row, col) from the end cell back to the start cell; returns the count. */
static int64_t backtrace_within_chunk(const int8_t *B, int64_t nrows,
                                      int64_t ncols, const int64_t *steps,
                                      int K, int64_t start_r, int64_t start_c,
                                      int64_t end_r, int64_t end_c, int64_t gro,
                                      int64_t gco, int64_t *out) {
  const int64_t max_iters = nrows + ncols;
  int64_t k = 0, r = end_r, c = end_c;
  int64_t it;
  for (it = 0; it < max_iters; ++it) {
    out[2 * k] = r + gro;
    out[2 * k + 1] = c + gco;
    ++k;
    if (r == start_r && c == start_c) break;
    const int64_t step_idx = B[r * ncols + c];
    if (step_idx < 0 || step_idx >= K) {
      if (r != start_r || c != start_c) {
        out[2 * k] = start_r + gro;
        out[2 * k + 1] = start_c + gco;
        ++k;
      }
      break;
    }
    const int64_t pr = r - steps[2 * step_idx], pc = c - steps[2 * step_idx + 1];
    if (pr < 0 || pc < 0 || pr >= nrows || pc >= ncols) {
      if (r != start_r || c != start_c) {
        out[2 * k] = start_r + gro;
        out[2 * k + 1] = start_c + gco;
        ++k;
      }
      break;
    }
    r = pr;
    c = pc;
  }
  if (it == max_iters) {
    if (r != start_r || c != start_c) {
      out[2 * k] = start_r + gro;
      out[2 * k + 1] = start_c + gco;
      ++k;
    }
  }
  return k;
}

/* Parflex.py:1032-1103 (backtrace_and_stitch), returns the number of points
 * written to path (already in forward order). */
int64_t or_backtrace_and_stitch(int64_t L1, int64_t L2, int64_t L,
                                const int64_t *steps, int K,
                                const int32_t *edge_P, const int8_t *B_all,
                                int64_t ci, int64_t cj, int64_t et, int64_t idx,
                                int64_t *path, int64_t path_cap) {
  int64_t n1, n2;
  or_tile_grid(L1, L2, L, &n1, &n2);
  int64_t n = 0;
  int64_t *seg = (int64_t *)malloc(sizeof(int64_t) * 2 * (2 * L + 8));
  int64_t cur_i = ci, cur_j = cj, cur_e = et, cur_idx = idx;
  /* The stitched chain moves to a tile with a smaller i+j at every step, so a
   * chunk key can never repeat; the visited set of the reference is kept
   * implicit. */
  for (int it = 0; it < 101; ++it) {
    if (cur_i < 0 || cur_j < 0 || cur_i >= n1 || cur_j >= n2) break;
    const tile_geo g = geo(cur_i, cur_j, L1, L2, L);
    int64_t end_r, end_c;
    if (cur_e == 0) {
      end_r = g.rows - 1;
      end_c = cur_idx;
    } else {
      end_r = cur_idx;
      end_c = g.cols - 1;
    }
    /* S at the end cell: the end cell is on the tile's last row or column. */
    int64_t S_val;
    if (end_r == g.rows - 1)
      S_val = edge_P[IDX4(cur_i, cur_j, 0, end_c)];
    else
      S_val = edge_P[IDX4(cur_i, cur_j, 1, end_r)];
    int64_t start_r, start_c;
    if (S_val >= 0) {
      start_r = 0;
      start_c = S_val;
    } else {
      start_r = -S_val;
      start_c = 0;
    }
    const int8_t *B = B_all + (size_t)(cur_i * n2 + cur_j) * L * L;
    const int64_t k = backtrace_within_chunk(B, g.rows, g.cols, steps, K,
                                             start_r, start_c, end_r, end_c,
                                             g.s1, g.s2, seg);
    for (int64_t q = 0; q < k; ++q) {
      const int64_t pr = seg[2 * q], pc = seg[2 * q + 1];
      if (n == 0 || path[2 * (n - 1)] != pr || path[2 * (n - 1) + 1] != pc) {
        if (n < path_cap) {
          path[2 * n] = pr;
          path[2 * n + 1] = pc;
        }
        ++n;
      }
    }
    const int64_t g_start_row = g.s1 + start_r, g_start_col = g.s2 + start_c;
    if (g_start_row == 0 || g_start_col == 0) break;
    const int corner = (start_r == 0 && start_c == 0);
    int64_t pi, pj, pe;
    if (corner) {
      pi = cur_i - 1;
      pj = cur_j - 1;
      pe = 0;
    } else if (S_val >= 0) {
      pi = cur_i - 1;
      pj = cur_j;
      pe = 0;
    } else {
      pi = cur_i;
      pj = cur_j - 1;
      pe = 1;
    }
    if (pi < 0 || pj < 0) break;
    if (pi >= n1 || pj >= n2) break;
    const tile_geo pg = geo(pi, pj, L1, L2, L);
    int64_t prev_idx, max_prev;
    if (corner) {
      prev_idx = pg.cols - 1;
      max_prev = pg.cols;
    } else {
      const int64_t plr = g_start_row - pg.s1, plc = g_start_col - pg.s2;
      prev_idx = pe == 0 ? plc : plr;
      max_prev = pe == 0 ? pg.cols : pg.rows;
    }
    if (prev_idx < 0 || prev_idx >= max_prev) break;
    cur_i = pi;
    cur_j = pj;
    cur_e = pe;
    cur_idx = prev_idx;
  }
  free(seg);
  /* path[::-1] */
  const int64_t m = n < path_cap ? n : path_cap;
  for (int64_t a = 0, b = m - 1; a < b; ++a, --b) {
    int64_t t0 = path[2 * a], t1 = path[2 * a + 1];
    path[2 * a] = path[2 * b];
    path[2 * a + 1] = path[2 * b + 1];
    path[2 * b] = t0;
    path[2 * b + 1] = t1;
  }
  return n;
}

/* sync_tile_overlap_edges (Parflex.py:919-969): three passes in order, in
 * place; edge_lens [n1][n2][2] (cols, rows). */
void or_sync_overlap(double *D_arr, double *L_arr, const int64_t *edge_lens, int64_t n1,
                     int64_t n2, int64_t L) {
  for (int64_t i = 0; i < n1; ++i)
    for (int64_t j = 0; j + 1 < n2; ++j) {
      const int64_t el = edge_lens[2 * (i * n2 + j)];
      if (el == 0) continue;
      const double dl = D_arr[IDX4(i, j, 0, el - 1)], dr = D_arr[IDX4(i, j + 1, 0, 0)];
      if (isfinite(dl) && !isfinite(dr)) {
        D_arr[IDX4(i, j + 1, 0, 0)] = dl;
        L_arr[IDX4(i, j + 1, 0, 0)] = L_arr[IDX4(i, j, 0, el - 1)];
      } else if (!isfinite(dl) && isfinite(dr)) {
        D_arr[IDX4(i, j, 0, el - 1)] = dr;
        L_arr[IDX4(i, j, 0, el - 1)] = L_arr[IDX4(i, j + 1, 0, 0)];
      }
    }
  for (int64_t i = 0; i + 1 < n1; ++i)
    for (int64_t j = 0; j < n2; ++j) {
      const int64_t el = edge_lens[2 * (i * n2 + j) + 1];
      if (el == 0) continue;
      const double db = D_arr[IDX4(i, j, 1, el - 1)], dt = D_arr[IDX4(i + 1, j, 1, 0)];
      if (isfinite(db) && !isfinite(dt)) {
        D_arr[IDX4(i + 1, j, 1, 0)] = db;
        L_arr[IDX4(i + 1, j, 1, 0)] = L_arr[IDX4(i, j, 1, el - 1)];
      } else if (!isfinite(db) && isfinite(dt)) {
        D_arr[IDX4(i, j, 1, el - 1)] = dt;
        L_arr[IDX4(i, j, 1, el - 1)] = L_arr[IDX4(i + 1, j, 1, 0)];
      }
    }
  for (int64_t i = 0; i < n1; ++i)
    for (int64_t j = 0; j < n2; ++j) {
      const int64_t et = edge_lens[2 * (i * n2 + j)], er = edge_lens[2 * (i * n2 + j) + 1];
      if (et == 0 || er == 0) continue;
      const double dtc = D_arr[IDX4(i, j, 0, et - 1)], drc = D_arr[IDX4(i, j, 1, er - 1)];
      if (isfinite(dtc) && !isfinite(drc)) {
        D_arr[IDX4(i, j, 1, er - 1)] = dtc;
        L_arr[IDX4(i, j, 1, er - 1)] = L_arr[IDX4(i, j, 0, et - 1)];
      } else if (!isfinite(dtc) && isfinite(drc)) {
        D_arr[IDX4(i, j, 0, et - 1)] = drc;
        L_arr[IDX4(i, j, 0, et - 1)] = L_arr[IDX4(i, j, 1, er - 1)];
      } else if (isfinite(dtc) && isfinite(drc) && fabs(dtc - drc) > 1e-10) {
        D_arr[IDX4(i, j, 1, er - 1)] = dtc;
        L_arr[IDX4(i, j, 1, er - 1)] = L_arr[IDX4(i, j, 0, et - 1)];
      }
    }

}

static int stage2_impl(const double *C, int64_t L1, int64_t L2, int64_t L,
                       const int64_t *steps, int K, double beta,
                       const double *edge_D, const int32_t *edge_P, const int8_t *B_all,
                       double *D_arr, double *L_arr, int64_t *edge_lens,
                       double *top_D, double *top_L, double *right_D, double *right_L,
                       double *seg_norm, double *seg_D, double *seg_L, int64_t *seg_i64,
                       int64_t max_segs, int64_t *best_out, double *best_norm_out,
                       int64_t *path, int64_t path_cap, int64_t *path_len, int presync) {
  int64_t n1, n2;
  or_tile_grid(L1, L2, L, &n1, &n2);
  const size_t n4 = (size_t)n1 * n2 * 2 * L;
  int64_t *esr = (int64_t *)calloc(n4, sizeof(int64_t));
  int64_t *esc = (int64_t *)calloc(n4, sizeof(int64_t));
  double *olap = (double *)calloc(n4, sizeof(double));
  unsigned char *valid = (unsigned char *)calloc(n4, 1);
  unsigned char *bot = (unsigned char *)calloc((size_t)n1 * n2 * (L + 1), 1);
  unsigned char *left = (unsigned char *)calloc((size_t)n1 * n2 * (L + 1), 1);
  int64_t *rows = (int64_t *)malloc(sizeof(int64_t) * n1 * n2);
  int64_t *cols = (int64_t *)malloc(sizeof(int64_t) * n1 * n2);
  int64_t *bnd = (int64_t *)malloc(sizeof(int64_t) * n1 * n2 * 4);
  if (!esr || !esc || !olap || !valid || !bot || !left || !rows || !cols || !bnd) return -1;

  /* _build_edge_data (Parflex.py:509-564) and _edge_starts_from_S
   * (Parflex.py:347-373). */
  for (int64_t i = 0; i < n1; ++i)
    for (int64_t j = 0; j < n2; ++j) {
      const tile_geo g = geo(i, j, L1, L2, L);
      const int64_t t = i * n2 + j;
      rows[t] = g.rows;
      cols[t] = g.cols;
      bnd[4 * t] = g.s1;
      bnd[4 * t + 1] = g.e1;
      bnd[4 * t + 2] = g.s2;
      bnd[4 * t + 3] = g.e2;
      edge_lens[2 * t] = g.cols;
      edge_lens[2 * t + 1] = g.rows;
      for (int e = 0; e < 2; ++e) {
        const int64_t elen = e == 0 ? g.cols : g.rows;
        for (int64_t p = 0; p < L; ++p) {
          D_arr[IDX4(i, j, e, p)] = INFINITY;
          L_arr[IDX4(i, j, e, p)] = INFINITY;
        }
        for (int64_t p = 0; p < elen; ++p) {
          const int64_t s = edge_P[IDX4(i, j, e, p)];
          int64_t sr, sc;
          decode_start(s, &sr, &sc);
          esr[IDX4(i, j, e, p)] = sr;
          esc[IDX4(i, j, e, p)] = sc;
          olap[IDX4(i, j, e, p)] = C[(g.s1 + sr) * L2 + (g.s2 + sc)];
          if (s > 0) {
            if (s < L + 1) bot[t * (L + 1) + s] = 1;
          } else {
            const int64_t a = -s;
            if (a < L + 1) left[t * (L + 1) + a] = 1;
          }
        }
      }
    }

  /* _build_valid_mask (Parflex.py:582-624). */
  for (int64_t i = 0; i < n1; ++i)
    for (int64_t j = 0; j < n2; ++j) {
      const int64_t t = i * n2 + j;
      const int64_t elen0 = cols[t], elen1 = rows[t];
      if (i == n1 - 1) {
        for (int64_t p = 0; p < elen0; ++p) valid[IDX4(i, j, 0, p)] = 1;
      } else {
        const int64_t ta = (i + 1) * n2 + j;
        for (int64_t p = 0; p < elen0 && p < L + 1; ++p)
          if (bot[ta * (L + 1) + p]) valid[IDX4(i, j, 0, p)] = 1;
        valid[IDX4(i, j, 0, 0)] = 1;
        valid[IDX4(i, j, 0, elen0 - 1)] = 1;
      }
      if (j == n2 - 1) {
        for (int64_t p = 0; p < elen1; ++p) valid[IDX4(i, j, 1, p)] = 1;
      } else {
        const int64_t tr = i * n2 + j + 1;
        for (int64_t p = 0; p < elen1 && p < L + 1; ++p)
          if (left[tr * (L + 1) + p]) valid[IDX4(i, j, 1, p)] = 1;
        valid[IDX4(i, j, 1, 0)] = 1;
        valid[IDX4(i, j, 1, elen1 - 1)] = 1;
      }
    }

  /* _nb_initialize_chunks_sparse (Parflex.py:631-760). */
  {
    const int64_t nr = rows[0], nc = cols[0];
    for (int e = 0; e < 2; ++e) {
      const int64_t elen = e == 0 ? nc : nr;
      for (int64_t p = 0; p < elen; ++p) {
        if (!valid[IDX4(0, 0, e, p)]) continue;
        const int64_t sr = esr[IDX4(0, 0, e, p)], sc = esc[IDX4(0, 0, e, p)];
        const int64_t lr = e == 0 ? nr - 1 : p, lc = e == 0 ? p : nc - 1;
        if (sr == 0 || sc == 0) {
          D_arr[IDX4(0, 0, e, p)] = edge_D[IDX4(0, 0, e, p)];
          L_arr[IDX4(0, 0, e, p)] = (double)(i64abs(lr - sr) + i64abs(lc - sc));
        }
      }
    }
  }
  for (int64_t j = 1; j < n2; ++j) {
    const int64_t t = j, tp = j - 1;
    const int64_t nr = rows[t], nc = cols[t];
    const int64_t pe = cols[tp];
    D_arr[IDX4(0, j, 0, 0)] = D_arr[IDX4(0, j - 1, 0, pe - 1)];
    L_arr[IDX4(0, j, 0, 0)] = L_arr[IDX4(0, j - 1, 0, pe - 1)];
    for (int e = 0; e < 2; ++e) {
      const int64_t elen = e == 0 ? nc : nr;
      for (int64_t p = 1; p < elen; ++p) {
        if (!valid[IDX4(0, j, e, p)]) continue;
        const int64_t sr = esr[IDX4(0, j, e, p)], sc = esc[IDX4(0, j, e, p)];
        const int64_t lr = e == 0 ? nr - 1 : p, lc = e == 0 ? p : nc - 1;
        if (sr == 0) {
          D_arr[IDX4(0, j, e, p)] = edge_D[IDX4(0, j, e, p)];
          L_arr[IDX4(0, j, e, p)] = (double)(i64abs(lr - sr) + i64abs(lc - sc));
        } else if (sc == 0) {
          const int64_t gsr = bnd[4 * t] + sr, gsc = bnd[4 * t + 2];
          const int64_t plr = gsr - bnd[4 * tp], plc = gsc - bnd[4 * tp + 2];
          const int64_t pnr = rows[tp], pnc = cols[tp];
          int64_t pedge, ppos, pelen;
          if (plr == pnr - 1) {
            pedge = 0;
            ppos = plc;
            pelen = pnc;
          } else if (plc == pnc - 1) {
            pedge = 1;
            ppos = plr;
            pelen = pnr;
          } else
            continue;
          if (ppos < 0 || ppos >= pelen) continue;
          const double prev = D_arr[IDX4(0, j - 1, pedge, ppos)];
          if (!isfinite(prev)) continue;
          const double ov = olap[IDX4(0, j, e, p)];
          D_arr[IDX4(0, j, e, p)] = (edge_D[IDX4(0, j, e, p)] + prev) - ov;
          L_arr[IDX4(0, j, e, p)] = L_arr[IDX4(0, j - 1, pedge, ppos)] +
                                    (double)(i64abs(lr - sr) + i64abs(lc - sc));
        }
      }
    }
  }
  for (int64_t i = 1; i < n1; ++i) {
    const int64_t t = i * n2, tp = (i - 1) * n2;
    const int64_t nr = rows[t], nc = cols[t];
    const int64_t pe = rows[tp];
    D_arr[IDX4(i, 0, 1, 0)] = D_arr[IDX4(i - 1, 0, 1, pe - 1)];
    L_arr[IDX4(i, 0, 1, 0)] = L_arr[IDX4(i - 1, 0, 1, pe - 1)];
    for (int e = 0; e < 2; ++e) {
      const int64_t elen = e == 0 ? nc : nr;
      for (int64_t p = 1; p < elen; ++p) {
        if (!valid[IDX4(i, 0, e, p)]) continue;
        const int64_t sr = esr[IDX4(i, 0, e, p)], sc = esc[IDX4(i, 0, e, p)];
        const int64_t lr = e == 0 ? nr - 1 : p, lc = e == 0 ? p : nc - 1;
        if (sc == 0) {
          D_arr[IDX4(i, 0, e, p)] = edge_D[IDX4(i, 0, e, p)];
          L_arr[IDX4(i, 0, e, p)] = (double)(i64abs(lr - sr) + i64abs(lc - sc));
        } else if (sr == 0) {
          const int64_t gsr = bnd[4 * t], gsc = bnd[4 * t + 2] + sc;
          const int64_t plr = gsr - bnd[4 * tp], plc = gsc - bnd[4 * tp + 2];
          const int64_t pnr = rows[tp], pnc = cols[tp];
          int64_t pedge, ppos, pelen;
          if (plr == pnr - 1) {
            pedge = 0;
            ppos = plc;
            pelen = pnc;
          } else if (plc == pnc - 1) {
            pedge = 1;
            ppos = plr;
            pelen = pnr;
          } else
            continue;
          if (ppos < 0 || ppos >= pelen) continue;
          const double prev = D_arr[IDX4(i - 1, 0, pedge, ppos)];
          if (!isfinite(prev)) continue;
          const double ov = olap[IDX4(i, 0, e, p)];
          D_arr[IDX4(i, 0, e, p)] = (edge_D[IDX4(i, 0, e, p)] + prev) - ov;
          L_arr[IDX4(i, 0, e, p)] = L_arr[IDX4(i - 1, 0, pedge, ppos)] +
                                    (double)(i64abs(lr - sr) + i64abs(lc - sc));
        }
      }
    }
  }

  /* _nb_dp_fill_chunks_sparse (Parflex.py:763-855). */
  for (int64_t i = 1; i < n1; ++i)
    for (int64_t j = 1; j < n2; ++j) {
      const int64_t t = i * n2 + j;
      const int64_t nr = rows[t], nc = cols[t];
      for (int e = 0; e < 2; ++e) {
        const int64_t elen = e == 0 ? nc : nr;
        for (int64_t p = 0; p < elen; ++p) {
          if (!valid[IDX4(i, j, e, p)]) continue;
          if (p == 0) {
            if (e == 0) {
              const int64_t pel = cols[i * n2 + j - 1];
              const double v = D_arr[IDX4(i, j - 1, 0, pel - 1)];
              const double l = L_arr[IDX4(i, j - 1, 0, pel - 1)];
              if (isfinite(v)) {
                D_arr[IDX4(i, j, 0, 0)] = v;
                L_arr[IDX4(i, j, 0, 0)] = l;
              }
            } else {
              const int64_t pel = rows[(i - 1) * n2 + j];
              const double v = D_arr[IDX4(i - 1, j, 1, pel - 1)];
              const double l = L_arr[IDX4(i - 1, j, 1, pel - 1)];
              if (isfinite(v)) {
                D_arr[IDX4(i, j, 1, 0)] = v;
                L_arr[IDX4(i, j, 1, 0)] = l;
              }
            }
            continue;
          }
          const int64_t lr = e == 0 ? nr - 1 : p, lc = e == 0 ? p : nc - 1;
          const int64_t sr = esr[IDX4(i, j, e, p)], sc = esc[IDX4(i, j, e, p)];
          int64_t pi, pj;
          if (sr == 0) {
            pi = i - 1;
            pj = j;
          } else if (sc == 0) {
            pi = i;
            pj = j - 1;
          } else {
            D_arr[IDX4(i, j, e, p)] = edge_D[IDX4(i, j, e, p)];
            L_arr[IDX4(i, j, e, p)] = (double)(i64abs(lr - sr) + i64abs(lc - sc));
            continue;
          }
          if (pi < 0 || pj < 0) continue;
          const int64_t tp = pi * n2 + pj;
          const int64_t gsr = bnd[4 * t] + sr, gsc = bnd[4 * t + 2] + sc;
          const int64_t plr = gsr - bnd[4 * tp], plc = gsc - bnd[4 * tp + 2];
          const int64_t pnr = rows[tp], pnc = cols[tp];
          int64_t pedge, ppos, pelen;
          if (plr == pnr - 1) {
            pedge = 0;
            ppos = plc;
            pelen = pnc;
          } else if (plc == pnc - 1) {
            pedge = 1;
            ppos = plr;
            pelen = pnr;
          } else
            continue;
          if (ppos < 0 || ppos >= pelen) continue;
          const double prev = D_arr[IDX4(pi, pj, pedge, ppos)];
          const double plen = L_arr[IDX4(pi, pj, pedge, ppos)];
          if (!isfinite(prev)) continue;
          const double ov = olap[IDX4(i, j, e, p)];
          D_arr[IDX4(i, j, e, p)] = (prev + edge_D[IDX4(i, j, e, p)]) - ov;
          L_arr[IDX4(i, j, e, p)] = plen + (double)(i64abs(lr - sr) + i64abs(lc - sc));
        }
      }
    }

  if (!presync) or_sync_overlap(D_arr, L_arr, edge_lens, n1, n2, L);
  if (presync) {
    free(esr);
    free(esc);
    free(olap);
    free(valid);
    free(bot);
    free(left);
    free(rows);
    free(cols);
    free(bnd);
    return 0;
  }

  /* _build_edge_lookups (Parflex.py:1110-1150). */
  int64_t *tci = (int64_t *)malloc(sizeof(int64_t) * L2 * 4);
  int64_t *rci = (int64_t *)malloc(sizeof(int64_t) * L1 * 4);
  for (int64_t g = 0; g < 4 * L2; ++g) tci[g] = -1;
  for (int64_t g = 0; g < 4 * L1; ++g) rci[g] = -1;
  {
    const int64_t grt = L1 - 1, gcr = L2 - 1;
    for (int64_t i = 0; i < n1; ++i)
      for (int64_t j = 0; j < n2; ++j) {
        const tile_geo g = geo(i, j, L1, L2, L);
        if (g.s1 <= grt && grt < g.e1) {
          const int64_t lr = grt - g.s1;
          for (int64_t gc = g.s2; gc < g.e2; ++gc) {
            if (tci[4 * gc] >= 0) continue;
            const int64_t lc = gc - g.s2;
            if (lr == g.rows - 1) {
              tci[4 * gc] = i;
              tci[4 * gc + 1] = j;
              tci[4 * gc + 2] = 0;
              tci[4 * gc + 3] = lc;
            } else if (lc == g.cols - 1) {
              tci[4 * gc] = i;
              tci[4 * gc + 1] = j;
              tci[4 * gc + 2] = 1;
              tci[4 * gc + 3] = lr;
            }
          }
        }
        if (g.s2 <= gcr && gcr < g.e2) {
          const int64_t lc = gcr - g.s2;
          for (int64_t gr = g.s1; gr < g.e1; ++gr) {
            if (rci[4 * gr] >= 0) continue;
            const int64_t lr = gr - g.s1;
            if (lc == g.cols - 1) {
              rci[4 * gr] = i;
              rci[4 * gr + 1] = j;
              rci[4 * gr + 2] = 1;
              rci[4 * gr + 3] = lr;
            } else if (lr == g.rows - 1) {
              rci[4 * gr] = i;
              rci[4 * gr + 1] = j;
              rci[4 * gr + 2] = 0;
              rci[4 * gr + 3] = lc;
            }
          }
        }
      }
  }

  /* _nb_scan_edges (Parflex.py:1153-1222) with the buffer of
   * Parflex.py:1526 and :1248. */
  const double mn = (double)(L1 < L2 ? L1 : L2), mx = (double)(L1 < L2 ? L2 : L1);
  const double buffer_global = mn * (1.0 - ((1.0 - beta) * mn) / mx);
  const int64_t buf = (int64_t)buffer_global;
  for (int64_t g = 0; g < L2; ++g) {
    top_D[g] = NAN;
    top_L[g] = 0.0;
  }
  for (int64_t g = 0; g < L1; ++g) {
    right_D[g] = NAN;
    right_L[g] = 0.0;
  }
  for (int64_t s = 0; s < 2 * max_segs; ++s) {
    seg_norm[s] = INFINITY;
    seg_D[s] = INFINITY;
    seg_L[s] = 0.0;
    for (int f = 0; f < 7; ++f) seg_i64[f * 2 * max_segs + s] = -1;
  }
  const double LARGE = 1e9;
  double best_norm = LARGE;
  for (int f = 0; f < 6; ++f) best_out[f] = -1;
  best_out[6] = 0;
  for (int side = 0; side < 2; ++side) {
    const int64_t n = side == 0 ? L2 : L1;
    const int64_t *lk = side == 0 ? tci : rci;
    for (int64_t g = 0; g < n; ++g) {
      if (buf > 0 && g < buf) continue;
      const int64_t ci = lk[4 * g], cj = lk[4 * g + 1], et = lk[4 * g + 2], idx = lk[4 * g + 3];
      if (ci < 0) continue;
      const int64_t pel = edge_lens[2 * (ci * n2 + cj) + et];
      if (idx < 0 || idx >= pel) continue;
      const double Dv = D_arr[IDX4(ci, cj, et, idx)], Lv = L_arr[IDX4(ci, cj, et, idx)];
      if (side == 0) {
        if (!isfinite(Dv) || Dv >= LARGE || Lv <= 0.0) continue;
      } else {
        if (!isfinite(Dv) || Lv <= 0.0) continue;
      }
      const double norm = Dv / Lv;
      if (side == 0) {
        top_D[g] = Dv;
        top_L[g] = Lv;
      } else {
        right_D[g] = Dv;
        right_L[g] = Lv;
      }
      if (norm < best_norm) {
        best_norm = norm;
        best_out[0] = side == 0 ? L1 - 1 : g;
        best_out[1] = side == 0 ? g : L2 - 1;
        best_out[2] = ci;
        best_out[3] = cj;
        best_out[4] = et;
        best_out[5] = idx;
        best_out[6] = 1;
      }
      const int64_t sidx = g / L;
      if (sidx < max_segs) {
        const int64_t s = side * max_segs + sidx;
        if (norm < seg_norm[s]) {
          seg_norm[s] = norm;
          seg_D[s] = Dv;
          seg_L[s] = Lv;
          seg_i64[0 * 2 * max_segs + s] = ci;
          seg_i64[1 * 2 * max_segs + s] = cj;
          seg_i64[2 * 2 * max_segs + s] = et;
          seg_i64[3 * 2 * max_segs + s] = idx;
          seg_i64[4 * 2 * max_segs + s] = side == 0 ? L1 - 1 : g;
          seg_i64[5 * 2 * max_segs + s] = side == 0 ? g : L2 - 1;
          seg_i64[6 * 2 * max_segs + s] = 1;
        }
      }
    }
  }
  *best_norm_out = best_norm;
  int rc = 0;
  *path_len = 0;
  if (best_out[6] == 0) {
    rc = 1;
  } else if (path) {
    *path_len = or_backtrace_and_stitch(L1, L2, L, steps, K, edge_P, B_all, best_out[2],
                                        best_out[3], best_out[4], best_out[5], path,
                                        path_cap);
  }
  free(tci);
  free(rci);
  free(esr);
  free(esc);
  free(olap);
  free(valid);
  free(bot);
  free(left);
  free(rows);
  free(cols);
  free(bnd);
  return rc;
}

int or_stage2(const double *C, int64_t L1, int64_t L2, int64_t L,
              const int64_t *steps, int K, double beta,
              const double *edge_D, const int32_t *edge_P, const int8_t *B_all,
              double *D_arr, double *L_arr, int64_t *edge_lens,
              double *top_D, double *top_L, double *right_D, double *right_L,
              double *seg_norm, double *seg_D, double *seg_L, int64_t *seg_i64,
              int64_t max_segs, int64_t *best_out, double *best_norm_out,
              int64_t *path, int64_t path_cap, int64_t *path_len) {
  return stage2_impl(C, L1, L2, L, steps, K, beta, edge_D, edge_P, B_all, D_arr, L_arr,
                     edge_lens, top_D, top_L, right_D, right_L, seg_norm, seg_D, seg_L, seg_i64,
                     max_segs, best_out, best_norm_out, path, path_cap, path_len, 0);
}

/* propagate_tile_edge_costs (Parflex.py:862-912): D_arr, L_arr before the
 * overlap sync, and edge_lens. */
int or_propagate(const double *C, int64_t L1, int64_t L2, int64_t L, const double *edge_D,
                 const int32_t *edge_P, double *D_arr, double *L_arr, int64_t *edge_lens) {
  return stage2_impl(C, L1, L2, L, NULL, 0, 0.0, edge_D, edge_P, NULL, D_arr, L_arr, edge_lens,
                     NULL, NULL, NULL, NULL, NULL, NULL, NULL, NULL, 0, NULL, NULL, NULL, 0, NULL,
                     1);
}

int or_parflex(const double *C, int64_t L1, int64_t L2, int64_t L,
               const int64_t *steps, const double *weights, int K, double beta,
               int nthreads, double *best_cost, int64_t *path,
               int64_t path_cap, int64_t *path_len) {
  int64_t n1, n2;
  or_tile_grid(L1, L2, L, &n1, &n2);
  const size_t n4 = (size_t)n1 * n2 * 2 * L;
  const int64_t max_segs_a = (L2 + L - 1) / L, max_segs_b = (L1 + L - 1) / L;
  int64_t max_segs = max_segs_a > max_segs_b ? max_segs_a : max_segs_b;
  if (max_segs < 1) max_segs = 1;
  double *eD = (double *)malloc(sizeof(double) * n4);
  int32_t *eP = (int32_t *)malloc(sizeof(int32_t) * n4);
  int8_t *B = (int8_t *)malloc((size_t)n1 * n2 * L * L);
  double *Da = (double *)malloc(sizeof(double) * n4);
  double *La = (double *)malloc(sizeof(double) * n4);
  int64_t *el = (int64_t *)malloc(sizeof(int64_t) * n1 * n2 * 2);
  double *tD = (double *)malloc(sizeof(double) * L2), *tL = (double *)malloc(sizeof(double) * L2);
  double *rD = (double *)malloc(sizeof(double) * L1), *rL = (double *)malloc(sizeof(double) * L1);
  double *sn = (double *)malloc(sizeof(double) * 2 * max_segs);
  double *sD = (double *)malloc(sizeof(double) * 2 * max_segs);
  double *sL = (double *)malloc(sizeof(double) * 2 * max_segs);
  int64_t *si = (int64_t *)malloc(sizeof(int64_t) * 14 * max_segs);
  int64_t best_out[7];
  int rc = -1;
  if (eD && eP && B && Da && La && el && tD && tL && rD && rL && sn && sD && sL && si) {
    rc = or_stage1(C, L1, L2, L, steps, weights, K, nthreads, eD, eP, B, NULL);
    if (rc == 0)
      rc = or_stage2(C, L1, L2, L, steps, K, beta, eD, eP, B, Da, La, el, tD, tL, rD, rL, sn,
                     sD, sL, si, max_segs, best_out, best_cost, path, path_cap, path_len);
  }
  free(eD);
  free(eP);
  free(B);
  free(Da);
  free(La);
  free(el);
  free(tD);
  free(tL);
  free(rD);
  free(rL);
  free(sn);
  free(sD);
  free(sL);
  free(si);
  return rc;
}

/* FlexDTW.py:1207-1210 (L2norm) and Parflex.py:1470 (C = 1 - F1n^T F2n). */
void or_cost_matrix(const double *F1, const double *F2, int64_t Kf, int64_t N,
                    int64_t M, double *C) {
  double *n1 = (double *)malloc(sizeof(double) * N);
  double *n2 = (double *)malloc(sizeof(double) * M);
  for (int64_t i = 0; i < N; ++i) {
    double s = 0.0;
    for (int64_t k = 0; k < Kf; ++k) {
      const double v = F1[k * N + i];
      s += v * v;
    }
    n1[i] = sqrt(s) + 1e-9;
  }
  for (int64_t j = 0; j < M; ++j) {
    double s = 0.0;
    for (int64_t k = 0; k < Kf; ++k) {
      const double v = F2[k * M + j];
      s += v * v;
    }
    n2[j] = sqrt(s) + 1e-9;
  }
  for (int64_t i = 0; i < N; ++i)
    for (int64_t j = 0; j < M; ++j) {
      double s = 0.0;
      for (int64_t k = 0; k < Kf; ++k) s += (F1[k * N + i] / n1[i]) * (F2[k * M + j] / n2[j]);
      C[i * M + j] = 1.0 - s;
    }
  free(n1);
  free(n2);
}
