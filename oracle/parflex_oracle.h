/*
 * CPU oracle for the ParFlex hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library restates, in plain C with IEEE
 * double arithmetic and no FMA contraction, the reference algorithm that the
 * HIP kernels in parflexdtw_amd/csrc must reproduce.  Only tests/, the smoke
 * check and the CPU-baseline leg of bench.py may load it, and only as the
 * checker (or as the timed CPU baseline), never as the product path.
 *
 * Reference files restated (ishitajain21/ParFlexDTW):
 *   FlexDTW.py:1320-1372   per-tile DP (flexdtw)
 *   FlexDTW.py:1098-1148   find_best_endpoint
 *   FlexDTW.py:1155-1180   backtrace_flexdtw
 *   Parflex.py:376-469     tiling (run_flexdtw_on_tiles)
 *   Parflex.py:347-373     _edge_starts_from_S
 *   Parflex.py:509-564     _build_edge_data
 *   Parflex.py:582-624     _build_valid_mask
 *   Parflex.py:631-760     _nb_initialize_chunks_sparse
 *   Parflex.py:763-855     _nb_dp_fill_chunks_sparse
 *   Parflex.py:919-969     sync_tile_overlap_edges
 *   Parflex.py:976-1012    _nb_backtrace_within_chunk
 *   Parflex.py:1032-1103   backtrace_and_stitch
 *   Parflex.py:1110-1150   _build_edge_lookups
 *   Parflex.py:1153-1222   _nb_scan_edges
 *   Parflex.py:1229-1356   stage2_scan_and_stitch
 *   cpu_parflex.py:380-428 stage-1 worker-pool schedule (min(n_tiles,16) workers)
 */
#ifndef PARFLEX_ORACLE_H
#define PARFLEX_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Per-tile FlexDTW DP: fills full D (f64), P (i32), B (i8), each R*W row-major.
 * C is read at C[r*ldc + c].  steps is K pairs (dr, dc). */
void or_flexdtw_dp(const double *C, int64_t ldc, int64_t R, int64_t W,
                   const int64_t *steps, const double *weights, int K,
                   double *D, int32_t *P, int8_t *B);

/* find_best_endpoint: returns number of candidates; fills avg[] (candidate
 * order of the reference), best_cost, best_r, best_c. */
int64_t or_find_best_endpoint(const double *D, const int32_t *P, int64_t R,
                              int64_t W, int64_t buffer, double *avg,
                              double *best_cost, int64_t *best_r,
                              int64_t *best_c);

/* backtrace_flexdtw + reverse: writes points (row, col) pairs, returns count. */
int64_t or_backtrace(const int8_t *B, int64_t R, int64_t W,
                     const int64_t *steps, int64_t r, int64_t c,
                     int64_t *out_pts);

/* Tile grid of Parflex.py:389-391. */
void or_tile_grid(int64_t L1, int64_t L2, int64_t L, int64_t *n1, int64_t *n2);

/* Stage 1 over all tiles of C (L1 x L2) with tile length L.
 *   edge_D  [n1][n2][2][L] f64 (inf beyond the edge length)
 *   edge_P  [n1][n2][2][L] i32 (0 beyond the edge length)
 *   B_all   int8, tile (i,j) at offset i*n2*L*L + j*L*L, row-major R x W
 *           (may be NULL to skip)
 *   best    [n1][n2] per-tile best_cost (buffer=1), may be NULL
 * nthreads workers, each taking whole tiles (cpu_parflex schedule). */
int or_stage1(const double *C, int64_t L1, int64_t L2, int64_t L,
              const int64_t *steps, const double *weights, int K,
              int nthreads, double *edge_D, int32_t *edge_P, int8_t *B_all,
              double *best);

/* Stage 2 of Parflex.py on stage-1 edge outputs.
 * Outputs (caller-allocated):
 *   D_arr, L_arr [n1][n2][2][L] f64
 *   edge_lens    [n1][n2][2] i64
 *   top_D, top_L [L2]; right_D, right_L [L1]
 *   seg_norm/seg_D/seg_L [2][max_segs] f64; seg_i64 [7][2][max_segs]
 *     (ci, cj, et, idx, grow, gcol, valid)
 *   best_out[7] (g_row, g_col, ci, cj, et, idx, found)
 *   path (capacity path_cap points as (row, col)); *path_len
 * Returns 0, or 1 when no endpoint was found (the reference raises). */
int or_stage2(const double *C, int64_t L1, int64_t L2, int64_t L,
              const int64_t *steps, int K, double beta,
              const double *edge_D, const int32_t *edge_P, const int8_t *B_all,
              double *D_arr, double *L_arr, int64_t *edge_lens,
              double *top_D, double *top_L, double *right_D, double *right_L,
              double *seg_norm, double *seg_D, double *seg_L, int64_t *seg_i64,
              int64_t max_segs, int64_t *best_out, double *best_norm,
              int64_t *path, int64_t path_cap, int64_t *path_len);

/* propagate_tile_edge_costs (Parflex.py:862-912): the D_arr / L_arr of
 * stage 2 before the overlap sync, and edge_lens.  Returns 0 or -1. */
int or_propagate(const double *C, int64_t L1, int64_t L2, int64_t L, const double *edge_D,
                 const int32_t *edge_P, double *D_arr, double *L_arr, int64_t *edge_lens);

/* sync_tile_overlap_edges (Parflex.py:919-969) in place on D_arr, L_arr
 * [n1][n2][2][L] with edge_lens [n1][n2][2]. */
void or_sync_overlap(double *D_arr, double *L_arr, const int64_t *edge_lens, int64_t n1,
                     int64_t n2, int64_t L);

/* backtrace_and_stitch from an arbitrary edge slot (for per-segment paths). */
int64_t or_backtrace_and_stitch(int64_t L1, int64_t L2, int64_t L,
                                const int64_t *steps, int K,
                                const int32_t *edge_P, const int8_t *B_all,
                                int64_t ci, int64_t cj, int64_t et, int64_t idx,
                                int64_t *path, int64_t path_cap);

/* End-to-end parflex(C, steps, weights, beta, L): returns 0 / 1 (no
 * endpoint) / -1 (allocation failure). */
int or_parflex(const double *C, int64_t L1, int64_t L2, int64_t L,
               const int64_t *steps, const double *weights, int K, double beta,
               int nthreads, double *best_cost, int64_t *path,
               int64_t path_cap, int64_t *path_len);

/* Cost matrix C = 1 - L2norm(F1)^T L2norm(F2), F1 (K x N), F2 (K x M)
 * row-major (feature-major), FP64 with sequential k summation. */
void or_cost_matrix(const double *F1, const double *F2, int64_t Kf, int64_t N,
                    int64_t M, double *C);

#ifdef __cplusplus
}
#endif
#endif
