/*
 * parflex_hip.h -- C ABI of the MI355X (gfx950) ParFlex kernels.
 *
 * libparflex_hip.so exports exactly the functions below.  All pointers named
 * d_* are device pointers (e.g. torch.Tensor.data_ptr() of a ROCm tensor);
 * the library allocates no device memory in any of them: the caller owns and
 * sizes every buffer (the Python layer parflexdtw_amd/_plan.py computes the
 * sizes).  `stream` is a hipStream_t (torch.cuda.current_stream().cuda_stream)
 * and every call is asynchronous on it.  Every function returns 0 on success
 * or a negative pf_status; pf_last_error() returns a thread-local message.
 *
 * The reference (ishitajain21/ParFlexDTW) is Python: its GPU interface is the
 * set of Python functions in gpu_flexdtw.py / Parflex_gpu.py, which reach one
 * CUDA kernel through CuPy.  Each entry point below names the reference
 * function (file:line) whose work it replaces.
 */
#ifndef PARFLEX_HIP_H
#define PARFLEX_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum pf_status {
  PF_OK = 0,
  PF_EINVAL = -1,   /* bad argument (shape, step set, size) */
  PF_EHIP = -2,     /* a HIP runtime call failed */
  PF_ENOENDPOINT = -3 /* stage 2 found no endpoint (Parflex.py:1285-1286) */
};

/* Precision of C and of the DP accumulator. */
enum pf_dtype { PF_F64 = 0, PF_F32 = 1, PF_BF16 = 2 };

/* Thread-local description of the last error. */
const char *pf_last_error(void);

/* ABI version (major*100 + minor). */
int pf_version(void);

/* Step set: K <= 4 steps (dr, dc), 1 <= dr <= 2, 0 <= dc <= 2, with at least
 * one step of dr == 1 and dc <= 1 (every interior cell then has a valid
 * predecessor); weights as doubles. Validation only, no device work.
 * Replaces the step handling of gpu_flexdtw.py:326-328. */
int pf_check_steps(const int32_t *steps, const double *weights, int K);

/*
 * Tile table: int64[8] per tile, row-major, on the device:
 *   [0] c_off  element offset of the tile's (0,0) cell in the C buffer
 *   [1] ldc    row stride of C in elements
 *   [2] R      tile rows       [3] W  tile columns
 *   [4] b_off  offset (uint32 words) of the tile's packed back-pointers
 *   [5] e_off  offset (elements) of the tile's edge slots (2*Lmax per tile)
 *   [6] pair   owning pair     [7] (i << 32) | j tile coordinates
 * Packed back-pointers: word (rb, c) = b_off + rb*W + c holds the 2-bit step
 * indices of rows 16*rb .. 16*rb+15 of column c, row r at bit 2*(r%16).
 * Edge slots: edge_D/edge_P[e_off + e*Lmax + pos], e = 0 last row (pos =
 * column), e = 1 last column (pos = row); edge_C[e_off + e*Lmax + pos] holds
 * C of the FIRST row (e = 0) and FIRST column (e = 1) of the tile, the start
 * cells of the tile's paths.
 */

/* Stage 1, all tiles of a tile table in one launch: per-tile FlexDTW DP
 * writing packed back-pointers and D/P on the tile edges only.
 * Replaces the tile loop of Parflex_gpu.run_flexdtw_on_tiles
 * (Parflex_gpu.py:461-563) with its per-tile flexdtw_chunk_from_global_C
 * (gpu_flexdtw.py:288-402) and _edge_starts_from_S (Parflex.py:347-373).
 * dtype: PF_F64 (C f64, DP f64) or PF_F32 (C f32, DP f32, edges f64) or
 * PF_BF16 (C bf16, DP f32, edges f64).
 * Launch shape: each tile is cut into `strips` column strips, one workgroup
 * of `waves` waves per strip (strips*waves*64*8 >= Lmax); strips of a tile
 * hand their boundary columns to the right every `rows_per_block` rows
 * (1..64) through d_xch (ntiles*strips*Lmax entries of 32 bytes for PF_F64,
 * 16 bytes otherwise) and d_flags (int32, ntiles*strips + 1).  With
 * strips == 1 both may be null.  d_edge_D, d_edge_P, d_edge_C: ntiles*2*Lmax
 * entries each (f64, int32, f64). */
int pf_stage1(int dtype, const void *d_C, const int64_t *d_tiles, int64_t ntiles,
              int64_t Lmax, const int32_t *steps, const double *weights, int K,
              int strips, int waves, int rows_per_block, int32_t *d_flags,
              void *d_xch, uint32_t *d_Bpacked, double *d_edge_D, int32_t *d_edge_P,
              double *d_edge_C, void *stream);

/* pf_stage1 for callers that cannot allocate device memory (code under HIP
 * graph capture, serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  The strip-exchange counters and rows that
 * pf_stage1 takes as d_flags / d_xch come from a scratch array the library
 * holds in device memory for the whole process, PF_NOALLOC_SCRATCH_BYTES
 * per GPU; this call allocates nothing (no hipMalloc, no library
 * allocation).  Arguments and results as for pf_stage1; a launch whose
 * exchange needs more than the scratch (align256(4*(ntiles*strips+1)) +
 * pf_stage1_workspace's xch_bytes) is rejected with PF_EINVAL.  Every call
 * on one GPU shares that GPU's scratch: calls that overlap in time on one
 * GPU (other streams, other threads) race and must be serialised by the
 * caller. */
#define PF_NOALLOC_SCRATCH_BYTES (32u << 20)
int pf_stage1_noalloc(int dtype, const void *d_C, const int64_t *d_tiles, int64_t ntiles,
                      int64_t Lmax, const int32_t *steps, const double *weights, int K,
                      int strips, int waves, int rows_per_block, uint32_t *d_Bpacked,
                      double *d_edge_D, int32_t *d_edge_P, double *d_edge_C, void *stream);

/* Stage 1 computing the cost inside the sweep from features (the fused
 * form of tiled_stage1_from_features, Parflex_gpu.py:610-653): C = 1 -
 * F1n^T F2n is never stored; every 32 rows each wave computes its block of C
 * on the MFMA units from d_F1b / d_F2b (frames x Kp bf16, the L2-normalised
 * features from pf_features_bf16), with the operands and rounding of
 * pf_cost(PF_BF16), so the results equal pf_stage1(PF_BF16) on that cost.
 * Tile table as for pf_stage1 except columns 0 and 1, which hold the frame
 * (row in d_F1b / d_F2b) of the tile's first row and first column.  FP32 DP;
 * 2 or 4 columns per lane (strips*waves*64*4 >= Lmax).  Other arguments as
 * for pf_stage1 with dtype PF_BF16. */
int pf_stage1_features(const void *d_F1b, const void *d_F2b, int Kp, const int64_t *d_tiles,
                       int64_t ntiles, int64_t Lmax, const int32_t *steps, const double *weights,
                       int K, int strips, int waves, int rows_per_block, int32_t *d_flags,
                       void *d_xch, uint32_t *d_Bpacked, double *d_edge_D, int32_t *d_edge_P,
                       double *d_edge_C, void *stream);

/* Bounded waits of the stage-1 kernels.  A workgroup's waves, and the strips
 * of a tile, wait for one another (a wave for row r-1 of the wave on its
 * left, a strip for the rows the strip on its left has published); each wait
 * polls at most `polls` times (default 4194304, seconds of waiting).  A wave
 * whose wait gives up sets a device error word and finishes its rows without
 * waiting again, so every kernel ends.  pf_set_wait_limit sets the limit for
 * later launches (process-wide) and returns the previous one in *previous
 * (may be null).  No reference counterpart: the reference's kernel has no
 * inter-workgroup waits. */
int pf_set_wait_limit(int64_t polls, int64_t *previous);
/* Synchronises `stream`, then returns PF_EHIP (and clears the words) if a
 * stage-1 wait or a wait of pf_stage2_propagate_dag on the current device
 * gave up since the last call, else 0. */
int pf_device_status(void *stream);

/* L2-normalised features as frames x Kp bf16 (Kp = K rounded up to 16, zero
 * padded) from F (K x n f64, FlexDTW.py:1207-1210 normalisation). */
int pf_features_bf16(const double *d_F, int64_t K, int64_t n, void *d_out, void *stream);

/* Sizes of the strip-exchange workspace of pf_stage1: int32 count of
 * d_flags and byte size of d_xch (both 0 when strips == 1). */
int pf_stage1_workspace(int dtype, int64_t ntiles, int64_t Lmax, int strips,
                        int64_t *flags_count, int64_t *xch_bytes);

/* Per-tile DP with full D/P/B outputs for one tile C[r0:r1, c0:c1] of a
 * row-major FP64 matrix with leading dimension ldc (FlexDTW.py:1320-1364;
 * gpu_flexdtw.py:146-263,356-393).  D is f64 R*W, P int32, B int8. */
int pf_tile_full(const double *d_C, int64_t ldc, int64_t r0, int64_t r1,
                 int64_t c0, int64_t c1, const int32_t *steps,
                 const double *weights, int K, double *d_D, int32_t *d_P,
                 int8_t *d_B, void *stream);

/* find_best_endpoint (FlexDTW.py:1098-1148) on full D/P of one tile:
 * d_avg receives the (W-buffer)+(R-1-buffer) candidate averages in the
 * reference's order; d_best receives {best_cost, (double)best_r,
 * (double)best_c}. */
int pf_tile_endpoint(const double *d_D, const int32_t *d_P, int64_t R, int64_t W,
                     int64_t buffer, double *d_avg, double *d_best, void *stream);

/* backtrace_flexdtw + reverse (FlexDTW.py:1155-1180,1366-1368) on full int8
 * B from the endpoint in d_best (as written by pf_tile_endpoint).
 * d_path receives (row, col) int64 pairs, capacity R+W; *d_len the count. */
int pf_tile_backtrace(const int8_t *d_B, int64_t R, int64_t W,
                      const int32_t *steps, int K, const double *d_best,
                      int64_t *d_path, int64_t *d_len, void *stream);

/*
 * Stage 2 for a batch of pairs that share the tile length L (= Lmax).
 * Pair table: int64[16] per pair on the device:
 *   [0] L1 [1] L2 [2] n1 [3] n2 [4] tile_base (index of tile (0,0))
 *   [5] c_off (element offset of the pair's C) [6] ldc
 *   [7] buf (int(buffer_global), Parflex.py:1248)
 *   [8] lk_off (offset of the pair's L2+L1 candidate slots)
 *   [9] seg_off (offset of the pair's 2*max_segs segment slots) [10] max_segs
 *   [11] path_off (offset, in points, of the pair's path buffer; capacity L1+L2+8)
 *   [12..15] reserved (0)
 * Tile table: as for pf_stage1.  A tile whose column 0 is negative had no
 * stage 1 (the learned form's pruned tiles): its edges read as D = +inf,
 * P = 0, C = 0, and its slots of d_edge_D / d_edge_P / d_edge_C are never
 * read (they need no initialisation); its back-pointer offset must point
 * at readable zero words.
 * Workspace (device, caller-allocated, sizes from parflexdtw_amd/_plan.py):
 *   d_D_arr, d_L_arr, d_olap  f64 [ntiles*2*L]      (Parflex.py:509-564,631-855)
 *   d_bits                    uint32 [ntiles*2*bitw], bitw = (L+32)/32
 *   d_cand_D, d_cand_L        f64 [sum(L2+L1)]      (edge_summary top/right D, L)
 *   d_seg_f                   f64 [sum(2*max_segs)*4] (norm, D, L, order)
 *   d_seg_i                   int64 [sum(2*max_segs)*8] (ci,cj,et,idx,grow,gcol,found,order)
 *   d_best_f                  f64 [npairs]           best normalised cost
 *   d_best_i                  int64 [npairs*8]       (g_row,g_col,ci,cj,et,idx,found,order)
 *   d_chain                   int64 [npairs*101*8]   stitched segment list
 *   d_pts                     int64 [npairs*101*(2*L+8)*2] per-segment points
 *   d_path                    int64 [sum(L1+L2+8)*2]  stitched path (row, col)
 *   d_path_len                int64 [npairs]
 * Batch dimensions (host values): max_diag = max(n1+n2-1), max_tile_diag =
 * max(min(n1,n2)), max_tiles_pair = max(n1*n2), max_segs = max(max_segs).
 * Replaces Parflex_gpu.propagate_tile_edge_costs (Parflex_gpu.py:247-339),
 * Parflex.sync_tile_overlap_edges (Parflex.py:919-969) and
 * Parflex.stage2_scan_and_stitch (Parflex.py:1229-1356).  Only the path (and,
 * on request, the edge arrays) need to be copied back.
 */
int pf_stage2(const double *d_edge_C, const int64_t *d_pairs, int64_t npairs,
              const int64_t *d_tiles, int64_t ntiles, int64_t L, int64_t max_diag,
              int64_t max_tile_diag, int64_t max_tiles_pair, int64_t max_segs,
              const int32_t *steps, int K, const uint32_t *d_Bpacked,
              const double *d_edge_D, const int32_t *d_edge_P, double *d_D_arr,
              double *d_L_arr, double *d_olap, uint32_t *d_bits, double *d_cand_D,
              double *d_cand_L, double *d_seg_f, int64_t *d_seg_i, double *d_best_f,
              int64_t *d_best_i, int64_t *d_chain, int64_t *d_pts, int64_t *d_path,
              int64_t *d_path_len, void *stream);

/* The phases of pf_stage2 on their own, for callers that drive stage 2 step
 * by step as the reference's Python functions do.  `phases` is a bit set,
 * run in this order:
 *   PF_S2_PROPAGATE  edge data, start bitmaps, valid mask, init and fill:
 *                    D_arr, L_arr before the overlap sync
 *                    (propagate_tile_edge_costs, Parflex_gpu.py:247-339,
 *                    Parflex.py:862-912);
 *   PF_S2_SYNC       the three overlap passes on D_arr, L_arr in place
 *                    (sync_tile_overlap_edges, Parflex.py:919-969); the edge
 *                    lengths are the tile table's W (edge 0) and R (edge 1),
 *                    each 0 or at least 2;
 *   PF_S2_SCAN       candidate scan, segment and global bests
 *                    (stage2_scan_and_stitch, Parflex.py:1229-1290, with
 *                    _nb_scan_edges, Parflex.py:1153-1222);
 *   PF_S2_STITCH     backtrace_and_stitch of each pair's best endpoint
 *                    (Parflex.py:1032-1103, 1318-1324).
 * Arguments as for pf_stage2; buffers that no requested phase uses may be
 * null.  pf_stage2 is all four phases. */
#define PF_S2_PROPAGATE 1
#define PF_S2_SYNC 2
#define PF_S2_SCAN 4
#define PF_S2_STITCH 8
int pf_stage2_phases(int phases, const double *d_edge_C, const int64_t *d_pairs, int64_t npairs,
                     const int64_t *d_tiles, int64_t ntiles, int64_t L, int64_t max_diag,
                     int64_t max_tile_diag, int64_t max_tiles_pair, int64_t max_segs,
                     const int32_t *steps, int K, const uint32_t *d_Bpacked,
                     const double *d_edge_D, const int32_t *d_edge_P, double *d_D_arr,
                     double *d_L_arr, double *d_olap, uint32_t *d_bits, double *d_cand_D,
                     double *d_cand_L, double *d_seg_f, int64_t *d_seg_i, double *d_best_f,
                     int64_t *d_best_i, int64_t *d_chain, int64_t *d_pts, int64_t *d_path,
                     int64_t *d_path_len, void *stream);

/* The PF_S2_PROPAGATE phase with each tile anti-diagonal launched over a
 * list of its tiles, for a batch where most tiles had no stage 1 (the
 * learned form): the tiles with stage 1 get a thread per slot, the others
 * one per edge (only their corner copies take a value), instead of a thread
 * for every slot of every tile on the diagonal.  d_list: int32 tile-table
 * indices, every tile once, diagonal by diagonal (the diagonal of tile (i,
 * j) is i + j); h_diag (host): int64[3] per diagonal d = 0 .. ndiag-1:
 * {offset of its list in d_list, tiles with stage 1 (listed first), tiles
 * without}.  Results equal those of PF_S2_PROPAGATE; the caller then runs
 * pf_stage2_phases(PF_S2_SYNC | PF_S2_SCAN | PF_S2_STITCH).  No reference
 * counterpart (the reference computes every tile). */
int pf_stage2_propagate_listed(const double *d_edge_C, const int64_t *d_pairs, int64_t npairs,
                               const int64_t *d_tiles, int64_t ntiles, int64_t L,
                               const int32_t *d_list, const int64_t *h_diag, int64_t ndiag,
                               const double *d_edge_D, const int32_t *d_edge_P, double *d_D_arr,
                               double *d_L_arr, double *d_olap, uint32_t *d_bits, void *stream);

/* The same phase as one launch over the tiles with stage 1 only, each tile
 * after the tiles it reads (a tile DAG): the workgroups take the live tiles
 * in diagonal order from a ticket counter and wait for the done flags of
 * their live neighbours below, left and below-left; a finished tile copies
 * its corners into the position-0 slots of pruned neighbours (the only
 * slots of a pruned tile the phase changes).  d_list as for
 * pf_stage2_propagate_listed; d_diag: its per-diagonal table on the device
 * (as pf_select_tiles writes it at d_counts + 3); d_work: ntiles + 1 int32
 * (done flags and the counter, cleared here).  The overlap costs are taken
 * from d_edge_C as needed (no d_olap table).  Results equal those of
 * PF_S2_PROPAGATE (D everywhere it is read, L where D is finite).  Waits
 * poll at most pf_set_wait_limit times; one that gives up is reported by
 * pf_device_status.  No reference counterpart (the reference computes every
 * tile, one diagonal after another: Parflex.py:763-855). */
int pf_stage2_propagate_dag(const double *d_edge_C, const int64_t *d_pairs, int64_t npairs,
                            const int64_t *d_tiles, int64_t ntiles, int64_t L,
                            const int32_t *d_list, const int64_t *d_diag, int64_t ndiag,
                            const double *d_edge_D, const int32_t *d_edge_P, double *d_D_arr,
                            double *d_L_arr, uint32_t *d_bits, int32_t *d_work, void *stream);

/* backtrace_and_stitch (Parflex.py:1032-1103) for n requests, each
 * int64[5] = (pair, ci, cj, et, idx) in d_req; stage-1 outputs as for
 * pf_stage2.  Request q writes its path at d_path + q*path_cap*2 and its
 * length to d_path_len[q].  d_chain: int64[n*101*8]; d_pts: int64
 * [n*101*(2L+8)*2]. */
int pf_stitch(const int64_t *d_req, int64_t nreq, const int64_t *d_pairs,
              const int64_t *d_tiles, int64_t L, const int32_t *steps, int K,
              const uint32_t *d_Bpacked, const int32_t *d_edge_P, int64_t *d_chain,
              int64_t *d_pts, int64_t path_cap, int64_t *d_path, int64_t *d_path_len,
              void *stream);

/* Cost matrix C = 1 - L2norm(F1)^T L2norm(F2) (FlexDTW.py:1207-1210,
 * Parflex.py:1470; gpu_flexdtw.py:95-109).  F1 (K x N), F2 (K x M) f64
 * row-major.  out_dtype PF_F64: fixed k order, no FMA contraction
 * (bit-stable); PF_F32 / PF_BF16: MFMA-tiled BF16 product with FP32
 * accumulation, written as f32 or bf16.  d_C row stride ldc.
 * d_work: scratch for the normalised features, pf_cost_workspace bytes,
 * 16-byte aligned. */
int pf_cost(int out_dtype, const double *d_F1, const double *d_F2, int64_t K,
            int64_t N, int64_t M, void *d_C, int64_t ldc, double *d_work,
            void *stream);

/* Bytes of d_work that pf_cost needs for out_dtype (for pf_cost_batch pass
 * the total frame counts n1_total, n2_total as N, M). */
int pf_cost_workspace(int out_dtype, int64_t K, int64_t N, int64_t M, int64_t *bytes);

/* Rows x columns of C that one workgroup of the cost kernel for out_dtype
 * covers (the unit of blk0 in pf_cost_batch job tables). */
int pf_cost_tile(int out_dtype, int64_t *rows, int64_t *cols);

/* Cost matrices of many pairs in one launch (the batched form of
 * gpu_flexdtw.py:95-109 / Parflex.py:1470).  d_F1 (K, n1_total) and d_F2
 * (K, n2_total): the feature arrays of every pair concatenated along frames.
 * d_jobs: int64[8] per pair, (f1_off, f2_off, N, M, c_off, ldc, blk0, 0):
 * the pair's frames [f1_off, f1_off+N) x [f2_off, f2_off+M), its C at
 * element offset c_off with row stride ldc, and blk0 = sum over earlier
 * pairs of ceil(N/rows)*ceil(M/cols) (pf_cost_tile); total_blocks = the
 * same sum over all pairs.  Same arithmetic as pf_cost, bit for bit. */
int pf_cost_batch(int out_dtype, const double *d_F1, int64_t n1_total, const double *d_F2,
                  int64_t n2_total, int64_t K, const int64_t *d_jobs, int64_t njobs,
                  int64_t total_blocks, void *d_C, double *d_work, void *stream);

/*
 * Learned tile proposal (no reference counterpart: the reference computes
 * every tile).  A 4-layer 3x3 convolution network over the 1/pool-pooled
 * cost matrix scores each tile; only tiles above a threshold (plus a margin)
 * run stage 1.  Activations are NHWC bf16 with 32 channels; weights of one
 * layer are bf16 [32 cout][288], k = (dy*3 + dx)*32 + cin; master weights,
 * gradients and Adam moments are f32 in the same layout followed by 32
 * biases.
 */
/* Fp (K x ceil(n/pool)) f32: means of the L2-normalised frames of each
 * pool-frame block (FlexDTW.py:1207-1210 normalisation); pool must be 32. */
int pf_pool_features(const double *d_F, int64_t K, int64_t n, int pool, float *d_Fp,
                     void *stream);
/* The same for many frame runs in one launch (every pair of a batch):
 * d_seg holds nseg rows of int64 {f_off, frames, b_off, 0}; run s pools
 * frames [f_off, f_off+frames) of d_F (K x n, row stride n) into columns
 * [b_off, b_off + ceil(frames/32)) of d_Fp (K x nb, row stride nb).  Rows
 * are in increasing b_off; blocks of one run are contiguous. */
int pf_pool_features_batch(const double *d_F, int64_t K, int64_t n, const int64_t *d_seg,
                           int64_t nseg, int pool, float *d_Fp, int64_t nb, void *stream);
/* pf_pool_features_batch that also writes the BF16 features of every frame
 * the runs cover to d_Fb (n x Kp bf16, frame-major, Kp = K rounded up to 16,
 * 16-byte aligned): the values and layout of pf_features_bf16, bit for bit,
 * from the same reads of d_F.  Frames no run covers are not written.  No
 * reference counterpart: the learned form's pooling (FlexDTW.py:1207-1210
 * normalisation) and the fused form's feature preparation in one pass. */
int pf_pool_features_bf16_batch(const double *d_F, int64_t K, int64_t n, const int64_t *d_seg,
                                int64_t nseg, int pool, float *d_Fp, int64_t nb, void *d_Fb,
                                void *stream);
/* Input statistics of many pairs in one launch (K <= 16): pair p's pooled
 * features are columns [a_off, a_off+H) of d_A (K x lda) and [b_off, b_off+W)
 * of d_B (K x ldb), d_pairs rows int64 {a_off, H, b_off, W}; d_st + p*(2K+1)
 * receives the feature means of both sides and 1 / std of the double-centred
 * pooled cost (the statistics pf_pooled_cost computes, in the same order). */
int pf_input_stats_batch(const float *d_A, int64_t lda, const float *d_B, int64_t ldb,
                         int64_t K, const int64_t *d_pairs, int64_t npairs, float *d_st,
                         void *stream);
/* First layer (as pf_conv3x3 mode 3) computed from one pair's pooled
 * features and statistics, without an input map: equal bit for bit to
 * pf_pooled_cost(d_Xc) followed by pf_conv3x3 mode 3.  K <= 16. */
int pf_conv3x3_first_pooled(const float *d_A, int64_t lda, int64_t a_off, int64_t H,
                            const float *d_B, int64_t ldb, int64_t b_off, int64_t W, int64_t K,
                            const float *d_st, const void *d_Wt, const float *d_bias, void *d_Y,
                            void *stream);
/* Network input from the pooled features: the pooled cost P = 1 - A^T B
 * double-centred (row and column means removed, global mean added back) and
 * divided by its standard deviation, written as d_X (bf16 NHWC 32, channel 0,
 * other channels 0) and/or d_Xc (bf16 H*W, one channel); if d_P is not null,
 * also P itself (H x W f32).  d_work: pf_pooled_cost_workspace floats. */
int pf_pooled_cost_workspace(int64_t K, int64_t *floats);
int pf_pooled_cost(const float *d_A, int64_t H, const float *d_B, int64_t W, int64_t K,
                   void *d_X, void *d_Xc, float *d_P, float *d_work, void *stream);
/* One 3x3 convolution layer on MFMA.  mode 0: relu(conv + bias) -> bf16
 * NHWC 32; mode 1: channel 0 of conv + bias -> f32 H*W (the logit map);
 * mode 2: gradient w.r.t. the layer input, d_Wt = flipped weights
 * (pf_pack_weights), masked by d_M > 0 -> bf16 NHWC 32; mode 3: as mode 0
 * for the first layer from the one-channel input d_X = d_Xc of
 * pf_pooled_cost (only input channel 0 of the weights is read).  X, the
 * weights and the bias are 16-byte aligned, Y and M 8-byte aligned. */
int pf_conv3x3(int mode, const void *d_X, int64_t H, int64_t W, const void *d_Wt,
               const float *d_bias, const void *d_M, void *d_Y, void *stream);
/* Weight and bias gradient of one layer, accumulated into d_grad (f32
 * [32*288 + 32]); d_part: pf_conv3x3_wgrad_workspace floats. */
int pf_conv3x3_wgrad_workspace(int64_t H, int64_t W, int64_t *floats);
int pf_conv3x3_wgrad(const void *d_dY, const void *d_X, int64_t H, int64_t W, float *d_grad,
                     float *d_part, void *stream);
/* Weighted binary cross-entropy on the logit map d_Z (f32 H*W): writes the
 * logit gradient d_dZ (bf16 NHWC 32, channel 0, scaled by `scale`) and adds
 * the scaled loss to *d_loss. */
int pf_bce(const float *d_Z, const uint8_t *d_label, int64_t npix, float pos_weight,
           float scale, void *d_dZ, float *d_loss, void *stream);
/* Adam step (bias-corrected, step >= 1) on f32 master weights; clears d_g.
 * d_frozen (uint8 per weight, may be null) keeps marked weights fixed. */
int pf_adam(float *d_w, float *d_g, float *d_m, float *d_v, const uint8_t *d_frozen, int64_t n,
            float lr, float b1, float b2, float eps, int64_t step, void *stream);
/* bf16 weights of one layer and, if d_Wf is not null, the flipped copy used
 * by mode 2 of pf_conv3x3. */
int pf_pack_weights(const float *d_w, void *d_Wt, void *d_Wf, void *stream);
/* Per-tile score = max logit (map d_Z, f32 H*W) over the pooled cells that tile
 * (i, j) of tile length L covers (tiling of Parflex.py:388-427), the region
 * extended by ext cells on every side; d_score has n1*n2 entries in i-major
 * order. */
int pf_tile_scores(const float *d_Z, int64_t H, int64_t W, int pool, int64_t L, int64_t L1,
                   int64_t L2, int ext, float *d_score, void *stream);

/* The four layers of the network (pf_conv3x3_first_pooled, pf_conv3x3 mode 0
 * twice, pf_conv3x3 mode 1) for every pair of a batch in one launch: the
 * logit maps equal those of the per-layer calls bit for bit, with no
 * activation map in HBM.  d_A (K x lda) / d_B (K x ldb): pooled features of
 * every pair (pf_pool_features_batch); d_st: the statistics of
 * pf_input_stats_batch.  Pair table int64[8] per pair: {a_off, H, b_off, W,
 * z_off, st_off, 0, 0}: the pair's pooled columns of d_A and d_B, its map
 * size, its logit map at d_Z + z_off (f32 H x W) and its statistics at
 * d_st + st_off.  Job table int64[4] per job: {pair, x0, y0, y1}: output
 * columns x0 .. x0 + cols - 1 (clipped to W) and rows y0 .. y1 - 1, with
 * x0 a multiple of cols and y1 - y0 <= rows from pf_tile_net_geometry.
 * Weights W1..W4 bf16 [32][288] (pf_pack_weights), biases b1..b4 f32 [32],
 * all 16-byte aligned; K <= 16. */
int pf_tile_net_geometry(int64_t *cols_per_job, int64_t *rows_per_job);
int pf_tile_net(const float *d_A, int64_t lda, const float *d_B, int64_t ldb, int64_t K,
                const float *d_st, const int64_t *d_pairs, int64_t npairs, const int64_t *d_jobs,
                int64_t njobs, const void *d_W1, const float *d_b1, const void *d_W2,
                const float *d_b2, const void *d_W3, const float *d_b3, const void *d_W4,
                const float *d_b4, float *d_Z, void *stream);
/* pf_tile_scores for every pair of a batch in one launch: tile t (the
 * batch's tile order, pair by pair, i-major) belongs to pair d_tpair[t]
 * (int32), whose row of d_pairs (int64[8]) is {H, W, z_off, L1, L2, n2,
 * tile_base, 0}; d_score[t] as pf_tile_scores gives it for that pair's map
 * at d_Z + z_off. */
int pf_tile_scores_batch(const float *d_Z, const int64_t *d_pairs, const int32_t *d_tpair,
                         int64_t ntiles, int pool, int64_t L, int ext, float *d_score,
                         void *stream);

/* Tile selection of the learned form (this library's tile proposal) on the
 * device: the rule of tile_net.select_tiles with margin 0, tile t selected
 * when d_score[t] >= threshold in FP64, over the
 * plan's tile table d_plan (ntiles x int64[8], as for pf_stage1), one
 * workgroup.  d_frames (npairs x int64[2]): each pair's first frame in the
 * concatenated feature arrays; d_diag_off (max_diag + 1): first list slot
 * of every tile diagonal d = i + j (the plan's tiles per diagonal, summed).
 * Writes d_sel (uint8 per tile), d_full (the tile table with b_off packed
 * for the selected tiles after the zero region of ceil(L/16)*L words, 0
 * and c_off = -1 for the pruned), d_work (the selected tiles in plan order,
 * c_off / ldc replaced by the frames of the tile's first row and column, as
 * pf_stage1_features takes them), d_list (every tile once, diagonal by
 * diagonal, selected first, as pf_stage2_propagate_listed takes it) and
 * d_counts (int64[3 + 3 max_diag]: selected tiles, packed back-pointer words
 * with the zero region, selected cells, then h_diag of
 * pf_stage2_propagate_listed).  d_tmp: workspace of
 * 3 * ceil(ntiles / 1024) + 2 * max_diag int64.  max_diag <= 4096. */
int pf_select_tiles(const float *d_score, double threshold, const int64_t *d_plan, int64_t ntiles,
                    const int64_t *d_frames, int64_t L, const int64_t *d_diag_off,
                    int64_t max_diag, uint8_t *d_sel, int64_t *d_full, int64_t *d_work,
                    int32_t *d_list, int64_t *d_counts, int64_t *d_tmp, void *stream);

/* Selection check of the learned form: d_flag[p] |= 1 (int32, cleared by
 * the caller) when a tile holding a cell within `guard` frames of a point of
 * pair p's stitched path (d_path / d_path_len as written by pf_stage2, pair
 * table as for pf_stage2) is not selected in d_sel (uint8 per tile of the
 * tile table).  max_path >= every path length; 2*guard < L-1. */
int pf_selection_check(const int64_t *d_pairs, int64_t npairs, int64_t max_path,
                       const int64_t *d_path, const int64_t *d_path_len, const uint8_t *d_sel,
                       int64_t L, int64_t guard, int32_t *d_flag, void *stream);

/* Stitched paths of every pair (d_path / d_path_len as written by
 * pf_stage2, pair table as for pf_stage2) packed back to back as int64 (2, n)
 * arrays: pair p at offset 2 * sum_{q<p} len_q, rows then columns, so that
 * one copy of 2 * sum(len) values brings every path to the host in the
 * layout the reference returns (wp of shape (2, n), Parflex.py:1543-1548).
 * max_path >= every path length; d_out holds at least 2 * sum(len). */
int pf_pack_paths(const int64_t *d_pairs, int64_t npairs, int64_t max_path,
                  const int64_t *d_path, const int64_t *d_path_len, int64_t *d_out, void *stream);

/* Alignment errors of many warping paths against ground-truth time pairs,
 * and the error-rate-versus-tolerance histogram, in one launch.  Replaces
 * the per-file loop of 05_Evaluate_Benchmarks_L.ipynb#cell7-8 (and
 * eval_tools.py:33-46 with hop = 1) and the tolerance sweep of #cell23.
 *
 * Query q's path has d_lens[q] points; point j is (d_path[d_offs[2q] +
 * j*stride], d_path[d_offs[2q+1] + j*stride]) in frames (int64), rows
 * non-decreasing.  Stage 2's interleaved buffer is stride 2 with offsets
 * (2*path_off, 2*path_off + 1); a packed (2, n) array is stride 1 with
 * offsets (o, o + n).  Query q's beats are [d_gt_off[q], d_gt_off[q+1]) of
 * d_gt_x (query times, seconds) and d_gt_y (reference times, seconds);
 * max_beats >= every query's beat count.  For each beat b:
 *   d_err[b] = np.interp(gt_x[b], rows * hop, cols * hop) - gt_y[b]
 * with np.interp's semantics (left/right clamping, NaN x gives NaN, one
 * sample point gives that point), bit for bit; +inf for a query with an
 * empty path (05_Evaluate_Benchmarks_L.ipynb#cell8's value for a missing
 * hypothesis).  If d_hist is not null (uint64, ntol + 1 bins, accumulated:
 * the caller clears it), d_hist[k] += 1 for k = #{i : |err| > d_tol[i]}
 * (d_tol ascending, 1 <= ntol <= PF_EVAL_MAX_TOL), so that the error rate
 * at threshold i is sum_{k > i} d_hist[k] / #beats. */
#define PF_EVAL_MAX_TOL 4096
int pf_eval_paths(const int64_t *d_path, int64_t stride, const int64_t *d_offs,
                  const int64_t *d_lens, int64_t nq, const int64_t *d_gt_off, int64_t max_beats,
                  const double *d_gt_x, const double *d_gt_y, double hop, const double *d_tol,
                  int64_t ntol, double *d_err, uint64_t *d_hist, void *stream);

/* Measurement helper (no counterpart in the reference): `blocks` workgroups
 * of 256 threads, each lane running `iters` rounds of 16 independent FP32
 * fused multiply-adds, 16 * iters VALU wave instructions per wave; d_out
 * receives blocks * 256 floats.  bench.py times it to state the stage-1
 * kernel's VALU instruction rate as a fraction of the rate this dense FMA
 * loop reaches on the same GPU. */
int pf_valu_probe(float *d_out, int blocks, int iters, void *stream);

/* Fill helpers used instead of hipMemsetAsync (safe under graph capture). */
int pf_fill_f64(double *d_p, double v, int64_t n, void *stream);
int pf_fill_i64(int64_t *d_p, int64_t v, int64_t n, void *stream);

#ifdef __cplusplus
}
#endif
#endif
