// Stage 2 of ParFlex on gfx950: edge-cost propagation across tiles, overlap
// sync, global boundary scan, stitch and backtrace -- all on the device, for
// a batch of pairs sharing one tile length L.
//
// Semantics follow Parflex.py exactly (file:line at each kernel).  Layout:
// every per-tile quantity lives in slot t*2L + e*L + pos of a tile-indexed
// array (e = 0: tile's last row indexed by column, e = 1: last column indexed
// by row), so a pair's (n1, n2, 2, L) arrays of the reference are contiguous
// slices starting at tile_base*2L.
//
// Parallel structure:
//  * edge data, start sets and overlap costs: one thread per edge slot;
//  * init/fill (Parflex.py:631-855): tile (i,j) reads only tiles (i-1,j) and
//    (i,j-1), so the i-major sweep of the reference equals a sweep over tile
//    anti-diagonals; one launch per anti-diagonal, one thread per slot;
//  * sync (Parflex.py:919-969): each of the three passes touches disjoint
//    slot pairs (edges are >= 2 long), so each pass is one parallel launch;
//  * scan (Parflex.py:1153-1222): one workgroup per (pair, side, segment of L
//    positions) reduces (norm, scan order) lexicographically -- the strict '<'
//    of the sequential scan keeps the first minimum -- then one workgroup per
//    pair reduces the segment winners (global threshold 1e9);
//  * stitch (Parflex.py:1032-1103): one thread per request walks the <= 101
//    tiles (only edge P is needed), then one wave per segment walks the
//    packed back-pointers through a register window (pf_walk.hip), and one
//    workgroup per request concatenates, removes the duplicated overlap cells
//    and reverses.
#include "pf_common.h"

namespace pf {

struct Geo {
  int64_t s1, s2, rows, cols;
};

__device__ __forceinline__ Geo tile_geo(const PairDesc &pd, int64_t L, int64_t i, int64_t j) {
  const int64_t hop = L - 1;
  Geo g;
  g.s1 = i * hop;
  g.s2 = j * hop;
  const int64_t e1 = g.s1 + L < pd.L1 ? g.s1 + L : pd.L1;
  const int64_t e2 = g.s2 + L < pd.L2 ? g.s2 + L : pd.L2;
  g.rows = e1 - g.s1;
  g.cols = e2 - g.s2;
  return g;
}

__device__ __forceinline__ int64_t iabs64(int64_t v) { return v < 0 ? -v : v; }

__global__ void k_fill_f64(double *p, double v, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}
__global__ void k_fill_i64(int64_t *p, int64_t v, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// _build_edge_data (Parflex.py:509-564) + _edge_starts_from_S
// (Parflex.py:347-373): overlap cost C[start] per slot (from the first row /
// first column of C that stage 1 stored per tile), start-set bitmaps of
// every tile (bit s of map 0: some edge cell starts at (0, s), s > 0; bit a
// of map 1: some edge cell starts at (a, 0), a >= 0), and D_arr = L_arr =
// inf (L_arr of tiles with stage 1 only).
//
// A tile without stage 1 (column 0 < 0) gets D = inf only in the slots that
// some later step reads: position 0 and the corner of each edge (corner
// copies, the overlap sync), its last row when the tile above has stage 1
// or it is in the top tile row (chaining, the top scan), its last column
// when the tile to the right has stage 1 or it is in the right tile column.
// Every other slot of such a tile is never read: a tile without stage 1
// chains from nowhere, and chaining reads only the edges of the tiles below
// and to the left of a tile with stage 1.
constexpr int kMaxBitw = 8192 / 32 + 1;

__global__ __launch_bounds__(256) void k_s2_prep(const PairDesc *__restrict__ pairs,
                                                 const TileDesc *__restrict__ tiles, int64_t L,
                                                 const int32_t *__restrict__ eP,
                                                 const double *__restrict__ eC,
                                                 double *__restrict__ Darr,
                                                 double *__restrict__ Larr,
                                                 double *__restrict__ olap,
                                                 uint32_t *__restrict__ bits, int64_t bitw,
                                                 int32_t *__restrict__ dag, int64_t ntiles,
                                                 int sweep_all) {
  // A workgroup per tile at a time (tiles blockIdx.x, + gridDim.x, ...): its
  // 2L slots, and its two start bitmaps built in LDS and stored whole (the
  // workgroup owns them: no global atomics and no clearing pass).  With dag
  // (k_s2_dag follows): the tile's done flag and, in workgroup 0, the ticket
  // counter after the flags are cleared.  With sweep_all (k_s2_diag in its
  // full form follows: it writes every slot of every tile with stage 1 and
  // takes the overlap costs from eC) only the slots of tiles without stage
  // 1 are initialised here.
  __shared__ uint32_t lb[2][kMaxBitw];
  if (dag && blockIdx.x == 0 && threadIdx.x == 0) dag[ntiles] = 0;
  for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    if (dag && threadIdx.x == 0) dag[t] = 0;
    const TileDesc td = tiles[t];
    if (td.c_off < 0) {
      // A tile without stage 1 (column 0 < 0): P = 0, C = 0, arrays not
      // read.  Its slots are read only where a corner copy or the sync has
      // written them (an unreached slot, D = +inf, is never read), and its
      // overlap costs not at all (k_s2_diag takes 0).  Its start maps hold
      // (0, 0) only, stored directly (no LDS, no barrier: the workgroup
      // takes this branch as a whole).
      const PairDesc pd = pairs[td.pair];
      const int64_t i = td.ij >> 32, j = td.ij & 0xffffffffll;
      // edges whose every slot may be read
      const bool need0 = i + 1 >= pd.n1 || tiles[t + pd.n2].c_off >= 0;
      const bool need1 = j + 1 >= pd.n2 || tiles[t + 1].c_off >= 0;
      for (int64_t rem = threadIdx.x; rem < 2 * L; rem += blockDim.x) {
        const int e = rem >= L ? 1 : 0;
        const int64_t pos = rem - (int64_t)e * L;
        const int64_t last = e == 0 ? td.W - 1 : td.R - 1;
        if ((e == 0 ? need0 : need1) || pos == 0 || pos == last)
          Darr[t * 2 * L + rem] = __builtin_huge_val();
      }
      for (int w = threadIdx.x; w < 2 * bitw; w += blockDim.x)
        bits[t * 2 * bitw + w] = (w == bitw) ? 1u : 0u;
      continue;
    }
    for (int w = threadIdx.x; w < 2 * bitw; w += blockDim.x) lb[w >= bitw][w - (w >= bitw) * bitw] = 0u;
    __syncthreads();
    for (int64_t rem = threadIdx.x; rem < 2 * L; rem += blockDim.x) {
      const int64_t q = t * 2 * L + rem;
      const int e = rem >= L ? 1 : 0;
      const int64_t pos = rem - (int64_t)e * L;
      // (with dag or sweep_all, the sweep writes every slot of the tile and
      // takes the overlap costs from eC itself)
      const bool init = !dag && !sweep_all;
      if (init) {
        Darr[q] = __builtin_huge_val();
        Larr[q] = __builtin_huge_val();
      }
      const int64_t elen = e == 0 ? td.W : td.R;
      if (pos >= elen) {
        if (init) olap[q] = 0.0;
      } else {
        const int64_t s = eP[q];
        if (init) olap[q] = s > 0 ? eC[td.e_off + s] : eC[td.e_off + L - s];
        if (s > 0) atomicOr(&lb[0][s >> 5], 1u << (s & 31));
        else atomicOr(&lb[1][(-s) >> 5], 1u << ((-s) & 31));
      }
    }
    __syncthreads();
    for (int w = threadIdx.x; w < 2 * bitw; w += blockDim.x) {
      const int m = w >= bitw ? 1 : 0;
      const int64_t k = w - m * bitw;
      bits[(t * 2 + m) * bitw + k] = lb[m][k];
    }
    // lb is cleared for the next tile only after every thread has stored it
    __syncthreads();
  }
}

// Workgroups of k_s2_prep: one per tile (fewer workgroups looping over the
// tiles were slower: 214 against 180 us for the learned pass's 51 831 tiles).
static unsigned prep_grid(int64_t ntiles) {
  const int64_t g = 0x7fffffff;
  return (unsigned)(ntiles < g ? ntiles : g);
}

__device__ __forceinline__ bool bit_at(const uint32_t *bits, int64_t bitw, int64_t t, int map,
                                       int64_t pos) {
  return (bits[(t * 2 + map) * bitw + (pos >> 5)] >> (pos & 31)) & 1u;
}

// What slot (e, pos) of tile (i, j) of pair pd takes in the tile
// anti-diagonal sweep of _nb_initialize_chunks_sparse (Parflex.py:631-760)
// and _nb_dp_fill_chunks_sparse (Parflex.py:763-855), with the valid mask of
// _build_valid_mask (Parflex.py:582-624) evaluated per slot.  slot_plan
// reads only the tile's own edge data, overlap costs and the start bitmaps
// (all final before the sweep); slot_apply then reads the predecessor slot
// (pslot, a slot of tile (i-1, j), (i, j-1) or their corners) and writes.
enum SlotKind : int {
  kSlotNone = 0,   // the slot keeps its value
  kSlotSet = 1,    // D = eDv, L = len
  kSlotChain = 2,  // a path from pslot: if D[pslot] finite, D = (D[pslot] + eDv) - ol,
                   // L = L[pslot] + len
  kSlotCopy = 3,   // corner copy: D, L = D[pslot], L[pslot]
  kSlotCopyFin = 4 // corner copy where D[pslot] is finite
};
struct SlotPlan {
  int64_t slot, pslot;
  double eDv, ol, len;
  int kind;
};

// The overlap cost of a chained slot is olap[slot] (k_s2_prep's table) or,
// with olap null, taken from the tile's start-cell costs eC at e_off as
// k_s2_prep takes it.
__device__ __forceinline__ SlotPlan slot_plan(const PairDesc &pd,
                                              const TileDesc *__restrict__ tiles, int64_t L,
                                              int64_t i, int64_t j, int e, int64_t pos,
                                              const double *__restrict__ eD,
                                              const int32_t *__restrict__ eP,
                                              const double *__restrict__ olap,
                                              const double *__restrict__ eC, int64_t e_off,
                                              const uint32_t *__restrict__ bits, int64_t bitw) {
  SlotPlan sp;
  sp.kind = kSlotNone;
  sp.pslot = 0;
  sp.eDv = sp.ol = sp.len = 0.0;
  const int64_t n1 = pd.n1, n2 = pd.n2;
  const int64_t t = pd.tile_base + i * n2 + j;
  const int64_t slot = (t * 2 + e) * L + pos;
  sp.slot = slot;
  const Geo g = tile_geo(pd, L, i, j);
  const int64_t elen = e == 0 ? g.cols : g.rows;
  if (pos >= elen) return sp;
  // A tile without stage 1 (column 0 < 0) starts no path: every slot that
  // is read, but the corner copies (pos 0), keeps the D = +inf k_s2_prep
  // wrote, and the L of a slot with infinite D is never read (sync, scan and
  // chaining test D first).
  const bool live = tiles[t].c_off >= 0;
  if (!live && pos != 0) return sp;

  bool valid;
  if (e == 0) {
    valid = (i == n1 - 1) || pos == 0 || pos == g.cols - 1 ||
            bit_at(bits, bitw, t + n2, 0, pos);
  } else {
    valid = (j == n2 - 1) || pos == 0 || pos == g.rows - 1 ||
            bit_at(bits, bitw, t + 1, 1, pos);
  }

  const int64_t lr = e == 0 ? g.rows - 1 : pos, lc = e == 0 ? pos : g.cols - 1;
  // a tile without stage 1: D = +inf, P = 0, arrays not read
  const int64_t s = live ? eP[slot] : 0;
  const double eDv = live ? eD[slot] : __builtin_huge_val();
  const int64_t sr = s < 0 ? -s : 0, sc = s > 0 ? s : 0;
  const double seg_len = (double)(iabs64(lr - sr) + iabs64(lc - sc));
  auto set = [&]() {
    sp.kind = kSlotSet;
    sp.eDv = eDv;
    sp.len = seg_len;
  };
  auto copy = [&](int64_t pslot, bool finite_only) {
    sp.kind = finite_only ? kSlotCopyFin : kSlotCopy;
    sp.pslot = pslot;
  };

  // Chain from predecessor tile (pi, pj) as in Parflex.py:683-707/734-758/824-853.
  auto chain = [&](int64_t pi, int64_t pj) {
    const Geo pg = tile_geo(pd, L, pi, pj);
    const int64_t plr = (g.s1 + sr) - pg.s1, plc = (g.s2 + sc) - pg.s2;
    int pedge;
    int64_t ppos, pelen;
    if (plr == pg.rows - 1) {
      pedge = 0;
      ppos = plc;
      pelen = pg.cols;
    } else if (plc == pg.cols - 1) {
      pedge = 1;
      ppos = plr;
      pelen = pg.rows;
    } else {
      return;
    }
    if (ppos < 0 || ppos >= pelen) return;
    sp.kind = kSlotChain;
    sp.pslot = ((pd.tile_base + pi * n2 + pj) * 2 + pedge) * L + ppos;
    sp.eDv = eDv;
    sp.ol = !live ? 0.0 : (olap ? olap[slot] : (s > 0 ? eC[e_off + s] : eC[e_off + L - s]));
    sp.len = seg_len;
  };

  if (i == 0 && j == 0) {
    if (valid && (sr == 0 || sc == 0)) set();
    return sp;
  }
  if (i == 0) {
    if (pos == 0) {
      if (e == 0) {
        const Geo pg = tile_geo(pd, L, 0, j - 1);
        copy(((pd.tile_base + (j - 1)) * 2 + 0) * L + pg.cols - 1, false);
      }
      return sp;
    }
    if (!valid) return sp;
    if (sr == 0) {
      set();
    } else if (sc == 0) {
      chain(0, j - 1);
    }
    return sp;
  }
  if (j == 0) {
    if (pos == 0) {
      if (e == 1) {
        const Geo pg = tile_geo(pd, L, i - 1, 0);
        copy(((pd.tile_base + (i - 1) * n2) * 2 + 1) * L + pg.rows - 1, false);
      }
      return sp;
    }
    if (!valid) return sp;
    if (sc == 0) {
      set();
    } else if (sr == 0) {
      chain(i - 1, 0);
    }
    return sp;
  }
  if (!valid) return sp;
  if (pos == 0) {
    if (e == 0) {
      const Geo pg = tile_geo(pd, L, i, j - 1);
      copy(((t - 1) * 2 + 0) * L + pg.cols - 1, true);
    } else {
      const Geo pg = tile_geo(pd, L, i - 1, j);
      copy(((t - n2) * 2 + 1) * L + pg.rows - 1, true);
    }
    return sp;
  }
  if (sr == 0) {
    chain(i - 1, j);
  } else if (sc == 0) {
    chain(i, j - 1);
  } else {
    set();
  }
  return sp;
}

// Slot reads and writes: plain, or (AGENT) relaxed atomics at agent scope,
// which are coherent across the GPU's L2 caches access by access (the tile
// DAG's slots are read by workgroups on other XCDs within one launch).
template <bool AGENT>
__device__ __forceinline__ double slot_ld(const double *p) {
  if constexpr (AGENT) {
    return __builtin_bit_cast(double, __hip_atomic_load(reinterpret_cast<const uint64_t *>(p),
                                                        __ATOMIC_RELAXED,
                                                        __HIP_MEMORY_SCOPE_AGENT));
  } else {
    return *p;
  }
}
template <bool AGENT>
__device__ __forceinline__ void slot_st(double *p, double v) {
  if constexpr (AGENT) {
    __hip_atomic_store(reinterpret_cast<uint64_t *>(p), __builtin_bit_cast(uint64_t, v),
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  } else {
    *p = v;
  }
}

// Carries out a plan; a slot it writes gets (*dv, *lv).  Returns whether
// it wrote.
template <bool AGENT = false>
__device__ __forceinline__ bool slot_apply(const SlotPlan &sp, double *__restrict__ Darr,
                                           double *__restrict__ Larr, double *dv, double *lv) {
  double d, l;
  switch (sp.kind) {
    case kSlotSet:
      d = sp.eDv;
      l = sp.len;
      break;
    case kSlotChain: {
      const double prev = slot_ld<AGENT>(Darr + sp.pslot);
      if (!isfinite(prev)) return false;
      d = (prev + sp.eDv) - sp.ol;
      l = slot_ld<AGENT>(Larr + sp.pslot) + sp.len;
      break;
    }
    case kSlotCopy:
      d = slot_ld<AGENT>(Darr + sp.pslot);
      l = slot_ld<AGENT>(Larr + sp.pslot);
      break;
    case kSlotCopyFin:
      d = slot_ld<AGENT>(Darr + sp.pslot);
      if (!isfinite(d)) return false;
      l = slot_ld<AGENT>(Larr + sp.pslot);
      break;
    default:
      return false;
  }
  slot_st<AGENT>(Darr + sp.slot, d);
  slot_st<AGENT>(Larr + sp.slot, l);
  *dv = d;
  *lv = l;
  return true;
}

// With olap null (the full form), the overlap costs come from eC and a slot
// of a tile with stage 1 that takes no value is written +inf (k_s2_prep
// initialised only the tiles without stage 1).
__device__ __forceinline__ void diag_slot(const PairDesc &pd, const TileDesc *__restrict__ tiles,
                                          int64_t L, int64_t i, int64_t j, int e, int64_t pos,
                                          const double *__restrict__ eD,
                                          const int32_t *__restrict__ eP,
                                          const double *__restrict__ olap,
                                          const double *__restrict__ eC,
                                          const uint32_t *__restrict__ bits, int64_t bitw,
                                          double *__restrict__ Darr, double *__restrict__ Larr) {
  const TileDesc *td = tiles + pd.tile_base + i * pd.n2 + j;
  const int64_t e_off = olap ? 0 : td->e_off;
  const SlotPlan sp = slot_plan(pd, tiles, L, i, j, e, pos, eD, eP, olap, eC, e_off, bits, bitw);
  double d, l;
  if (!slot_apply(sp, Darr, Larr, &d, &l) && !olap && td->c_off >= 0) {
    Darr[sp.slot] = __builtin_huge_val();
    Larr[sp.slot] = __builtin_huge_val();
  }
}

// One tile anti-diagonal d of every pair: a thread per slot.
__global__ void k_s2_diag(int64_t d, const PairDesc *__restrict__ pairs,
                          const TileDesc *__restrict__ tiles, int64_t L,
                          const double *__restrict__ eD, const int32_t *__restrict__ eP,
                          const double *__restrict__ eC, const uint32_t *__restrict__ bits,
                          int64_t bitw, double *__restrict__ Darr, double *__restrict__ Larr) {
  const PairDesc pd = pairs[blockIdx.y];
  const int64_t n1 = pd.n1, n2 = pd.n2;
  if (d > n1 + n2 - 2) return;
  const int64_t i0 = d - (n2 - 1) > 0 ? d - (n2 - 1) : 0;
  const int64_t i1 = d < n1 - 1 ? d : n1 - 1;
  const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t tq = q / (2 * L);
  if (i0 + tq > i1) return;
  const int64_t rem = q - tq * 2 * L;
  const int e = (int)(rem / L);
  const int64_t pos = rem - (int64_t)e * L;
  const int64_t i = i0 + tq;
  diag_slot(pd, tiles, L, i, d - i, e, pos, eD, eP, nullptr, eC, bits, bitw, Darr, Larr);
}

// The same diagonal from a list of its tiles: `nlive` tiles with stage 1
// (every slot, a thread each) then tiles without (only their slots at
// position 0 take a value, the corner copies: a thread per edge).  list
// holds tile-table indices; the pair and (i, j) come from the table.
__global__ void k_s2_diag_list(const int32_t *__restrict__ list, int64_t nlive, int64_t npruned,
                               const PairDesc *__restrict__ pairs,
                               const TileDesc *__restrict__ tiles, int64_t L,
                               const double *__restrict__ eD, const int32_t *__restrict__ eP,
                               const double *__restrict__ olap, const uint32_t *__restrict__ bits,
                               int64_t bitw, double *__restrict__ Darr, double *__restrict__ Larr) {
  const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  int64_t k, pos;
  int e;
  if (q < nlive * 2 * L) {
    k = q / (2 * L);
    const int64_t rem = q - k * 2 * L;
    e = (int)(rem / L);
    pos = rem - (int64_t)e * L;
  } else {
    const int64_t u = q - nlive * 2 * L;
    if (u >= npruned * 2) return;
    k = nlive + (u >> 1);
    e = (int)(u & 1);
    pos = 0;
  }
  const int64_t t = list[k];
  const TileDesc td = tiles[t];
  const PairDesc pd = pairs[td.pair];
  diag_slot(pd, tiles, L, td.ij >> 32, td.ij & 0xffffffffll, e, pos, eD, eP, olap, nullptr, bits,
            bitw, Darr, Larr);
}

// The same sweep as one launch over the tiles with stage 1 only, each
// after the tiles it reads (a tile DAG instead of a launch per diagonal).
// The workgroups take tickets k = 0, 1, ... from a counter; ticket k is the
// k-th live tile of `list` in diagonal order (diag: per diagonal the first
// list slot and the live count, as pf_select_tiles writes them; the live
// tiles of a diagonal come first in its slots).  Tile (i, j) reads slots of
// (i-1, j) and (i, j-1) and, through the corner copies of pruned tiles,
// (i-1, j-1); it waits for the done flags of those of them that are live.
// Tickets are taken in diagonal order, so every tile a workgroup waits for
// holds a smaller ticket, taken by a running workgroup: the smallest
// unfinished ticket never waits, and every wait ends.
//
// A pruned tile's only written slots are its corner copies (position 0):
// edge 0 copies the edge-0 corner of (i, j-1), edge 1 the edge-1 corner of
// (i-1, j), always in the top tile row / left tile column, elsewhere where
// finite.  A pruned corner is +inf (from k_s2_prep), so the copy changes the
// slot (from +inf) only below or right of a live tile: a live tile pushes
// its corners into the position-0 slots of pruned tiles to its right and
// above when it is done, and pruned tiles take no ticket.  (The L of a slot
// whose D stays +inf is never read.)
//
// Each thread plans up to kDagPer slots before the wait (their loads do not
// depend on other tiles); slots beyond kDagThreads * kDagPer are planned and
// applied after it.  Waits poll at most spin_limit times; one that gives up
// sets g_s2_wait_error, and the tile goes on and publishes its flag, so
// every workgroup still ends.
//
// Ordering: every D / L slot access of this kernel and the done flags are
// relaxed atomics at agent scope (coherent across the XCDs' L2 caches, no
// cache-wide write-back or invalidation); a tile's slot stores have
// completed (s_waitcnt 0 in every wave, then the workgroup barrier) before
// its flag is stored, and a waiting tile reads its neighbours' slots only
// after its thread 0 has seen their flags (then the barrier).
constexpr int kDagThreads = 1024;
constexpr int kDagPer = 4;
constexpr int kDagMaxDiag = 4096;

__device__ int g_s2_wait_error;

__global__ __launch_bounds__(kDagThreads) void k_s2_dag(
    const int32_t *__restrict__ list, const int64_t *__restrict__ diag, int ndiag,
    const PairDesc *__restrict__ pairs, const TileDesc *__restrict__ tiles, int64_t L,
    const double *__restrict__ eD, const int32_t *__restrict__ eP,
    const double *__restrict__ eC, const uint32_t *__restrict__ bits, int64_t bitw,
    double *__restrict__ Darr, double *__restrict__ Larr, int32_t *__restrict__ flags,
    int32_t *__restrict__ ticket, int spin_limit) {
  __shared__ int32_t pre[kDagMaxDiag + 1];
  __shared__ int32_t s_t;
  __shared__ double cD[2], cL[2];  // this tile's corners (edge 0, edge 1) as they end
  const int tid = threadIdx.x;
  // exclusive prefix of the live counts over the diagonals (wave 0)
  if (tid < 64) {
    const int per = (ndiag + 63) / 64;
    const int d0 = tid * per;
    int sum = 0;
    for (int q = 0; q < per; ++q)
      if (d0 + q < ndiag) sum += (int)diag[3 * (d0 + q) + 1];
    int inc = sum;
    for (int o = 1; o < 64; o <<= 1) {
      const int u = __shfl_up(inc, o);
      if (tid >= o) inc += u;
    }
    int run = inc - sum;
    for (int q = 0; q < per; ++q) {
      if (d0 + q < ndiag) {
        pre[d0 + q] = run;
        run += (int)diag[3 * (d0 + q) + 1];
      }
    }
    if (tid == 63) pre[ndiag] = inc;
  }
  __syncthreads();
  const int total = pre[ndiag];
  const int64_t slots = 2 * L;
  for (;;) {
    if (tid == 0) {
      const int k = atomicAdd(ticket, 1);
      int t = -1;
      if (k < total) {
        int lo = 0, hi = ndiag - 1;  // the diagonal with pre[d] <= k < pre[d + 1]
        while (lo < hi) {
          const int mid = (lo + hi + 1) >> 1;
          if (pre[mid] <= k) lo = mid;
          else hi = mid - 1;
        }
        t = list[diag[3 * lo] + (k - pre[lo])];
      }
      s_t = t;
    }
    __syncthreads();
    const int64_t t = __builtin_amdgcn_readfirstlane(s_t);
    if (t < 0) break;
    const TileDesc td = wave_uniform(tiles[t]);
    const PairDesc pd = pairs[td.pair];
    const int64_t i = td.ij >> 32, j = td.ij & 0xffffffffll;
    const Geo g = tile_geo(pd, L, i, j);
    auto plan_at = [&](int64_t idx) {
      const int e = idx >= L ? 1 : 0;
      const int64_t pos = idx - (int64_t)e * L;
      return slot_plan(pd, tiles, L, i, j, e, pos, eD, eP, nullptr, eC, td.e_off, bits, bitw);
    };
    // every slot of the tile is written here: its value, or +inf (D and L)
    // where the sweep leaves k_s2_prep's initial value
    auto apply = [&](const SlotPlan &sp) {
      double d, l;
      if (!slot_apply<true>(sp, Darr, Larr, &d, &l)) {
        d = l = __builtin_huge_val();
        slot_st<true>(Darr + sp.slot, d);
        slot_st<true>(Larr + sp.slot, l);
      }
      const int64_t local = sp.slot - t * slots;  // e*L + pos
      if (local == g.cols - 1) {
        cD[0] = d;
        cL[0] = l;
      }
      if (local == L + g.rows - 1) {
        cD[1] = d;
        cL[1] = l;
      }
    };
    SlotPlan sp[kDagPer];
#pragma unroll
    for (int q = 0; q < kDagPer; ++q) {
      const int64_t idx = (int64_t)q * kDagThreads + tid;
      sp[q].slot = -1;  // no slot
      if (idx < slots) {
        sp[q] = plan_at(idx);
        // slots that take their own edge values, or keep +inf, need no other
        // tile
        if (sp[q].kind == kSlotSet || sp[q].kind == kSlotNone) {
          apply(sp[q]);
          sp[q].slot = -1;
        }
      }
    }
    if (tid == 0) {
      // tiles (i-1, j), (i, j-1), (i-1, j-1)
      for (int q = 0; q < 3; ++q) {
        const bool has = q == 0 ? i > 0 : (q == 1 ? j > 0 : (i > 0 && j > 0));
        if (!has) continue;
        const int64_t p = t - (q == 1 ? 0 : pd.n2) - (q == 0 ? 0 : 1);
        if (tiles[p].c_off < 0) continue;  // pruned: no flag
        for (int n = spin_limit;
             __hip_atomic_load(&flags[p], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0;
             --n) {
          if (n <= 0) {
            g_s2_wait_error = 1;
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kDagPer; ++q)
      if (sp[q].slot >= 0) apply(sp[q]);
    for (int64_t idx = (int64_t)kDagPer * kDagThreads + tid; idx < slots; idx += kDagThreads)
      apply(plan_at(idx));
    __builtin_amdgcn_s_waitcnt(0);  // this wave's slot stores have completed
    __syncthreads();
    if (tid == 0) {
      // corner copies into pruned tiles to the right (edge 0) and above (edge 1)
      if (j + 1 < pd.n2 && tiles[t + 1].c_off < 0 && (i == 0 || isfinite(cD[0]))) {
        const int64_t q = ((t + 1) * 2 + 0) * L;
        slot_st<true>(Darr + q, cD[0]);
        slot_st<true>(Larr + q, cL[0]);
      }
      if (i + 1 < pd.n1 && tiles[t + pd.n2].c_off < 0 && (j == 0 || isfinite(cD[1]))) {
        const int64_t q = ((t + pd.n2) * 2 + 1) * L;
        slot_st<true>(Darr + q, cD[1]);
        slot_st<true>(Larr + q, cL[1]);
      }
      __builtin_amdgcn_s_waitcnt(0);
      __hip_atomic_store(&flags[t], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

// Reads and clears the wait error word of k_s2_dag on the current device.
int take_s2_wait_error(hipStream_t st, int *word) {
  *word = 0;
  PF_CHECK_HIP(hipMemcpyFromSymbolAsync(word, HIP_SYMBOL(g_s2_wait_error), sizeof(int), 0,
                                        hipMemcpyDeviceToHost, st));
  PF_CHECK_HIP(hipStreamSynchronize(st));
  if (*word != 0) {
    const int zero = 0;
    PF_CHECK_HIP(hipMemcpyToSymbolAsync(HIP_SYMBOL(g_s2_wait_error), &zero, sizeof(int), 0,
                                        hipMemcpyHostToDevice, st));
    PF_CHECK_HIP(hipStreamSynchronize(st));
  }
  return PF_OK;
}

// sync_tile_overlap_edges (Parflex.py:919-969), pass `pass` (0, 1, 2).  The
// edge lengths are the tile table's W (edge 0) and R (edge 1), the
// reference's edge_lens; a zero length skips the tile as there.
__global__ void k_s2_sync(int pass, const PairDesc *__restrict__ pairs,
                          const TileDesc *__restrict__ tiles, int64_t L,
                          double *__restrict__ Darr, double *__restrict__ Larr) {
  const PairDesc pd = pairs[blockIdx.y];
  const int64_t n1 = pd.n1, n2 = pd.n2;
  const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (q >= n1 * n2) return;
  const int64_t i = q / n2, j = q - i * n2;
  const int64_t t = pd.tile_base + q;
  const int64_t cols = tiles[t].W, rows = tiles[t].R;
  int64_t a, b;  // slot a is the tile's own, b the other one
  if (pass == 0) {
    if (j + 1 >= n2 || cols == 0) return;
    a = (t * 2 + 0) * L + cols - 1;
    b = ((t + 1) * 2 + 0) * L;
  } else if (pass == 1) {
    if (i + 1 >= n1 || rows == 0) return;
    a = (t * 2 + 1) * L + rows - 1;
    b = ((t + n2) * 2 + 1) * L;
  } else {
    if (cols == 0 || rows == 0) return;
    a = (t * 2 + 0) * L + cols - 1;  // corner seen by the top edge
    b = (t * 2 + 1) * L + rows - 1;  // corner seen by the right edge
  }
  const double da = Darr[a], db = Darr[b];
  const bool fa = isfinite(da), fb = isfinite(db);
  if (fa && !fb) {
    Darr[b] = da;
    Larr[b] = Larr[a];
  } else if (!fa && fb) {
    Darr[a] = db;
    Larr[a] = Larr[b];
  } else if (pass == 2 && fa && fb && fabs(da - db) > 1e-10) {
    Darr[b] = da;
    Larr[b] = Larr[a];
  }
}

// Lexicographic (value, order) minimum; order < 0 means "no candidate".
__device__ __forceinline__ bool better(double v, int64_t o, double bv, int64_t bo) {
  if (o < 0) return false;
  if (bo < 0) return true;
  return v < bv || (v == bv && o < bo);
}

// _nb_scan_edges (Parflex.py:1153-1222) for one (pair, side, segment): the
// owning slot of global top cell (L1-1, g) is tile (n1-1, (g-1)//hop) edge 0
// (tile 0 for g = 0), of right cell (g, L2-1) tile ((g-1)//hop, n2-1) edge 1:
// the closed form of _build_edge_lookups (Parflex.py:1110-1150).
__global__ __launch_bounds__(256) void k_s2_scan(const PairDesc *__restrict__ pairs, int64_t L,
                                                 const double *__restrict__ Darr,
                                                 const double *__restrict__ Larr,
                                                 double *__restrict__ candD,
                                                 double *__restrict__ candL,
                                                 double *__restrict__ segf,
                                                 int64_t *__restrict__ segi) {
  __shared__ double sv[256];
  __shared__ int64_t so[256];
  const PairDesc pd = pairs[blockIdx.z];
  const int side = blockIdx.y;
  const int64_t seg = blockIdx.x;
  if (seg >= pd.max_segs) return;
  const int64_t n = side == 0 ? pd.L2 : pd.L1;
  const int64_t hop = L - 1;
  const int64_t g0 = seg * L;
  const int64_t g1 = g0 + L < n ? g0 + L : n;
  double bv = __builtin_huge_val();
  int64_t bo = -1;
  for (int64_t g = g0 + threadIdx.x; g < g1; g += blockDim.x) {
    const int64_t cslot = pd.lk_off + (side == 0 ? g : pd.L2 + g);
    candD[cslot] = __builtin_nan("");
    candL[cslot] = 0.0;
    if (pd.buf > 0 && g < pd.buf) continue;
    int64_t ci, cj, idx;
    if (side == 0) {
      ci = pd.n1 - 1;
      cj = g == 0 ? 0 : (g - 1) / hop;
      idx = g - cj * hop;
    } else {
      ci = g == 0 ? 0 : (g - 1) / hop;
      cj = pd.n2 - 1;
      idx = g - ci * hop;
    }
    const int64_t slot = ((pd.tile_base + ci * pd.n2 + cj) * 2 + side) * L + idx;
    const double D = Darr[slot], Lv = Larr[slot];
    if (!isfinite(D) || Lv <= 0.0) continue;
    if (side == 0 && D >= 1e9) continue;
    const double norm = D / Lv;
    candD[cslot] = D;
    candL[cslot] = Lv;
    if (better(norm, g, bv, bo)) {
      bv = norm;
      bo = g;
    }
  }
  sv[threadIdx.x] = bv;
  so[threadIdx.x] = bo;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s && better(sv[threadIdx.x + s], so[threadIdx.x + s], sv[threadIdx.x],
                                       so[threadIdx.x])) {
      sv[threadIdx.x] = sv[threadIdx.x + s];
      so[threadIdx.x] = so[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const int64_t k = pd.seg_off + side * pd.max_segs + seg;
    const int64_t g = so[0];
    double *f = segf + k * 4;
    int64_t *ii = segi + k * 8;
    if (g < 0) {
      f[0] = __builtin_huge_val();
      f[1] = __builtin_huge_val();
      f[2] = 0.0;
      f[3] = -1.0;
      for (int z = 0; z < 8; ++z) ii[z] = -1;
      ii[6] = 0;
      return;
    }
    int64_t ci, cj, idx;
    if (side == 0) {
      ci = pd.n1 - 1;
      cj = g == 0 ? 0 : (g - 1) / hop;
      idx = g - cj * hop;
    } else {
      ci = g == 0 ? 0 : (g - 1) / hop;
      cj = pd.n2 - 1;
      idx = g - ci * hop;
    }
    const int64_t slot = ((pd.tile_base + ci * pd.n2 + cj) * 2 + side) * L + idx;
    const int64_t order = side == 0 ? g : pd.L2 + g;
    f[0] = sv[0];
    f[1] = Darr[slot];
    f[2] = Larr[slot];
    f[3] = (double)order;
    ii[0] = ci;
    ii[1] = cj;
    ii[2] = side;
    ii[3] = idx;
    ii[4] = side == 0 ? pd.L1 - 1 : g;
    ii[5] = side == 0 ? g : pd.L2 - 1;
    ii[6] = 1;
    ii[7] = order;
  }
}

// Global winner per pair: the first (in scan order) minimum among segment
// winners below LARGE = 1e9 (Parflex.py:1167-1168).
__global__ __launch_bounds__(256) void k_s2_best(const PairDesc *__restrict__ pairs,
                                                 const double *__restrict__ segf,
                                                 const int64_t *__restrict__ segi,
                                                 double *__restrict__ bestf,
                                                 int64_t *__restrict__ besti) {
  __shared__ double sv[256];
  __shared__ int64_t so[256];
  const PairDesc pd = pairs[blockIdx.x];
  double bv = 1e9;
  int64_t bo = -1;
  for (int64_t k = threadIdx.x; k < 2 * pd.max_segs; k += blockDim.x) {
    const int64_t kk = pd.seg_off + k;
    if (segi[kk * 8 + 6] != 1) continue;
    const double v = segf[kk * 4];
    if (!(v < 1e9)) continue;
    const int64_t o = segi[kk * 8 + 7];
    if (better(v, o, bv, bo)) {
      bv = v;
      bo = kk;  // remember the segment slot; orders are distinct across slots
    }
  }
  // Reduce on (value, scan order).
  sv[threadIdx.x] = bv;
  so[threadIdx.x] = bo;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      const int64_t oa = so[threadIdx.x], ob = so[threadIdx.x + s];
      const int64_t ka = oa < 0 ? -1 : segi[oa * 8 + 7];
      const int64_t kb = ob < 0 ? -1 : segi[ob * 8 + 7];
      if (better(sv[threadIdx.x + s], kb, sv[threadIdx.x], ka)) {
        sv[threadIdx.x] = sv[threadIdx.x + s];
        so[threadIdx.x] = ob;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const int64_t kk = so[0];
    int64_t *o = besti + blockIdx.x * 8;
    if (kk < 0) {
      bestf[blockIdx.x] = 1e9;
      for (int z = 0; z < 8; ++z) o[z] = -1;
      o[6] = 0;
      return;
    }
    bestf[blockIdx.x] = sv[0];
    o[0] = segi[kk * 8 + 4];
    o[1] = segi[kk * 8 + 5];
    o[2] = segi[kk * 8 + 0];
    o[3] = segi[kk * 8 + 1];
    o[4] = segi[kk * 8 + 2];
    o[5] = segi[kk * 8 + 3];
    o[6] = 1;
    o[7] = segi[kk * 8 + 7];
  }
}

// Tile chain of backtrace_and_stitch (Parflex.py:1038-1101).  Request q is
// (pair, ci, cj, et, idx) -- or, with from_best, pair q's winner.  Chain
// entry s: (tile, end_r, end_c, start_r, start_c, s1, s2, npts).  The count
// goes to cnt[q].
__global__ void k_s2_chain(int from_best, const int64_t *__restrict__ req, int64_t nreq,
                           const int64_t *__restrict__ besti, const PairDesc *__restrict__ pairs,
                           const TileDesc *__restrict__ tiles, int64_t L,
                           const int32_t *__restrict__ eP, int64_t *__restrict__ chain,
                           int64_t *__restrict__ cnt) {
  const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (q >= nreq) return;
  int64_t pair, ci, cj, et, idx;
  if (from_best) {
    pair = q;
    if (besti[q * 8 + 6] != 1) {
      cnt[q] = 0;
      return;
    }
    ci = besti[q * 8 + 2];
    cj = besti[q * 8 + 3];
    et = besti[q * 8 + 4];
    idx = besti[q * 8 + 5];
  } else {
    pair = req[q * 5];
    ci = req[q * 5 + 1];
    cj = req[q * 5 + 2];
    et = req[q * 5 + 3];
    idx = req[q * 5 + 4];
  }
  const PairDesc pd = pairs[pair];
  int64_t *ch = chain + q * 101 * 8;
  int64_t ns = 0;
  for (int it = 0; it < 101; ++it) {
    if (ci < 0 || cj < 0 || ci >= pd.n1 || cj >= pd.n2) break;
    const Geo g = tile_geo(pd, L, ci, cj);
    const int64_t t = pd.tile_base + ci * pd.n2 + cj;
    int64_t er, ec;
    if (et == 0) {
      er = g.rows - 1;
      ec = idx;
    } else {
      er = idx;
      ec = g.cols - 1;
    }
    if (er < 0 || ec < 0 || er >= g.rows || ec >= g.cols) break;
    // both loads unconditional (a pruned tile's slot is in bounds; its value
    // is not used), so one memory round trip per tile
    const int64_t Sl = (er == g.rows - 1) ? eP[(t * 2 + 0) * L + ec] : eP[(t * 2 + 1) * L + er];
    const int64_t S = tiles[t].c_off < 0 ? 0 : Sl;
    const int64_t sr = S >= 0 ? 0 : -S, sc = S >= 0 ? S : 0;
    int64_t *en = ch + ns * 8;
    en[0] = t;
    en[1] = er;
    en[2] = ec;
    en[3] = sr;
    en[4] = sc;
    en[5] = g.s1;
    en[6] = g.s2;
    en[7] = 0;
    ++ns;
    const int64_t gsr = g.s1 + sr, gsc = g.s2 + sc;
    if (gsr == 0 || gsc == 0) break;
    const bool corner = (sr == 0 && sc == 0);
    int64_t pi, pj, pe;
    if (corner) {
      pi = ci - 1;
      pj = cj - 1;
      pe = 0;
    } else if (S >= 0) {
      pi = ci - 1;
      pj = cj;
      pe = 0;
    } else {
      pi = ci;
      pj = cj - 1;
      pe = 1;
    }
    if (pi < 0 || pj < 0) break;
    const Geo pg = tile_geo(pd, L, pi, pj);
    int64_t pidx, pmax;
    if (corner) {
      pidx = pg.cols - 1;
      pmax = pg.cols;
    } else {
      pidx = pe == 0 ? gsc - pg.s2 : gsr - pg.s1;
      pmax = pe == 0 ? pg.cols : pg.rows;
    }
    if (pidx < 0 || pidx >= pmax) break;
    ci = pi;
    cj = pj;
    et = pe;
    idx = pidx;
  }
  cnt[q] = ns;
}

// Concatenate the segments of request q, drop the repeated overlap cell at
// each join, reverse (Parflex.py:1060-1062, :1103).  One workgroup per
// request: the segments' lengths and join repeats in parallel, a serial
// prefix over <= 101 values, then every output point in one pass (its
// segment found by bisection over the prefix), 16 bytes a point.
constexpr int kCompactThreads = 1024;
__global__ __launch_bounds__(kCompactThreads) void k_s2_compact(
    const int64_t *__restrict__ chain, int64_t *__restrict__ cnt,
    const int64_t *__restrict__ pts, int64_t L, const PairDesc *__restrict__ pairs,
    int use_pair_off, int64_t path_cap, int64_t *__restrict__ path) {
  __shared__ int64_t off[102];
  __shared__ int64_t skip[101];
  __shared__ int64_t len_s[101];
  __shared__ int64_t total;
  const int64_t q = blockIdx.x;
  const int64_t ns = cnt[q];
  const int64_t *ch = chain + q * 101 * 8;
  const int64_t stride = (2 * L + 8) * 2;
  int64_t cap;
  int64_t *dst;
  if (use_pair_off) {
    const PairDesc pd = pairs[q];
    cap = pd.L1 + pd.L2 + 8;
    dst = path + pd.path_off * 2;
  } else {
    cap = path_cap;
    dst = path + q * path_cap * 2;
  }
  for (int64_t s = threadIdx.x; s < ns; s += blockDim.x) {
    const int64_t n = ch[s * 8 + 7];
    int64_t sk = 0;
    if (s > 0 && n > 0) {
      const int64_t *prev = pts + (q * 101 + s - 1) * stride;
      const int64_t pn = ch[(s - 1) * 8 + 7];
      const int64_t *cur = pts + (q * 101 + s) * stride;
      if (pn > 0 && prev[2 * (pn - 1)] == cur[0] && prev[2 * (pn - 1) + 1] == cur[1]) sk = 1;
    }
    len_s[s] = n;
    skip[s] = sk;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t o = 0;
    for (int64_t s = 0; s < ns; ++s) {
      // a join can only repeat a point once something has been emitted
      if (o == 0) skip[s] = 0;
      off[s] = o;
      o += len_s[s] - skip[s];
    }
    off[ns] = o;
    total = o < cap ? o : cap;
  }
  __syncthreads();
  const int64_t n_out = total;
  typedef long long ll2_t __attribute__((ext_vector_type(2)));
  for (int64_t pos = threadIdx.x; pos < n_out; pos += blockDim.x) {
    // the segment s with off[s] <= pos < off[s + 1] (segments emitting
    // nothing share their offset with the next one: the last such s)
    int64_t lo = 0, hi = ns - 1;
    while (lo < hi) {
      const int64_t mid = (lo + hi + 1) >> 1;
      if (off[mid] <= pos) lo = mid;
      else hi = mid - 1;
    }
    const int64_t k = skip[lo] + (pos - off[lo]);
    const ll2_t v = *reinterpret_cast<const ll2_t *>(pts + (q * 101 + lo) * stride + 2 * k);
    *reinterpret_cast<ll2_t *>(dst + 2 * (n_out - 1 - pos)) = v;
  }
  __syncthreads();
  if (threadIdx.x == 0) cnt[q] = n_out;
}

static inline unsigned nblk(int64_t n, int bs) {
  int64_t b = (n + bs - 1) / bs;
  if (b < 1) b = 1;
  if (b > (1 << 20)) b = 1 << 20;
  return (unsigned)b;
}

}  // namespace pf

using namespace pf;

namespace pf {
// Paths of every pair packed back to back in the (2, n) layout the host
// returns: pair p starts at off_p = 2 * sum_{q<p} len_q, its rows at
// [off_p, off_p + len_p) and its columns at [off_p + len_p, off_p + 2 len_p).
// Grid (point blocks, pairs); every block sums the lengths before its pair.
__global__ __launch_bounds__(256) void k_pack_paths(const PairDesc *__restrict__ pairs,
                                                    const int64_t *__restrict__ path,
                                                    const int64_t *__restrict__ path_len,
                                                    int64_t *__restrict__ out) {
  __shared__ int64_t red[256];
  const int64_t p = blockIdx.y;
  const int64_t n = path_len[p];
  if ((int64_t)blockIdx.x * blockDim.x >= n) return;
  int64_t s = 0;
  for (int64_t q = threadIdx.x; q < p; q += blockDim.x) s += path_len[q];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  const int64_t off = 2 * red[0];
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int64_t *src = path + (pairs[p].path_off + t) * 2;
  out[off + t] = src[0];
  out[off + n + t] = src[1];
}
}  // namespace pf

extern "C" int pf_pack_paths(const int64_t *d_pairs, int64_t npairs, int64_t max_path,
                             const int64_t *d_path, const int64_t *d_path_len, int64_t *d_out,
                             void *stream) {
  PF_REQUIRE(d_pairs && d_path && d_path_len && d_out && npairs >= 1 && max_path >= 0,
             "pf_pack_paths: bad arguments");
  PF_REQUIRE(npairs <= 65535, "pf_pack_paths: too many pairs for one launch");
  if (max_path == 0) return PF_OK;
  hipLaunchKernelGGL(k_pack_paths, dim3((unsigned)((max_path + 255) / 256), (unsigned)npairs),
                     dim3(256), 0, as_stream(stream), reinterpret_cast<const PairDesc *>(d_pairs),
                     d_path, d_path_len, d_out);
  PF_CHECK_HIP(hipGetLastError());
  return PF_OK;
}

extern "C" int pf_fill_f64(double *d_p, double v, int64_t n, void *stream) {
  if (n <= 0) return PF_OK;
  hipLaunchKernelGGL(k_fill_f64, dim3(nblk(n, 256)), dim3(256), 0, as_stream(stream), d_p, v, n);
  PF_CHECK_HIP(hipGetLastError());
  return PF_OK;
}

extern "C" int pf_fill_i64(int64_t *d_p, int64_t v, int64_t n, void *stream) {
  if (n <= 0) return PF_OK;
  hipLaunchKernelGGL(k_fill_i64, dim3(nblk(n, 256)), dim3(256), 0, as_stream(stream), d_p, v, n);
  PF_CHECK_HIP(hipGetLastError());
  return PF_OK;
}

static int stitch_common(hipStream_t st, int from_best, const int64_t *d_req, int64_t nreq,
                         const int64_t *d_besti, const PairDesc *pairs, const TileDesc *tiles,
                         int64_t L, const StepSet &ss, const uint32_t *Bp, const int32_t *eP,
                         int64_t *chain, int64_t *pts, int use_pair_off, int64_t path_cap,
                         int64_t *path, int64_t *cnt) {
  PF_REQUIRE(L >= 2 && L <= 8192, "stitch: tile length must be in [2, 8192]");
  hipLaunchKernelGGL(k_s2_chain, dim3(nblk(nreq, 64)), dim3(64), 0, st, from_best, d_req, nreq,
                     d_besti, pairs, tiles, L, eP, chain, cnt);
  PF_CHECK_HIP(hipGetLastError());
  launch_s2_walk(dim3(101, (unsigned)nreq), st, tiles, L, ss, Bp, chain, cnt, pts);
  PF_CHECK_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_s2_compact, dim3((unsigned)nreq), dim3(kCompactThreads), 0, st, chain, cnt,
                     pts, L,
                     pairs, use_pair_off, path_cap, path);
  PF_CHECK_HIP(hipGetLastError());
  return PF_OK;
}

extern "C" int pf_stage2_phases(int phases, const double *d_edge_C, const int64_t *d_pairs,
                                int64_t npairs, const int64_t *d_tiles, int64_t ntiles, int64_t L,
                                int64_t max_diag, int64_t max_tile_diag, int64_t max_tiles_pair,
                                int64_t max_segs, const int32_t *steps, int K,
                                const uint32_t *d_Bpacked, const double *d_edge_D,
                                const int32_t *d_edge_P, double *d_D_arr, double *d_L_arr,
                                double *d_olap, uint32_t *d_bits, double *d_cand_D,
                                double *d_cand_L, double *d_seg_f, int64_t *d_seg_i,
                                double *d_best_f, int64_t *d_best_i, int64_t *d_chain,
                                int64_t *d_pts, int64_t *d_path, int64_t *d_path_len,
                                void *stream) {
  StepSet ss;
  double w[4] = {1, 1, 1, 1};
  PF_REQUIRE(phases > 0 && (phases & ~(PF_S2_PROPAGATE | PF_S2_SYNC | PF_S2_SCAN |
                                       PF_S2_STITCH)) == 0,
             "pf_stage2_phases: unknown phase bits");
  if (phases & PF_S2_STITCH) {
    int rc = make_stepset(steps, w, K, &ss);
    if (rc) return rc;
  }
  PF_REQUIRE(npairs > 0 && ntiles > 0 && L >= 2, "pf_stage2: empty batch or bad tile length");
  PF_REQUIRE(max_diag >= 1 && max_tile_diag >= 1 && max_tiles_pair >= 1 && max_segs >= 1,
             "pf_stage2: bad batch dimensions");
  PF_REQUIRE(L <= 8192, "pf_stage2: tile length above 8192 is not supported");
  PF_REQUIRE(d_pairs && d_tiles && d_D_arr && d_L_arr, "pf_stage2: null buffer");
  PF_REQUIRE(!(phases & PF_S2_PROPAGATE) || (d_edge_C && d_edge_D && d_edge_P && d_olap && d_bits),
             "pf_stage2: null buffer (propagation)");
  PF_REQUIRE(!(phases & PF_S2_SCAN) ||
                 (d_cand_D && d_cand_L && d_seg_f && d_seg_i && d_best_f && d_best_i),
             "pf_stage2: null buffer (scan)");
  PF_REQUIRE(!(phases & PF_S2_STITCH) ||
                 (d_Bpacked && d_edge_P && d_best_i && d_chain && d_pts && d_path && d_path_len),
             "pf_stage2: null buffer (stitch)");
  PF_REQUIRE(npairs <= 65535 && max_segs <= (1 << 20), "pf_stage2: batch too large");
  PF_REQUIRE(ntiles < (1ll << 31), "pf_stage2: too many tiles");
  hipStream_t st = as_stream(stream);
  const PairDesc *pairs = reinterpret_cast<const PairDesc *>(d_pairs);
  const TileDesc *tiles = reinterpret_cast<const TileDesc *>(d_tiles);
  const int64_t bitw = (L + 32) / 32;
  if (phases & PF_S2_PROPAGATE) {
    // the full form: the diagonal sweep writes every slot and takes the
    // overlap costs from edge C, so prep builds only the start bitmaps
    hipLaunchKernelGGL(k_s2_prep, dim3(prep_grid(ntiles)), dim3(256), 0, st, pairs, tiles, L,
                       d_edge_P, d_edge_C, d_D_arr, d_L_arr, d_olap, d_bits, bitw,
                       (int32_t *)nullptr, ntiles, 1);
    PF_CHECK_HIP(hipGetLastError());
    for (int64_t d = 0; d < max_diag; ++d) {
      const int64_t ntd = (d + 1) < max_tile_diag ? (d + 1) : max_tile_diag;
      const int64_t nth = ntd * 2 * L;
      PF_REQUIRE((nth + 255) / 256 <= (1ll << 31) - 1, "pf_stage2: diagonal too large");
      hipLaunchKernelGGL(k_s2_diag, dim3((unsigned)((nth + 255) / 256), (unsigned)npairs),
                         dim3(256), 0, st, d, pairs, tiles, L, d_edge_D, d_edge_P, d_edge_C,
                         d_bits, bitw, d_D_arr, d_L_arr);
      PF_CHECK_HIP(hipGetLastError());
    }
  }
  if (phases & PF_S2_SYNC) {
    for (int pass = 0; pass < 3; ++pass) {
      hipLaunchKernelGGL(k_s2_sync,
                         dim3((unsigned)((max_tiles_pair + 255) / 256), (unsigned)npairs),
                         dim3(256), 0, st, pass, pairs, tiles, L, d_D_arr, d_L_arr);
      PF_CHECK_HIP(hipGetLastError());
    }
  }
  if (phases & PF_S2_SCAN) {
    hipLaunchKernelGGL(k_s2_scan, dim3((unsigned)max_segs, 2, (unsigned)npairs), dim3(256), 0,
                       st, pairs, L, d_D_arr, d_L_arr, d_cand_D, d_cand_L, d_seg_f, d_seg_i);
    PF_CHECK_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_s2_best, dim3((unsigned)npairs), dim3(256), 0, st, pairs, d_seg_f,
                       d_seg_i, d_best_f, d_best_i);
    PF_CHECK_HIP(hipGetLastError());
  }
  if (phases & PF_S2_STITCH)
    return stitch_common(st, 1, nullptr, npairs, d_best_i, pairs, tiles, L, ss, d_Bpacked,
                         d_edge_P, d_chain, d_pts, 1, 0, d_path, d_path_len);
  return PF_OK;
}

extern "C" int pf_stage2_propagate_listed(const double *d_edge_C, const int64_t *d_pairs,
                                          int64_t npairs, const int64_t *d_tiles, int64_t ntiles,
                                          int64_t L, const int32_t *d_list, const int64_t *h_diag,
                                          int64_t ndiag, const double *d_edge_D,
                                          const int32_t *d_edge_P, double *d_D_arr,
                                          double *d_L_arr, double *d_olap, uint32_t *d_bits,
                                          void *stream) {
  PF_REQUIRE(npairs > 0 && ntiles > 0 && L >= 2 && L <= 8192 && ndiag >= 1 && h_diag,
             "pf_stage2_propagate_listed: bad arguments");
  PF_REQUIRE(d_edge_C && d_pairs && d_tiles && d_list && d_edge_D && d_edge_P && d_D_arr &&
                 d_L_arr && d_olap && d_bits,
             "pf_stage2_propagate_listed: null buffer");
  PF_REQUIRE(ntiles < (1ll << 31), "pf_stage2_propagate_listed: too many tiles");
  int64_t listed = 0;
  for (int64_t d = 0; d < ndiag; ++d) {
    PF_REQUIRE(h_diag[3 * d] == listed && h_diag[3 * d + 1] >= 0 && h_diag[3 * d + 2] >= 0,
               "pf_stage2_propagate_listed: diagonal lists must be contiguous and in order");
    listed += h_diag[3 * d + 1] + h_diag[3 * d + 2];
  }
  PF_REQUIRE(listed == ntiles, "pf_stage2_propagate_listed: every tile must be listed once");
  hipStream_t st = as_stream(stream);
  const PairDesc *pairs = reinterpret_cast<const PairDesc *>(d_pairs);
  const TileDesc *tiles = reinterpret_cast<const TileDesc *>(d_tiles);
  const int64_t bitw = (L + 32) / 32;
  hipLaunchKernelGGL(k_s2_prep, dim3(prep_grid(ntiles)), dim3(256), 0, st, pairs, tiles, L,
                     d_edge_P, d_edge_C, d_D_arr, d_L_arr, d_olap, d_bits, bitw,
                     (int32_t *)nullptr, ntiles, 0);
  PF_CHECK_HIP(hipGetLastError());
  for (int64_t d = 0; d < ndiag; ++d) {
    const int64_t nl = h_diag[3 * d + 1], np = h_diag[3 * d + 2];
    const int64_t nth = nl * 2 * L + np * 2;
    if (nth == 0) continue;
    PF_REQUIRE((nth + 255) / 256 <= (1ll << 31) - 1, "pf_stage2_propagate_listed: diagonal too large");
    hipLaunchKernelGGL(k_s2_diag_list, dim3((unsigned)((nth + 255) / 256)), dim3(256), 0, st,
                       d_list + h_diag[3 * d], nl, np, pairs, tiles, L, d_edge_D, d_edge_P,
                       d_olap, d_bits, bitw, d_D_arr, d_L_arr);
    PF_CHECK_HIP(hipGetLastError());
  }
  return PF_OK;
}

extern "C" int pf_stage2_propagate_dag(const double *d_edge_C, const int64_t *d_pairs,
                                       int64_t npairs, const int64_t *d_tiles, int64_t ntiles,
                                       int64_t L, const int32_t *d_list, const int64_t *d_diag,
                                       int64_t ndiag, const double *d_edge_D,
                                       const int32_t *d_edge_P, double *d_D_arr,
                                       double *d_L_arr, uint32_t *d_bits, int32_t *d_work,
                                       void *stream) {
  PF_REQUIRE(npairs > 0 && ntiles > 0 && L >= 2 && L <= 8192 && ndiag >= 1,
             "pf_stage2_propagate_dag: bad arguments");
  PF_REQUIRE(ndiag <= kDagMaxDiag, "pf_stage2_propagate_dag: more than 4096 tile diagonals");
  PF_REQUIRE(d_edge_C && d_pairs && d_tiles && d_list && d_diag && d_edge_D && d_edge_P &&
                 d_D_arr && d_L_arr && d_bits && d_work,
             "pf_stage2_propagate_dag: null buffer");
  PF_REQUIRE(ntiles < (1ll << 31) - 1, "pf_stage2_propagate_dag: too many tiles");
  hipStream_t st = as_stream(stream);
  const PairDesc *pairs = reinterpret_cast<const PairDesc *>(d_pairs);
  const TileDesc *tiles = reinterpret_cast<const TileDesc *>(d_tiles);
  const int64_t bitw = (L + 32) / 32;
  hipLaunchKernelGGL(k_s2_prep, dim3(prep_grid(ntiles)), dim3(256), 0, st, pairs, tiles, L,
                     d_edge_P, d_edge_C, d_D_arr, d_L_arr, (double *)nullptr, d_bits, bitw,
                     d_work, ntiles, 0);
  PF_CHECK_HIP(hipGetLastError());
  int dev = 0, cus = 0;
  PF_CHECK_HIP(hipGetDevice(&dev));
  PF_CHECK_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  // two resident workgroups per CU; workgroups beyond the live tiles find no
  // ticket and end
  const unsigned grid = (unsigned)(2 * (cus > 0 ? cus : 1));
  hipLaunchKernelGGL(k_s2_dag, dim3(grid), dim3(kDagThreads), 0, st, d_list, d_diag, (int)ndiag,
                     pairs, tiles, L, d_edge_D, d_edge_P, d_edge_C, d_bits, bitw, d_D_arr,
                     d_L_arr, d_work, d_work + ntiles, wait_limit());
  PF_CHECK_HIP(hipGetLastError());
  return PF_OK;
}

extern "C" int pf_stage2(const double *d_edge_C, const int64_t *d_pairs, int64_t npairs,
                         const int64_t *d_tiles, int64_t ntiles, int64_t L, int64_t max_diag,
                         int64_t max_tile_diag, int64_t max_tiles_pair, int64_t max_segs,
                         const int32_t *steps, int K, const uint32_t *d_Bpacked,
                         const double *d_edge_D, const int32_t *d_edge_P, double *d_D_arr,
                         double *d_L_arr, double *d_olap, uint32_t *d_bits, double *d_cand_D,
                         double *d_cand_L, double *d_seg_f, int64_t *d_seg_i, double *d_best_f,
                         int64_t *d_best_i, int64_t *d_chain, int64_t *d_pts, int64_t *d_path,
                         int64_t *d_path_len, void *stream) {
  return pf_stage2_phases(PF_S2_PROPAGATE | PF_S2_SYNC | PF_S2_SCAN | PF_S2_STITCH, d_edge_C,
                          d_pairs, npairs, d_tiles, ntiles, L, max_diag, max_tile_diag,
                          max_tiles_pair, max_segs, steps, K, d_Bpacked, d_edge_D, d_edge_P,
                          d_D_arr, d_L_arr, d_olap, d_bits, d_cand_D, d_cand_L, d_seg_f, d_seg_i,
                          d_best_f, d_best_i, d_chain, d_pts, d_path, d_path_len, stream);
}

extern "C" int pf_stitch(const int64_t *d_req, int64_t nreq, const int64_t *d_pairs,
                         const int64_t *d_tiles, int64_t L, const int32_t *steps, int K,
                         const uint32_t *d_Bpacked, const int32_t *d_edge_P, int64_t *d_chain,
                         int64_t *d_pts, int64_t path_cap, int64_t *d_path, int64_t *d_path_len,
                         void *stream) {
  StepSet ss;
  double w[4] = {1, 1, 1, 1};
  int rc = make_stepset(steps, w, K, &ss);
  if (rc) return rc;
  if (nreq == 0) return PF_OK;
  PF_REQUIRE(nreq > 0 && nreq <= 65535 && L >= 2 && path_cap > 0, "pf_stitch: bad arguments");
  PF_REQUIRE(d_req && d_pairs && d_tiles && d_Bpacked && d_edge_P && d_chain && d_pts &&
                 d_path && d_path_len,
             "pf_stitch: null buffer");
  return stitch_common(as_stream(stream), 0, d_req, nreq, nullptr,
                       reinterpret_cast<const PairDesc *>(d_pairs),
                       reinterpret_cast<const TileDesc *>(d_tiles), L, ss, d_Bpacked, d_edge_P,
                       d_chain, d_pts, 0, path_cap, d_path, d_path_len);
}
