// Cell-list build: the device form of Clear_cells + Particle_To_Cell
// (src/SPH_find_pairs.h:45-92).
//
//   1. k_hash     cell id of every particle (reference truncation) and its
//                 arrival rank in the cell; lanes of a wavefront that fall in
//                 the same cell share one atomic increment (the input is in
//                 cell order from the previous build, so a wavefront spans only
//                 a handful of cells); removed particles go to the extra
//                 bucket `ncells`
//   2. scan       exclusive scan of the per-cell counts -> cell_start (CSR);
//                 also lists the non-empty pair-loop tiles
//   3. k_scatter  arrival position = cell_start[cell] + rank, with the seq
//                 of the particle placed there
//   4. k_reorder  each arrival position ranks its seq among those of its cell
//                 (reference particle order), which makes the result identical
//                 to the reference's serial push_back in ascending particle
//                 index and independent of the atomic arrival order, and moves
//                 the SoA fields to the ranked slot; also lists the wall
//                 particles for the wall-density gathers and clears the counts
//
// Steps 1-3 are a one-digit counting sort on the cell key (the key range is
// the cell count, so a single digit covers it and its scan is the cell-start
// table the pair loops need anyway); the rank in step 4 supplies the
// stability.  All passes read the particle count from device memory, so the
// sequence can be captured in a HIP graph.  The per-cell counters are cleared
// by the reorder that follows their scan, so no separate clearing pass is
// needed.  The scan also lists the pair loops' non-empty tiles.
#include "tsph_internal.h"
#include "tsph_physics.h"
#include "tsph_tiles.h"
#include "tsph_walls.h"

namespace tsph {

namespace {

constexpr int kScanBlock = 256;
// cells per thread of the scan: 4 for grids of up to kScanWide cells, 16 above
// (so that grids of up to 4M cells keep at most 1024 blocks: one look-back round)
constexpr int kScanWide = 1 << 20;
inline int scan_items(const tsph_ctx *c) { return c->ncells_tot + 1 <= kScanWide ? 4 : 16; }

// blocks of the cell-list scan (ncells + 1 counts)
inline int scan_blocks(const tsph_ctx *c) {
  const int tile = kScanBlock * scan_items(c);
  return (c->ncells_tot + 1 + tile - 1) / tile;
}

__device__ __forceinline__ unsigned long long lanemask_lt() {
  return (1ull << (threadIdx.x & 63)) - 1ull;
}

// PRED: the cell of the position the symplectic predictor gives (the
// predictor itself is applied by the reorder that follows).  DECIDE: the
// removal decision of RemoveParticlesOutOfDomain (k_rm_decide's) is taken here,
// from the removal states of the corrector that precedes the build.  The
// reference positions of removed particles are marked absent by k_scatter, so
// no decision in this pass reads a state that this pass changes.
// The wall-density pass that precedes a build in the symplectic step (WALLS:
// TSPH_WALL_MDBC or TSPH_WALL_BT; 0: none) runs in the first wb.walls +
// wb.sweep workgroups of the hash launch (64 threads each), beside the hash
// rather than before it: it reads the previous cell list and writes wall
// densities, the hash reads positions and writes the new cell counts.
struct WallPass {
  Soa d;
  const int *cell_start, *nwall, *wall_list;
  int walls, sweep;  // wall and sweep workgroups
  unsigned long long *counters;  // the density count (or nullptr)
};

template <bool PRED, bool DECIDE, int WALLS>
__global__ void k_hash(const double2 *__restrict__ r, const double2 *__restrict__ v,
                       uint8_t *__restrict__ type,
                       const uint32_t *__restrict__ seq, const int *__restrict__ scal, Consts k0,
                       const Consts *__restrict__ cases, const uint8_t *__restrict__ cid, double dt0,
                       int ncells_tot, int *__restrict__ cell_count, int *__restrict__ hcell,
                       int *__restrict__ hrank, uint8_t *__restrict__ seqstate,
                       unsigned long long *__restrict__ counters, WallPass wp) {
  int hb = blockIdx.x;  // hash workgroup
  if (WALLS) {
    const int b = blockIdx.x;
    if (b < wp.walls + wp.sweep) {
      if (WALLS == TSPH_WALL_MDBC)
        walls::mdbc_block(b, wp.walls, wp.sweep, wp.d, wp.cell_start, wp.nwall, wp.wall_list, k0,
                          cases, wp.counters, scal);
      else
        walls::bt_block(b, wp.walls, wp.sweep, wp.d, wp.cell_start, wp.nwall, wp.wall_list, k0,
                        cases, wp.counters, scal);
      return;
    }
    hb = b - wp.walls - wp.sweep;
  }
  const int n = scal[S_NP];
  const int i = hb * blockDim.x + threadIdx.x;
  const bool live = i < n;
  int c = -1;
  if (live) {
    bool removed = type[i] == kRemoved;
    if (DECIDE && !removed && removal_decided(seqstate, seq[i])) {
      removed = true;
      type[i] = kRemoved;
      atomicAdd(&counters[C_REMOVED], 1ull);
    }
    if (removed) {
      c = ncells_tot;
    } else {
      const Consts k = case_consts(k0, cases, cid, i);
      double2 ri = r[i];
      if (PRED) ri = predicted_r(ri, v[i], cases ? k.dt : dt0);
      const int cx = cell_coord(ri.x, k.x0, k.of, k.kh);
      const int cy = cell_coord(ri.y, k.y0, k.of, k.kh);
      c = cy + cx * k.ncy;  // POS(x,y) = y + x*Y, src/SPH_defs.h:36
      if (c < 0 || c >= k.ncells) {
        // undefined behaviour in the reference; counted and clamped here
        atomicAdd(&counters[C_OUTGRID], 1ull);
        c = c < 0 ? 0 : k.ncells - 1;
      }
      c += k.cell_base;  // the case's segment of the global cell range
    }
  }
  // group the lanes by cell; the lowest lane of each group does the atomic
  const int lane = threadIdx.x & 63;
  unsigned long long rem = __ballot(live);
  int leader = 0;
  unsigned long long grp = 0;
  while (rem) {
    const int l = __builtin_ctzll(rem);
    const int key = __shfl(c, l, 64);
    const unsigned long long m = __ballot(c == key) & rem;
    if ((m >> lane) & 1ull) {
      leader = l;
      grp = m;
    }
    rem &= ~m;
  }
  int base = 0;
  if (live && lane == leader) base = atomicAdd(&cell_count[c], __popcll(grp));
  base = __shfl(base, leader, 64);
  if (live) {
    hcell[i] = c;
    hrank[i] = base + __popcll(grp & lanemask_lt());
  }
}

template <int NT>
__device__ __forceinline__ int block_exclusive_scan(int v, int *lds, int &total) {
  constexpr int NW = NT / 64;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) lds[wid] = x;
  __syncthreads();
  int wo = 0, t = 0;
#pragma unroll
  for (int w = 0; w < NW; ++w) {
    const int s = lds[w];
    if (w < wid) wo += s;
    t += s;
  }
  total = t;
  __syncthreads();
  return wo + x - v;
}

// The pair loops' tiles (runs of `rows` cells of a grid column, per case) are
// listed by the scan too: a tile belongs to the scan block that holds its
// first cell.  A block stages its counts in LDS with the counts of the next
// kMaxTileRows cells after it, marks the tiles whose first cell is in the
// block (non-empty: one of their cells holds a particle), and lists them in
// tile order.  The counts are cleared later, by the reorder, so the scan reads
// them intact (the fallback of a stalled block included).
struct TileScan {
  int ncy, ncells, cell_base;  // the single case (cases == nullptr)
  const Consts *cases;
  const int *tile_base;
  int ncases;
  int rows;
  int *list;
  // with cost_in (one cost table: tile t at cost_in[kCostTables t]): each
  // listed tile's cost (or own-particle estimate) at its list position in
  // cost_list, for the chunk cut of a pooled pass
  const int *cost_in;
  int *cost_list;
  CostEstimate cm;  // the estimate for a tile without a recorded cost
  // with cost_list: per scan block, the listed tiles' summed cost and the
  // list position of the first (tiles::cut_chunks_scan)
  int *blk_cost, *blk_tile0;
};

// the case whose cell range holds cell c (batched grids)
__device__ __forceinline__ int case_of(const TileScan &ts, int c) {
  int lo = 0, hi = ts.ncases;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (ts.cases[mid].cell_base <= c) lo = mid;
    else hi = mid;
  }
  return lo;
}

// the grid geometry of one case, in registers: cells [cell_base, next) belong
// to it (next: the following case's first cell)
struct CaseGrid {
  int cse, cell_base, ncy, ncells, tile_base, next;
};

__device__ __forceinline__ CaseGrid case_grid(const TileScan &ts, int cse) {
  if (!ts.cases) return {0, ts.cell_base, ts.ncy, ts.ncells, 0, INT_MAX};
  const Consts &k = ts.cases[cse];
  return {cse, k.cell_base, k.ncy, k.ncells, ts.tile_base[cse],
          cse + 1 < ts.ncases ? ts.cases[cse + 1].cell_base : INT_MAX};
}

// the grid of cell c, starting from g (cells ascending); a reload only at a
// case boundary
__device__ __forceinline__ void advance_grid(const TileScan &ts, int c, CaseGrid &g) {
  if (c < g.next) return;
  int cse = g.cse;
  while (cse + 1 < ts.ncases && ts.cases[cse + 1].cell_base <= c) ++cse;
  g = case_grid(ts, cse);
}

// {tile, cells of the tile} if cell c (in grid g) is a tile's first cell,
// else {-1, 0}
__device__ __forceinline__ int2 tile_at(const TileScan &ts, int c, const CaseGrid &g) {
  const int lc = c - g.cell_base;
  if (lc < 0 || lc >= g.ncells) return make_int2(-1, 0);
  const int cx = lc / g.ncy, cy = lc - cx * g.ncy;
  if (cy % ts.rows) return make_int2(-1, 0);
  const int t = g.tile_base + cx * ((g.ncy + ts.rows - 1) / ts.rows) + cy / ts.rows;
  return make_int2(t, min(ts.rows, g.ncy - cy));
}

// the thread's counts (v) and, for the first kMaxTileRows threads, one count
// of the next block (halo)
template <int IT>
__device__ __forceinline__ void load_counts(const int *__restrict__ in, int len, int block,
                                            int (&v)[IT], int &halo) {
  constexpr int kTile = kScanBlock * IT;
  const int b0 = block * kTile, base = b0 + threadIdx.x * IT;
#pragma unroll
  for (int j = 0; j < IT; ++j) v[j] = (base + j < len) ? in[base + j] : 0;
  const int c = b0 + kTile + threadIdx.x;
  halo = threadIdx.x < kMaxTileRows && c < len ? in[c] : 0;
}

// the block's counts + halo staged in s_cnt; tl[j] >= 0: cell base + j
// starts a non-empty tile; returns the number of such tiles.  On a batched
// grid one thread finds the case of the block's first cell (s_case) and every
// thread steps from there to its own cells.
template <int IT>
__device__ __forceinline__ int stage_counts(int len, int block, const TileScan ts,
                                            const int (&v)[IT], int halo, int (&tl)[IT],
                                            int (&own)[IT], int *s_cnt, int *s_case) {
  constexpr int kTile = kScanBlock * IT;
  const int b0 = block * kTile, base = b0 + threadIdx.x * IT;
#pragma unroll
  for (int j = 0; j < IT; ++j) s_cnt[threadIdx.x * IT + j] = v[j];
  if (threadIdx.x < kMaxTileRows) s_cnt[kTile + threadIdx.x] = halo;
  if (ts.cases && threadIdx.x == 0) *s_case = case_of(ts, b0);
  __syncthreads();
  int nt = 0;
  CaseGrid g = case_grid(ts, ts.cases ? *s_case : 0);
  advance_grid(ts, base, g);
#pragma unroll
  for (int j = 0; j < IT; ++j) {
    advance_grid(ts, base + j, g);
    const int2 tn = base + j < len ? tile_at(ts, base + j, g) : make_int2(-1, 0);
    int o = 0;
    for (int r = 0; r < tn.y; ++r) o += s_cnt[threadIdx.x * IT + j + r];
    tl[j] = o ? tn.x : -1;
    own[j] = o;
    nt += o != 0;
  }
  return nt;
}

// Single-pass scan with decoupled look-back.  Every block publishes its sums
// (cells, tiles) in one 64-bit word (bits 0-30 cells, 31-61 tiles, 62 set
// once the word holds the inclusive prefix instead of the block's own sums,
// 63 set once written).  A block then walks back over the words of the blocks
// before it, 1024 at a time (four per thread), until it meets an inclusive
// prefix, adds up what it passed, and publishes its own inclusive prefix
// (a grid of at most 1024 blocks reads all its words in one round and
// publishes no inclusive prefixes).  Blocks cover 256 IT cells: IT = 16 above
// 1M cells keeps grids of up to 4M cells in one round.
// k_scatter clears the words for the next build.  Blocks are dispatched in
// index order, so a block waits only for blocks that have started, and those
// publish their own sums before they wait themselves.  A block whose wait
// exceeds spin_limit polls (the context's scan_wait; -1: it does not poll)
// forms its sums from the counts of the cells before it instead (and counts
// the stall): a late publisher costs time, never correctness, and every wait
// ends.
constexpr int kLookBack = 4;  // words per thread and look-back round
constexpr unsigned long long kScanField = 0x7fffffffull;
constexpr unsigned long long kScanSet = 1ull << 63, kScanIncl = 1ull << 62;

__device__ __forceinline__ unsigned long long scan_word(int cells, int tiles, bool incl) {
  return kScanSet | (incl ? kScanIncl : 0ull) | ((unsigned long long)tiles << 31) |
         (unsigned long long)cells;
}

// block-wide sums of a and b, and the minimum of m
__device__ __forceinline__ void block_reduce3(int &a, int &b, int &m, int *lds) {
  for (int o = 32; o > 0; o >>= 1) {
    a += __shfl_xor(a, o, 64);
    b += __shfl_xor(b, o, 64);
    m = min(m, __shfl_xor(m, o, 64));
  }
  const int wid = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    lds[3 * wid] = a;
    lds[3 * wid + 1] = b;
    lds[3 * wid + 2] = m;
  }
  __syncthreads();
  a = b = 0;
  m = INT_MAX;
  for (int w = 0; w < kScanBlock / 64; ++w) {
    a += lds[3 * w];
    b += lds[3 * w + 1];
    m = min(m, lds[3 * w + 2]);
  }
  __syncthreads();
}

template <int IT>
__global__ __launch_bounds__(kScanBlock) void k_scan(const int *__restrict__ in, int len,
                                                     unsigned long long *__restrict__ state,
                                                     int *__restrict__ out, int *__restrict__ scal,
                                                     unsigned long long *__restrict__ counters,
                                                     TileScan ts, int spin_limit, int nwall_idx) {
  __shared__ int lds[3 * kScanBlock / 64];
  constexpr int kTile = kScanBlock * IT;
  __shared__ int s_cnt[kTile + kMaxTileRows];
  __shared__ int s_case;
  const int nb = gridDim.x, blk = blockIdx.x;
  int v[IT], tl[IT], own[IT], halo;
  load_counts<IT>(in, len, blk, v, halo);
  const int nt = stage_counts<IT>(len, blk, ts, v, halo, tl, own, s_cnt, &s_case);
  // the listed tiles' recorded costs, in flight during the look-back
  int w[IT];
#pragma unroll
  for (int j = 0; j < IT; ++j) w[j] = ts.cost_in && tl[j] >= 0 ? ts.cost_in[kCostTables * tl[j]] : 0;
  int s = 0;
#pragma unroll
  for (int j = 0; j < IT; ++j) s += v[j];
  int total, tiles;
  int ex = block_exclusive_scan<kScanBlock>(s, lds, total);
  int et = block_exclusive_scan<kScanBlock>(nt, lds, tiles);
  // a grid that one look-back round covers needs no inclusive prefixes
  const bool chain = nb > kLookBack * kScanBlock;
  if (threadIdx.x == 0)
    __hip_atomic_store(&state[blk], scan_word(total, tiles, chain && blk == 0), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  // pc, pt = sums of the blocks before this one.  Look-back: a round reads
  // the words of the kLookBack * kScanBlock nearest blocks not yet covered
  // (entry j = thread + kScanBlock q: block hi - j), all loads in flight
  // together.
  int pc = 0, pt = 0;
  bool stalled = false;
  if (!chain) {
    // one round over all the blocks before this one, farthest first (they
    // published first)
    for (int b = threadIdx.x; b < blk && !stalled; b += kScanBlock) {
      unsigned long long word = 0;
      for (int spin = 0;; ++spin) {
        if (spin > spin_limit) {
          stalled = true;
          break;
        }
        word = __hip_atomic_load(&state[b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (word & kScanSet) break;
        __builtin_amdgcn_s_sleep(1);
      }
      pc += (int)(word & kScanField);
      pt += (int)((word >> 31) & kScanField);
    }
    if (!__syncthreads_or(stalled)) {
      int unused = 0;
      block_reduce3(pc, pt, unused, lds);
    }
  }
  for (int hi = chain ? blk - 1 : -1; hi >= 0;) {
    unsigned long long word[kLookBack];
#pragma unroll
    for (int q = 0; q < kLookBack; ++q) {
      const int b = hi - (int)threadIdx.x - kScanBlock * q;
      // before block 0: an empty inclusive prefix
      word[q] = b < 0 ? kScanSet | kScanIncl
                : spin_limit >= 0 ? __hip_atomic_load(&state[b], __ATOMIC_RELAXED,
                                                      __HIP_MEMORY_SCOPE_AGENT)
                                  : 0ull;
    }
#pragma unroll
    for (int q = 0; q < kLookBack; ++q) {
      const int b = hi - (int)threadIdx.x - kScanBlock * q;
      for (int spin = 1; !(word[q] & kScanSet) && !stalled; ++spin) {
        if (spin > spin_limit) {
          stalled = true;
          break;
        }
        __builtin_amdgcn_s_sleep(1);
        word[q] = __hip_atomic_load(&state[b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
    if (__syncthreads_or(stalled)) break;
    // the nearest inclusive prefix ends the walk; the words up to it count
    int f = INT_MAX, c = 0, t = 0;
    if (chain) {
#pragma unroll
      for (int q = 0; q < kLookBack; ++q)
        if (word[q] & kScanIncl) f = min(f, (int)threadIdx.x + kScanBlock * q);
      block_reduce3(c, t, f, lds);
    }
#pragma unroll
    for (int q = 0; q < kLookBack; ++q)
      if ((int)threadIdx.x + kScanBlock * q <= f) {
        c += (int)(word[q] & kScanField);
        t += (int)((word[q] >> 31) & kScanField);
      }
    int unused = 0;
    block_reduce3(c, t, unused, lds);
    pc += c;
    pt += t;
    if (f != INT_MAX) break;
    hi -= kLookBack * kScanBlock;
  }
  if (__syncthreads_or(stalled)) {
    // the same sums from the counts of the cells before the block
    pc = pt = 0;
    for (int c = threadIdx.x; c < blk * kTile; c += kScanBlock) {
      pc += in[c];
      const int2 tn = tile_at(ts, c, case_grid(ts, ts.cases ? case_of(ts, c) : 0));
      int own = 0;
      for (int r = 0; r < tn.y; ++r) own += in[c + r];
      pt += own != 0;
    }
    int unused = 0;
    block_reduce3(pc, pt, unused, lds);
    if (threadIdx.x == 0) atomicAdd(&counters[C_SCANSTALL], 1ull);
  }
  if (chain && threadIdx.x == 0 && blk > 0)
    __hip_atomic_store(&state[blk], scan_word(pc + total, pt + tiles, true), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  ex += pc;
  et += pt;
  const int base = blk * kTile + threadIdx.x * IT;
  int csum = 0;
#pragma unroll
  for (int j = 0; j < IT; ++j) {
    if (base + j < len) out[base + j] = ex;
    ex += v[j];
    if (tl[j] >= 0) {
      if (ts.cost_list) {
        const int wc = w[j] > 0 ? w[j] : ts.cm.own * own[j] + ts.cm.tile;
        ts.cost_list[et] = wc;
        csum += wc;
      }
      ts.list[et++] = tl[j];
    }
  }
  if (ts.cost_list) {
    int unused0 = 0, unused1 = 0;
    block_reduce3(csum, unused0, unused1, lds);
    if (threadIdx.x == 0) {
      ts.blk_cost[blk] = csum;
      ts.blk_tile0[blk] = pt;
    }
  }
  if (threadIdx.x == 0) {
    if (blk == 0) scal[nwall_idx] = 0;  // the reorder that follows counts the walls
    if (blk == nb - 1) {
      out[len] = pc + total;
      scal[S_NTILES] = pt + tiles;
    }
  }
}

// Arrival position of every particle: cell_start[cell] + its rank from the
// hash.  RANK (the reorder ranks the cells' particles by seq): with the seq of
// the particle placed there and the span of its cell (arrival rank | cell
// count << 16; all ones for a cell of more than 0xffff particles), so that the
// reorder needs no cell offsets; otherwise k_order / k_order_tiles order
// `arrival` (then perm) in place.  Also marks the reference positions of removed particles (the trailing
// bucket) absent in the removal states, and clears k_scan's words for the
// next build.  CUT: the first `cuts` workgroups (one wavefront per scan block)
// cut the new tile list into the chunks of the pooled pair pass that follows
// from the scan's per-block sums (cut_chunks_scan), beside the placement.
constexpr int kScatterBlock = 256;
// particles up to which the reorder ranks the cells' particles itself (on the
// 740k-particle dam break that saves 6 us of a 0.28 ms step; on 4M and 7.8M
// particles the 12 more bytes per particle it moves cost 2-3% of the step)
constexpr int kRankMax = 2000000;
// ordering workgroups (of 256 threads) per compute unit, single case
constexpr int kOrderWaves = 4;
template <bool CUT, bool RANK>
__global__ __launch_bounds__(kScatterBlock) void k_scatter(
    const int *__restrict__ scal, const int *__restrict__ hcell, const int *__restrict__ hrank,
    const int *__restrict__ cell_start, int *__restrict__ arrival,
    uint32_t *__restrict__ arrival_seq, uint32_t *__restrict__ arrival_span,
    int *__restrict__ cellof, int ncells_tot,
    const uint32_t *__restrict__ seq, uint8_t *__restrict__ seqstate,
    unsigned long long *__restrict__ scan_state, int scan_blocks, int groups, int cuts,
    const int *__restrict__ list_cost, const int *__restrict__ blk_cost,
    const int *__restrict__ blk_tile0, int *__restrict__ bound) {
  int blk = blockIdx.x, nblk = gridDim.x;
  if constexpr (CUT) {
    __shared__ long long red[kScatterBlock / 64], pfx[kScatterBlock / 64];
    if (blk < cuts) {
      tiles::cut_chunks_scan(scal, list_cost, blk_cost, blk_tile0, scan_blocks, groups, bound,
                             blk, red, pfx);
      return;
    }
    blk -= cuts;
    nblk -= cuts;
  }
  const int n = scal[S_NP];
  const int i = blk * blockDim.x + threadIdx.x;
  for (int b = i; b < scan_blocks; b += nblk * blockDim.x) scan_state[b] = 0;
  if (i >= n) return;
  const int c = hcell[i];
  if (c == ncells_tot) seqstate[seq[i]] = 0;
  const int s0 = cell_start[c], rank = hrank[i];
  const int pos = s0 + rank;
  arrival[pos] = i;
  if (RANK) {
    const int cnt = cell_start[c + 1] - s0;
    arrival_seq[pos] = seq[i];
    arrival_span[pos] = cnt <= 0xffff ? (uint32_t)rank | (uint32_t)cnt << 16 : 0xffffffffu;
  }
  cellof[pos] = c;
}

// Orders each cell's slots by reference order (seq).  Cells of up to 16
// particles are sorted in registers (all loads in flight together, then an
// odd-even transposition network); larger cells use an insertion sort.
// Also clears the per-cell counts (ncells + 1 of them) for the next build.
// orders cell c's slots by seq
__device__ __forceinline__ void order_cell(const int *__restrict__ cell_start, int c,
                                           const uint32_t *__restrict__ seq,
                                           int *__restrict__ perm) {
  constexpr int R = 16;
  const int s = cell_start[c], e = cell_start[c + 1];
  const int cnt = e - s;
  if (cnt <= 1) return;
  if (cnt <= R) {
    int pv[R];
    uint32_t kv[R];
#pragma unroll
    for (int a = 0; a < R; ++a) {
      pv[a] = a < cnt ? perm[s + a] : 0;
      kv[a] = 0xffffffffu;
    }
#pragma unroll
    for (int a = 0; a < R; ++a)
      if (a < cnt) kv[a] = seq[pv[a]];
    bool sorted = true;
#pragma unroll
    for (int a = 1; a < R; ++a) sorted &= kv[a - 1] <= kv[a];
    if (sorted) return;
#pragma unroll
    for (int pass = 0; pass < R; ++pass) {
#pragma unroll
      for (int a = pass & 1; a + 1 < R; a += 2) {
        if (kv[a] > kv[a + 1]) {
          const uint32_t tk = kv[a]; kv[a] = kv[a + 1]; kv[a + 1] = tk;
          const int tp = pv[a]; pv[a] = pv[a + 1]; pv[a + 1] = tp;
        }
      }
    }
#pragma unroll
    for (int a = 0; a < R; ++a)
      if (a < cnt) perm[s + a] = pv[a];
    return;
  }
  for (int a = s + 1; a < e; ++a) {
    const int pa = perm[a];
    const uint32_t ka = seq[pa];
    int b = a - 1;
    while (b >= s) {
      const int pb = perm[b];
      if (seq[pb] <= ka) break;
      perm[b + 1] = pb;
      --b;
    }
    perm[b + 1] = pa;
  }
}

// per cell of a batched grid: its slots ordered by seq, its count cleared
__global__ void k_order(const int *__restrict__ cell_start, int ncells,
                        const uint32_t *__restrict__ seq, int *__restrict__ perm,
                        int *__restrict__ cell_count) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c <= ncells) cell_count[c] = 0;
  if (c < ncells) order_cell(cell_start, c, seq, perm);
}

// The same over the cells of the listed (non-empty) tiles only, for a single
// case: every occupied cell lies in a listed tile, so the other cells need
// neither ordering nor clearing (their counts are 0).  A bounded grid strides
// over the (tile, row) items.
__global__ void k_order_tiles(const int *__restrict__ cell_start, const int *__restrict__ scal,
                              const int *__restrict__ tile_list, int rows, int ncy, int ncells,
                              const uint32_t *__restrict__ seq, int *__restrict__ perm,
                              int *__restrict__ cell_count) {
  const int ntiles = scal[S_NTILES];
  if (blockIdx.x == 0 && threadIdx.x == 0) cell_count[ncells] = 0;  // the removed bucket
  const int items = ntiles * rows;
  const int nty = (ncy + rows - 1) / rows;
  for (int item = blockIdx.x * blockDim.x + threadIdx.x; item < items;
       item += gridDim.x * blockDim.x) {
    const int it = item / rows, r = item - it * rows;
    const int tile = tile_list[it];
    const int cx = tile / nty, cy = (tile - cx * nty) * rows + r;
    if (cy >= ncy) continue;
    const int c = cx * ncy + cy;
    cell_count[c] = 0;
    order_cell(cell_start, c, seq, perm);
  }
}

// slot of arrival position x in (cell, seq) order: the cell's first slot plus
// the number of the cell's particles that precede x's in seq (ties, which
// distinct seqs never give, broken by arrival position so that every arrival
// position of the cell gets a slot of its own).  The cell's keys are read
// kRankBatch at a time, all of them in flight together (one round for all
// but the fullest cells).
constexpr int kRankBatch = 16;
__device__ __forceinline__ int ranked_slot(const int *__restrict__ cell_start,
                                           const int *__restrict__ cellof,
                                           const uint32_t *__restrict__ span,
                                           const uint32_t *__restrict__ key, int x) {
  const uint32_t sp = span[x];
  int s, e;
  if (sp != 0xffffffffu) {
    s = x - (int)(sp & 0xffffu);
    e = s + (int)(sp >> 16);
  } else {
    const int c = cellof[x];
    s = cell_start[c];
    e = cell_start[c + 1];
  }
  const uint32_t mine = key[x];
  int slot = s;
  for (int a0 = s; a0 < e; a0 += kRankBatch) {
    uint32_t kv[kRankBatch];
#pragma unroll
    for (int u = 0; u < kRankBatch; ++u) kv[u] = a0 + u < e ? key[a0 + u] : 0u;
#pragma unroll
    for (int u = 0; u < kRankBatch; ++u) {
      const int a = a0 + u;
      slot += a < e && (kv[u] < mine || (kv[u] == mine && a < x));
    }
  }
  return slot;
}

// Moves every field from its arrival position to its ranked slot (the state
// after the build is the reference's state, permuted; perm records the
// permutation), lists the wall particles, clears the cell counts for the next
// build and sets the new particle count (removed particles sit after n_new and
// are dropped).
// With EOS the pressure is recomputed as Density_to_pressure does
// (src/SPH_pressure.h:65-89) and the FP32 records of the mixed path are
// written, for the step sequences in which that stage follows the cell build.
// With PRED the symplectic predictor (src/SPH_integration.h:266-287) is
// applied to the moved particles, for the step sequences in which it
// immediately precedes the cell build.  ad_skip (a type mask): a and drho of
// those types are not moved, for the step sequences in which a pair pass that
// overwrites them for exactly those types follows the build.
template <bool EOS, bool PRED, bool RANK>
__global__ void k_reorder(const int *__restrict__ cell_start, int ncells,
                          const int *__restrict__ arrival, const uint32_t *__restrict__ arrival_seq,
                          const uint32_t *__restrict__ arrival_span, int *__restrict__ perm,
                          int *__restrict__ cell_count,
                          Soa src, Soa dst, int *__restrict__ scal, int *__restrict__ wall_list,
                          const int *__restrict__ cellof, Consts k0,
                          const Consts *__restrict__ cases, Mixed mx, double dt0,
                          unsigned ad_skip, unsigned long long *__restrict__ counters,
                          int nwall_idx, bool move_olds) {
  const int n_new = cell_start[ncells];
  const int x = blockIdx.x * blockDim.x + threadIdx.x;  // arrival position
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    scal[S_NP] = n_new;
    if (RANK) cell_count[ncells] = 0;  // the removed bucket
  }
  int wall = 0, k = x;
  unsigned bad = 0;
  if (x < n_new) {
    const int j = arrival[x];
    const int cell = RANK || (EOS && mx.rv) ? cellof[x] : 0;
    if (RANK) {
      k = ranked_slot(cell_start, cellof, arrival_span, arrival_seq, x);
      cell_count[cell] = 0;
      perm[k] = j;
    }
    const uint8_t t = src.type[j];
    wall = t == TSPH_WALL;
    double2 r = src.r[j], v = src.v[j];
    const bool move_ad = !((ad_skip >> t) & 1u);
    double2 a = make_double2(0., 0.);
    double drho = 0.;
    if (PRED || move_ad) {
      a = src.a[j];
      drho = src.drho[j];
    }
    double rho = src.rho[j];
    const Consts kc = EOS || PRED ? case_consts(k0, cases, src.cid, j) : k0;
    if (src.cid) dst.cid[k] = src.cid[j];
    if (PRED) {
      const double dt = cases ? kc.dt : dt0;
      dst.r_o[k] = r;
      dst.v_o[k] = v;
      dst.rho_o[k] = rho;
      predict(r, v, rho, a, drho, dt);
    } else if (move_olds) {
      dst.r_o[k] = src.r_o[j];
      dst.v_o[k] = src.v_o[j];
      dst.rho_o[k] = src.rho_o[j];
    }
    dst.r[k] = r;
    dst.v[k] = v;
    if (move_ad) {
      dst.a[k] = a;
      dst.drho[k] = drho;
    }
    dst.rho[k] = rho;
    dst.type[k] = t;
    dst.seq[k] = src.seq[j];
    dst.index[k] = src.index[j];
    // with the EOS the old pressure is read only where it is kept
    double p;
    if (EOS && t != TSPH_OUTLET && t != TSPH_INLET) {
      bad = rho > 1500. || rho < 0.;
      p = tait64(rho, kc.rho0, kc.cs);
    } else {
      p = src.p[j];
    }
    dst.p[k] = p;
    if (src.gnr) dst.gnr[k] = src.gnr[j];
    double2 nv = make_double2(0., 0.);
    if (src.n) {
      nv = src.n[j];
      dst.n[k] = nv;
    }
    if (EOS && mx.rv)
      write_mixed_record(mx, kc, cell - kc.cell_base, k, r, v, rho, p, src.n != nullptr, nv);
    if (src.omega) {
      dst.omega[k] = src.omega[j];
      dst.omegarho[k] = src.omegarho[j];
      dst.omegav[k] = src.omegav[j];
      dst.omega_o[k] = src.omega_o[j];
      dst.omegarho_o[k] = src.omegarho_o[j];
      dst.omegav_o[k] = src.omegav_o[j];
      dst.domega[k] = src.domega[j];
      dst.domegarho[k] = src.domegarho[j];
      dst.domegav[k] = src.domegav[j];
      if (src.gradrho) {
        dst.gradrho[k] = src.gradrho[j];
        dst.gradvx[k] = src.gradvx[j];
        dst.gradvy[k] = src.gradvy[j];
      }
    }
    if (src.v_oo) {
      dst.v_oo[k] = src.v_oo[j];
      dst.rho_oo[k] = src.rho_oo[j];
      dst.rho_tmp[k] = src.rho_tmp[j];
    }
    if (src.rc) {
      dst.rc[k] = src.rc[j];
      dst.rhoc[k] = src.rhoc[j];
    }
  }
  const unsigned long long m = __ballot(wall);
  if (m) {
    const int lane = threadIdx.x & 63;
    int base = 0;
    if (lane == __builtin_ctzll(m)) base = atomicAdd(&scal[nwall_idx], __popcll(m));
    base = __shfl(base, __builtin_ctzll(m), 64);
    if (wall) wall_list[base + __popcll(m & lanemask_lt())] = k;
  }
  if (EOS) {
    const unsigned long long w = __ballot(bad);
    if (w && (threadIdx.x & 63) == 0) atomicAdd(&counters[C_BADRHO], (unsigned long long)__popcll(w));
  }
}

// r_o, v_o, rho_o of slot k taken from the generation before the last build
// (cur <- alt through perm), for a build that left them behind
__global__ void k_gather_olds(const int *__restrict__ scal, const int *__restrict__ perm, Soa src,
                              Soa dst) {
  const int n = scal[S_NP];
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int j = perm[k];
  dst.r_o[k] = src.r_o[j];
  dst.v_o[k] = src.v_o[j];
  dst.rho_o[k] = src.rho_o[j];
}

// exclusive scan of the cell counts into cell_start, with the tile list of the
// current tile_rows (and, for cost_table >= 0, the listed tiles' costs)
int scan_cells(tsph_ctx *c, hipStream_t s, int *cs, int nwall_idx, int cost_table) {
  const int len = c->ncells_tot + 1;
  const TileScan ts{c->k.ncy, c->k.ncells, c->k.cell_base, c->cases, c->tile_base, c->ncases,
                    c->tile_rows, c->tile_list,
                    cost_table >= 0 ? c->scal + S_NSCAL + cost_table : nullptr,
                    cost_table >= 0 ? c->list_cost : nullptr,
                    cost_estimate(c->prm.precision == TSPH_PREC_MIXED), c->scan_cost,
                    c->scan_tile0};
  if (scan_items(c) == 4)
    hipLaunchKernelGGL(k_scan<4>, dim3(scan_blocks(c)), dim3(kScanBlock), 0, s, c->cell_count,
                       len, c->scan_state, cs, c->scal, c->counters, ts, c->scan_wait, nwall_idx);
  else
    hipLaunchKernelGGL(k_scan<16>, dim3(scan_blocks(c)), dim3(kScanBlock), 0, s, c->cell_count,
                       len, c->scan_state, cs, c->scal, c->counters, ts, c->scan_wait, nwall_idx);
  TSPH_HIP(hipGetLastError());
  c->list_groups = 0;  // a pooled pass cuts the new list
  c->list_cost_table = cost_table;
  return 0;
}

}  // namespace

int ensure_olds(tsph_ctx *c) {
  if (!c->olds_stale) return 0;
  hipLaunchKernelGGL(k_gather_olds, dim3(grid_for(c->n_ub, 256)), dim3(256), 0, c->stream, c->scal,
                     c->perm, c->alt, c->cur);
  TSPH_HIP(hipGetLastError());
  c->olds_stale = false;
  return 0;
}

int launch_cell_list(tsph_ctx *c, bool eos, bool predict, double dt, unsigned ad_skip,
                     bool decide, bool skip_olds, int walls, bool count_density) {
  // a build that moves r_o, v_o, rho_o needs them current (the predictor's
  // build rewrites them instead)
  if (!predict && !skip_olds)
    if (int e = ensure_olds(c)) return e;
  const Consts &k = c->k;
  const int nb = 256;
  const int g = grid_for(c->n_ub, nb);
  hipStream_t s = c->stream;
  const int nct = c->ncells_tot;
  // this build's cell offsets and wall count go to the buffers the current
  // cell list does not use (swapped in at the end)
  int *cs = c->cell_start_next;
  const int wslot = c->nwall_idx == S_NWALL ? S_NWALL_B : S_NWALL;
  WallPass wp{c->cur, c->cell_start, c->scal + c->nwall_idx, c->wall_list, 0, 0,
               count_density ? c->counters : nullptr};
  int hg = g, hbs = nb;
  if (walls) {
    wp.walls = walls::wall_blocks(c);
    wp.sweep = count_density ? walls::sweep_blocks(c->n_ub) : 0;
    hbs = 64;
    hg = wp.walls + wp.sweep + grid_for(c->n_ub, hbs);
    c->records_valid = false;
  }
#define TSPH_HASH(P, D, W)                                                                     \
  hipLaunchKernelGGL((k_hash<P, D, W>), dim3(hg), dim3(hbs), 0, s, c->cur.r, c->cur.v,            \
                     c->cur.type, c->cur.seq, c->scal, k, c->cases, c->cur.cid, dt, nct,          \
                     c->cell_count, c->hcell, c->hrank, c->seqstate, c->counters, wp)
#define TSPH_HASH_W(P, D)                                                                      \
  if (walls == TSPH_WALL_MDBC) TSPH_HASH(P, D, TSPH_WALL_MDBC);                               \
  else if (walls == TSPH_WALL_BT) TSPH_HASH(P, D, TSPH_WALL_BT);                              \
  else TSPH_HASH(P, D, 0)
  if (predict && decide) {
    TSPH_HASH_W(true, true);
  } else if (predict) {
    TSPH_HASH_W(true, false);
  } else if (decide) {
    TSPH_HASH_W(false, true);
  } else {
    TSPH_HASH_W(false, false);
  }
#undef TSPH_HASH_W
#undef TSPH_HASH
  // the pooled pair pass that follows (if any) has its chunks cut in the
  // first workgroups of the scatter, from the costs the scan lists beside the
  // tiles and its per-block sums of them
  const int cut = c->cut_groups;
  c->cut_groups = 0;
  if (cut > c->bound_cap) {
    set_error("tile list: more pooled workgroups than allocated");
    return TSPH_ESTATE;
  }
  if (int e = scan_cells(c, s, cs, wslot, cut > 0 ? c->cut_table : -1)) return e;
  const int sg = grid_for(c->n_ub, kScatterBlock);
  // the per-cell order: ranked by the reorder (one launch fewer), or by its
  // own pass (fewer bytes moved); by size, up to kRankMax particles ranked
  const bool rank = c->cell_order == 1 || (c->cell_order == 0 && c->n_ub <= kRankMax);
  int *arrival = rank ? c->arrival : c->perm;
#define TSPH_SCATTER(CUT, RANK, GRID, CUTS, GROUPS)                                            \
  hipLaunchKernelGGL((k_scatter<CUT, RANK>), dim3(GRID), dim3(kScatterBlock), 0, s, c->scal,       \
                     c->hcell, c->hrank, cs, arrival, c->arrival_seq, c->arrival_span, c->cellof,   \
                     nct, c->cur.seq, c->seqstate, c->scan_state, scan_blocks(c), GROUPS, CUTS,     \
                     c->list_cost, c->scan_cost, c->scan_tile0, c->tile_bound)
  if (cut > 0) {
    // one cutting wavefront per scan block, in the first workgroups
    const int cuts = (scan_blocks(c) + kScatterBlock / 64 - 1) / (kScatterBlock / 64);
    if (rank) TSPH_SCATTER(true, true, cuts + sg, cuts, cut);
    else TSPH_SCATTER(true, false, cuts + sg, cuts, cut);
    c->list_groups = cut;
    c->list_table = c->cut_table;
  } else {
    if (rank) TSPH_SCATTER(false, true, sg, 0, 0);
    else TSPH_SCATTER(false, false, sg, 0, 0);
  }
#undef TSPH_SCATTER
  if (!rank) {
    if (c->cases)
      hipLaunchKernelGGL(k_order, dim3(grid_for(nct + 1, nb)), dim3(nb), 0, s, cs, nct, c->cur.seq,
                         c->perm, c->cell_count);
    else
      hipLaunchKernelGGL(k_order_tiles, dim3(kOrderWaves * std::max(1, c->num_cu)), dim3(nb), 0, s,
                         cs, c->scal, c->tile_list, c->tile_rows, k.ncy, nct, c->cur.seq, c->perm,
                         c->cell_count);
  }
  const Mixed mx = c->prm.precision == TSPH_PREC_MIXED ? c->mx : Mixed{};
#define TSPH_REORDER(E, P, R)                                                                 \
  hipLaunchKernelGGL((k_reorder<E, P, R>), dim3(g), dim3(nb), 0, s, cs, nct, arrival,              \
                     c->arrival_seq, c->arrival_span, c->perm, c->cell_count, c->cur, c->alt,     \
                     c->scal, c->wall_list, c->cellof, k, c->cases, mx, dt, ad_skip, c->counters, \
                     wslot, !skip_olds)
#define TSPH_REORDER_R(E, P)                                                                  \
  if (rank) TSPH_REORDER(E, P, true);                                                         \
  else TSPH_REORDER(E, P, false)
  if (eos && predict) {
    TSPH_REORDER_R(true, true);
  } else if (eos) {
    TSPH_REORDER_R(true, false);
  } else if (predict) {
    TSPH_REORDER_R(false, true);
  } else {
    TSPH_REORDER_R(false, false);
  }
#undef TSPH_REORDER_R
#undef TSPH_REORDER
  TSPH_HIP(hipGetLastError());
  std::swap(c->cur, c->alt);
  std::swap(c->cell_start, c->cell_start_next);
  c->nwall_idx = wslot;
  c->olds_stale = skip_olds && !predict;
  c->parity ^= 1;
  c->cells_valid = true;
  c->tiles_valid = true;  // listed by the scan
  c->records_valid = eos && c->prm.precision == TSPH_PREC_MIXED;
  if (predict || decide) c->marks_valid = false;
  return 0;
}

}  // namespace tsph
