// Tile machinery shared by the pair loops (tsph_pairs.hip, tsph_ale.hip):
// one wavefront per run of `rows` cells of a grid column, persistent waves that
// walk XCD-banded slices of the list of non-empty tiles, the cell-offset table
// of the three neighbour slices in LDS, and the lane-private neighbour lists.
#pragma once

#include "tsph_internal.h"
#include "tsph_physics.h"

namespace tsph {
namespace tiles {

constexpr int kWave = 64;          // one wavefront per tile
constexpr int kCellCap = 11;       // LDS records budgeted per cell of the slices
constexpr int kList = 32;          // per-lane neighbour list entries in LDS
constexpr int kGroup = 8;          // candidates tested together in the mixed pair loop
constexpr int kStageU = 4;         // staged slots per lane whose loads are in flight together
constexpr float kFar32 = 1e30f;    // coordinate of the mixed loop's sentinel slots

// adds the wavefront's counts to its own slots (kWaveCtr0 + 4 w, w: the
// wavefront's index in the launch); no two wavefronts of a launch share a slot
__device__ __forceinline__ void wave_count(unsigned pairs, unsigned cand, unsigned act,
                                           unsigned long long *counters) {
  unsigned long long a = pairs, b = cand, c = act;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    a += __shfl_xor(a, o, 64);
    b += __shfl_xor(b, o, 64);
    c += __shfl_xor(c, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    const int w = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    unsigned long long *p = counters + kWaveCtr0 + 4 * w;
    p[0] += a;
    p[1] += b;
    p[2] += c;
  }
}

// exact reference inclusion test: qq = |dr|/h <= 2 (src/SPH_kernel.h:63).  The
// rounded sqrt and division are only needed within a relative 1e-12 of the
// boundary (their errors are ~1e-16); elsewhere r^2 against 4h^2 decides.
__device__ __forceinline__ bool in_support64(double drx, double dry, double h) {
  const double r2 = (drx * drx) + (dry * dry);
  const double c2 = 4.0 * h * h;
  if (r2 > c2 * (1.0 + 1e-12)) return false;
  if (r2 < c2 * (1.0 - 1e-12)) return true;
  return sqrt(r2) / h <= 2.0;
}

// --------------------------------------------------------------------------
// tiles: one workgroup = one wavefront = one run of `rows` cells of a column
// --------------------------------------------------------------------------
// Dynamic LDS layout: int tab[3 * (kMaxTileRows + 3)] (cell_start of the rows
// of the three neighbour slices), the staged neighbour records, the lane lists.
constexpr int kTabInts = 3 * (kMaxTileRows + 3);
constexpr int kTabBytes = ((kTabInts * 4 + 15) / 16) * 16;

struct Tile {
  int cx, cy0, cy1, row_lo, nrows, w;
  int own0, own1;   // particle range of the tile
  int lq[4];        // LDS start of slice q (lq[3] = records staged)
  int base[3];      // LDS index - global index, per slice
};

// element q of a 3-array held in registers, without dynamic indexing (which
// would place the array in scratch memory)
__device__ __forceinline__ int sel3(const int (&a)[3], int q) {
  return q == 0 ? a[0] : (q == 1 ? a[1] : a[2]);
}

// Each wavefront of a tile kernel works on its own tile with its own LDS area;
// a workgroup holds one wavefront (static walk) or up to kPoolWaves of them
// (pooled walk), which share nothing but the pool's counter and tile order.
// LDS layout: pool_hdr(DYN) bytes (the pooled walk's counter, then kPoolSort
// tile positions), then one area of lds_wave bytes per wavefront.
constexpr int kPoolWaves = 4;
constexpr int kPoolSort = 64;  // chunks of at most this many tiles go in decreasing cost
constexpr int kPoolHdr = 16 + 4 * kPoolSort;
constexpr int kStaticHdr = 16;
constexpr int pool_hdr(bool pooled) { return pooled ? kPoolHdr : kStaticHdr; }

__device__ __forceinline__ int tile_lane() { return threadIdx.x & (kWave - 1); }

template <bool DYN>
__device__ __forceinline__ unsigned char *wave_lds(unsigned char *smem, int lds_wave) {
  return smem + pool_hdr(DYN) + (threadIdx.x >> 6) * lds_wave;
}

// Hand-over between the lanes of one wavefront through LDS (staged records,
// lists): a wavefront's LDS instructions execute in issue order, so only the
// compiler has to be kept from moving accesses across this point.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// Persistent waves over the list of non-empty tiles (tile order).  Workgroup
// b runs on XCD b % 8.  Two walks:
//   * static (DYN = false, one wavefront per workgroup): the list is cut into
//     8 contiguous bands and the waves of XCD x stride through band x, so
//     tiles that are neighbours in the list, which share slices, are staged
//     from the same L2;
//   * pooled (DYN = true): the list is cut into one chunk per workgroup of
//     about equal cost, by the times the tiles took in the previous pass
//     (chunk bounds written by k_cut_chunks), the chunks of XCD x again
//     contiguous; the workgroup's wavefronts take the tiles of their chunk
//     from a counter in LDS, the costliest first (a chunk of up to kPoolSort
//     tiles is sorted by recorded cost when the workgroup starts), so that
//     the tiles left for the end are the cheap ones.  Waves of one compute
//     unit thereby balance each other, and the compute units each other
//     through the costs, without any device-wide atomic (those serialise
//     across the XCDs).
// Either walk records each tile's time (tq_mark) for the next pooled cut.
struct TileQueue {
  int ntl;   // static: tiles in the list
  int band;  // static: band of the wave
  int nwb;   // static: waves per band
  int wb;    // static: this wave's rank in its band
  int j;     // static: round of the current tile
  int lo, hi;   // pooled: the chunk
  int *ctr;     // pooled: the chunk's next tile (LDS)
  int *order;   // pooled, sorted chunk: list positions in decreasing cost (LDS)
  bool sorted;  // pooled: the chunk is taken in the order of `order`
  int *cost; // per-tile costs of the wave's table (tile t at cost[kCostTables t])
  int cur;   // global tile being timed (-1: none)
  int old;   // its recorded cost (lane 0)
  unsigned long long t0;  // its start
};

__device__ __forceinline__ int band_lo(int ntl, int b) { return (int)(((long long)ntl * b) >> 3); }

template <bool DYN>
__device__ __forceinline__ TileQueue tq_begin(int *__restrict__ scal, const int *__restrict__ bound,
                                              const int *__restrict__ tile_list,
                                              unsigned char *smem, int table = 0) {
  TileQueue tq;
  tq.ntl = 0;
  tq.band = blockIdx.x & 7;
  tq.nwb = gridDim.x >> 3;
  tq.wb = blockIdx.x >> 3;
  tq.j = -1;
  tq.lo = tq.hi = 0;
  tq.ctr = reinterpret_cast<int *>(smem);
  tq.order = reinterpret_cast<int *>(smem + 16);
  tq.sorted = false;
  tq.cost = scal + S_NSCAL + table;
  tq.cur = -1;
  tq.old = 0;
  tq.t0 = 0;
  if (!DYN) {
    tq.ntl = scal[S_NTILES];
  } else {
    const int chunk = tq.band * tq.nwb + tq.wb;
    tq.lo = bound[chunk];
    tq.hi = bound[chunk + 1];
    tq.sorted = tq.hi - tq.lo <= kPoolSort;
    if (threadIdx.x == 0) *tq.ctr = tq.lo;
    if (tq.sorted && threadIdx.x < kWave) {
      // bitonic sort of (cost, position) over the first wavefront's lanes,
      // decreasing cost, ties in list order; unused lanes sort last
      const int l = threadIdx.x;
      int key = -1, idx = l;
      if (tq.lo + l < tq.hi) key = tq.cost[kCostTables * tile_list[tq.lo + l]];
      for (int k = 2; k <= kWave; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) {
          const int ok = __shfl_xor(key, j, kWave), oi = __shfl_xor(idx, j, kWave);
          const bool first = ok > key || (ok == key && oi < idx);  // the other ranks first
          const bool keep_first = ((l & k) == 0) == ((l & j) == 0);
          if (keep_first == first) {
            key = ok;
            idx = oi;
          }
        }
      tq.order[l] = tq.lo + idx;
    }
    __syncthreads();
  }
  return tq;
}

// Records the wave's current tile's cost (the time since it began, in 10 ns
// ticks of the device's constant clock, averaged with the tile's recorded
// cost if it has one) and begins global tile `tile` (-1: none, after the
// last tile).
__device__ __forceinline__ void tq_mark(TileQueue &tq, int tile) {
  const unsigned long long now = wall_clock64();
  if (tile_lane() == 0) {
    if (tq.cur >= 0) {
      const unsigned long long u = now - tq.t0;
      int w = u < 1 ? 1 : (u > (1u << 24) ? (1 << 24) : (int)u);
      if (tq.old > 0) w = (tq.old + w + 1) >> 1;
      tq.cost[kCostTables * tq.cur] = w;
    }
    // the new tile's record, read now and used when it ends
    tq.old = tile >= 0 ? tq.cost[kCostTables * tile] : 0;
  }
  tq.cur = tile;
  tq.t0 = now;
}

// pooled walk: lane 0 takes the next tile of the chunk (read with tq_take)
template <bool DYN>
__device__ __forceinline__ int tq_ticket(const TileQueue &tq) {
  int t = 0;
  if (DYN && tile_lane() == 0) t = atomicAdd(tq.ctr, 1);
  return t;
}

// list index of the wave's next tile (raw: the value from tq_ticket), or -1
template <bool DYN>
__device__ __forceinline__ int tq_take(TileQueue &tq, int raw) {
  if (!DYN) {
    const int idx = band_lo(tq.ntl, tq.band) + tq.wb + (++tq.j) * tq.nwb;
    return idx < band_lo(tq.ntl, tq.band + 1) ? idx : -1;
  }
  const int v = __builtin_amdgcn_readfirstlane(raw);
  if (v >= tq.hi) return -1;
  return tq.sorted ? tq.order[v - tq.lo] : v;
}

// Loads the cell offsets of the three slices (one global load per lane, all in
// flight together) and derives the tile ranges.  False for an empty tile.
__device__ __forceinline__ bool tile_begin(const Consts &k, const int *__restrict__ cell_start,
                                           int tile, int rows, int *tab, Tile &t) {
  const int nty = (k.ncy + rows - 1) / rows;
  t.cx = tile / nty;
  t.cy0 = (tile - t.cx * nty) * rows;
  t.cy1 = min(t.cy0 + rows, k.ncy);
  t.row_lo = max(t.cy0 - 1, 0);
  t.nrows = min(t.cy1 + 1, k.ncy) - t.row_lo;
  t.w = t.nrows + 1;
  for (int e = tile_lane(); e < 3 * t.w; e += kWave) {
    const int q = e / t.w, r = e - q * t.w;
    const int col = t.cx + q - 1;
    tab[e] = (col >= 0 && col < k.ncx) ? cell_start[col * k.ncy + t.row_lo + r] : 0;
  }
  wave_sync();
  t.own0 = tab[t.w + (t.cy0 - t.row_lo)];
  t.own1 = tab[t.w + (t.cy1 - t.row_lo)];
  int l = 0;
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const int g0 = tab[q * t.w], g1 = tab[q * t.w + t.nrows];
    t.lq[q] = l;
    t.base[q] = l - g0;
    l += g1 - g0;
  }
  t.lq[3] = l;
  return t.own1 > t.own0;
}

// case of global tile t in a batched ensemble (tile_base: ncases+1 bounds)
__device__ __forceinline__ int case_of_tile(const int *__restrict__ tile_base, int ncases, int t) {
  int lo = 0, hi = ncases;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (tile_base[mid] <= t) lo = mid;
    else hi = mid;
  }
  return lo;
}

// own particles of global tile t (its `rows` cells of one column)
__device__ __forceinline__ int tile_own(const int *__restrict__ cell_start, const Consts &k0,
                                        const Consts *__restrict__ cases,
                                        const int *__restrict__ tile_base, int ncases, int rows,
                                        int t) {
  int lt = t, cell_base = k0.cell_base, ncy = k0.ncy;
  if (cases) {
    const int cse = case_of_tile(tile_base, ncases, t);
    lt = t - tile_base[cse];
    cell_base = cases[cse].cell_base;
    ncy = cases[cse].ncy;
  }
  const int *cs = cell_start + cell_base;
  const int nty = (ncy + rows - 1) / rows;
  const int cx = lt / nty, cy0 = (lt - cx * nty) * rows, cy1 = min(cy0 + rows, ncy);
  return cs[cx * ncy + cy1] - cs[cx * ncy + cy0];
}

// exclusive scan of a over the block (blockDim.x threads, a multiple of 64);
// total: the block's sum.  lds: one slot per wavefront.
template <typename T>
__device__ __forceinline__ T block_exscan(T a, T &total, T *lds) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  T x = a;
  for (int o = 1; o < 64; o <<= 1) {
    const T y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  __syncthreads();
  if (lane == 63) lds[wid] = x;
  __syncthreads();
  T before = 0;
  total = 0;
  for (int w = 0; w < nw; ++w) {
    const T v = lds[w];
    if (w < wid) before += v;
    total += v;
  }
  return before + x - a;
}

// Chunk starts of a pooled walk of `groups` workgroups over the listed tiles,
// cut by cost: listed tile t has the cost interval [W0, W0 + w) (W0: the costs
// of the tiles before it), and chunk g starts at the tile whose interval holds
// g W / groups (W: the total), so that every chunk holds about W / groups.
// The costs come by list position (list_cost, written with the list).  The
// list is cut in segments of kCutSeg entries, one workgroup of kCutBlock
// threads each (cut_segments: workgroups beyond the list exit at once), which
// share nothing: each reads the whole cost list, consecutive threads on
// consecutive entries and kCutBatch loads of each thread in flight together,
// to form the total and the costs of the segments before its own, and keeps
// its own segment in LDS; each thread then scans a contiguous run of it and
// finds the chunk starts there by comparing products (g W against (W0 + w)
// groups) rather than by divisions.
constexpr int kCutBlock = 1024, kCutBatch = 16, kCutSeg = kCutBlock * kCutBatch;

// workgroups that cut a list of at most `tiles` entries
inline int cut_segments(int tiles) { return tiles / kCutSeg + 1; }

// the cut of segment `seg` by one workgroup of kCutBlock threads; wl: kCutSeg
// ints, red: kCutBlock / 64 long longs of LDS
__device__ __forceinline__ void cut_chunks_block(const int *__restrict__ scal,
                                                 const int *__restrict__ list_cost, int groups,
                                                 int *__restrict__ bound, int seg, int *wl,
                                                 long long *red) {
  const int n = scal[S_NTILES];
  const int s0 = seg * kCutSeg;
  if (seg == 0) {
    if (threadIdx.x == 0) bound[groups] = n;
    if (n == 0)
      for (int g = threadIdx.x; g < groups; g += kCutBlock) bound[g] = 0;
  }
  if (s0 >= n) return;
  // the total and the costs before this segment; this segment's costs to LDS
  long long all = 0, before = 0;
  for (int b0 = 0; b0 < n; b0 += kCutSeg) {
    int w[kCutBatch];
#pragma unroll
    for (int u = 0; u < kCutBatch; ++u) {
      const int e = b0 + u * kCutBlock + (int)threadIdx.x;
      w[u] = e < n ? list_cost[e] : 0;
    }
    long long part = 0;
#pragma unroll
    for (int u = 0; u < kCutBatch; ++u) part += w[u];
    all += part;
    if (b0 < s0) before += part;
    if (b0 == s0) {
#pragma unroll
      for (int u = 0; u < kCutBatch; ++u) wl[u * kCutBlock + threadIdx.x] = w[u];
    }
  }
  long long total = 0, carry = 0;
  block_exscan(all, total, red);
  block_exscan(before, carry, red);
  const int m = min(kCutSeg, n - s0);
  const int per = (m + kCutBlock - 1) / kCutBlock;
  const int lo = min((int)threadIdx.x * per, m), hi = min(lo + per, m);
  long long r = 0;
  for (int e = lo; e < hi; ++e) r += wl[e];
  long long segsum = 0;
  long long p = carry + block_exscan(r, segsum, red);
  // first chunk whose start g W / groups >= p
  long long g = (p * groups + total - 1) / total;
  for (int e = lo; e < hi; ++e) {
    p += wl[e];
    const long long end = p * groups;  // chunk g starts in this tile while g W < end
    for (; g < groups && g * total < end; ++g) bound[g] = s0 + e;
  }
}

// The same cut right after the cell-list scan that listed the tiles, from its
// per-block sums: the listed tiles of scan block b sit at list positions
// [tile0[b], tile0[b + 1]) (the last block's end: the list length) and cost
// bsum[b] together.  Workgroup wg (blockDim.x threads, one wavefront per scan
// block, nb blocks) forms the costs before its blocks and the total from the
// nb block sums; each wavefront then reads its block's tile costs 64 at a
// time, scans them across the lanes, and writes the starts of the chunks that
// begin at its tiles.  red, pfx: blockDim.x / 64 long longs of LDS each.
__device__ __forceinline__ void cut_chunks_scan(const int *__restrict__ scal,
                                                const int *__restrict__ list_cost,
                                                const int *__restrict__ bsum,
                                                const int *__restrict__ tile0, int nb, int groups,
                                                int *__restrict__ bound, int wg, long long *red,
                                                long long *pfx) {
  const int n = scal[S_NTILES];
  const int nw = blockDim.x >> 6, wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (wg == 0) {
    if (threadIdx.x == 0) bound[groups] = n;
    if (n == 0)
      for (int g = threadIdx.x; g < groups; g += blockDim.x) bound[g] = 0;
  }
  if (n == 0 || wg * nw >= nb) return;
  long long carry = 0;
  for (int r0 = 0; r0 < nb; r0 += blockDim.x) {
    const int q = r0 + threadIdx.x;
    long long sum = 0;
    const long long ex = block_exscan<long long>(q < nb ? bsum[q] : 0, sum, red) + carry;
    const int qw = q - wg * nw;
    if (qw >= 0 && qw < nw) pfx[qw] = ex;
    carry += sum;
  }
  __syncthreads();
  const long long total = carry;
  const int b = wg * nw + wid;
  if (b >= nb || total <= 0) return;
  long long p = pfx[wid];
  const int t1 = b + 1 < nb ? tile0[b + 1] : n;
  for (int e0 = tile0[b]; e0 < t1; e0 += 64) {
    const int e = e0 + lane;
    const long long w = e < t1 ? list_cost[e] : 0;
    long long x = w;
    for (int o = 1; o < 64; o <<= 1) {
      const long long y = __shfl_up(x, o, 64);
      if (lane >= o) x += y;
    }
    if (e < t1) {
      const long long w0 = p + x - w;  // the costs before tile e
      long long g = (w0 * groups + total - 1) / total;
      const long long end = (w0 + w) * groups;
      for (; g < groups && g * total < end; ++g) bound[g] = e;
    }
    p += __shfl(x, 63, 64);
  }
}

__device__ __forceinline__ int slice_of(const Tile &t, int l) {
  return l >= t.lq[2] ? 2 : (l >= t.lq[1] ? 1 : 0);
}

// --------------------------------------------------------------------------
// Generic tile walks.  A pass supplies a policy P:
//   P::Lds            pointers into the staged-slot area of LDS
//   P::lds(base, cap) carves that area (double2 / float2 arrays first)
//   P::kSlotBytes     bytes per staged slot
//   stage(L, l, j[, ox, oy])  slot l <- particle j (FP32 walk: tile-frame origin)
//   active(i)         whether particle i is active in the pass
//   Own own(i[, ox, oy]); Acc{} accumulators
//   pair_lds(acc, own, L, l[, drx, dry, r2]) / pair_glob(acc, own, j[, ...])
//   finish(acc, own, i)
// walk64 visits neighbours in the reference order (9-cell stencil order,
// ascending index in a cell) and tests support exactly as the reference does;
// walk32 visits the three column slices in memory order with an FP32 test.
// --------------------------------------------------------------------------
template <bool DYN, class P>
__device__ __forceinline__ void walk64(const P &pol, const Consts &k, const int *__restrict__ cell_start,
                                       const int *__restrict__ cellof, int rows,
                                       int *__restrict__ scal, const int *__restrict__ tile_list,
                                       const int *__restrict__ bound, int cap, int lds_wave,
                                       unsigned long long *__restrict__ counters) {
  extern __shared__ __align__(16) unsigned char smem[];
  unsigned char *wl = wave_lds<DYN>(smem, lds_wave);
  int *tab = reinterpret_cast<int *>(wl);
  uint16_t *s_list = reinterpret_cast<uint16_t *>(wl + kTabBytes);
  const typename P::Lds L = P::lds(wl + kTabBytes + sizeof(uint16_t) * kList * kWave, cap);
  const int lane = tile_lane();
  if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&counters[C_PASSES], 1ull);
  unsigned npair = 0, ncand = 0, nact = 0;
  TileQueue tq = tq_begin<DYN>(scal, bound, tile_list, smem);
  for (int it = tq_take<DYN>(tq, tq_ticket<DYN>(tq)); it >= 0;) {
    wave_sync();
    const int ticket = tq_ticket<DYN>(tq);  // read once this tile's loads are waited for
    Tile t;
    tile_begin(k, cell_start, tile_list[it], rows, tab, t);
    tq_mark(tq, tile_list[it]);
    const bool staged = t.lq[3] <= cap;
    if (staged)
      for (int l = lane; l < t.lq[3]; l += kWave) pol.stage(L, l, l - sel3(t.base, slice_of(t, l)));
    wave_sync();
    const int next = tq_take<DYN>(tq, ticket);
    uint16_t *my_list = s_list + lane;
    for (int i0 = t.own0; i0 < t.own1; i0 += kWave) {
      const int i = i0 + lane;
      int li = 0, cy = 0;
      bool act = false;
      if (i < t.own1) {
        li = i + t.base[1];
        act = pol.active(i);
        cy = cellof[i] - t.cx * k.ncy;
      }
      nact += act;
      const typename P::Own o = pol.own(act ? i : -1);
      typename P::Acc acc{};
      if (!staged) {
        if (act) {
          const int c = cellof[i];
          for (int q = 0; q < 9; ++q) {
            int dx, dy;
            stencil_dxy(q, dx, dy);
            if (t.cx + dx < 0 || t.cx + dx >= k.ncx || cy + dy < 0 || cy + dy >= k.ncy) continue;
            const int cl = c + dy + dx * k.ncy;
            for (int j = cell_start[cl]; j < cell_start[cl + 1]; ++j) {
              if (j == i) continue;
              ++ncand;
              const double2 nr = pol.pos_glob(j);
              if (!in_support64(o.r.x - nr.x, o.r.y - nr.y, k.h)) continue;
              ++npair;
              pol.pair_glob(acc, o, j);
            }
          }
        }
      } else {
        // a cell's candidates kGroup at a time, their LDS loads in flight
        // together, appended without branches (every slot of the group is
        // written, the count advances by the test); the exact test of
        // in_support64, with its rounded square root only for the rare
        // candidates within a relative 1e-12 of the boundary.  Slots past the
        // cell's end are masked; they lie in the staged area or the next
        // array of the wave's slots (positions come first).
        const double c2 = 4.0 * k.h * k.h, c2lo = c2 * (1.0 - 1e-12), c2hi = c2 * (1.0 + 1e-12);
        int q = -1, lj = 0, lend = 0;
        bool more = act;
        while (__any(more)) {
          int cnt = 0;
          while (more && cnt <= kList - kGroup) {
            if (lj >= lend) {
              ++q;
              if (q == 9) { more = false; break; }
              int dx, dy;
              stencil_dxy(q, dx, dy);
              const int row = cy + dy;
              if (row < 0 || row >= k.ncy) { lend = lj; continue; }
              const int e = (dx + 1) * t.w + (row - t.row_lo);
              const int bq = sel3(t.base, dx + 1);
              lj = tab[e] + bq;
              lend = tab[e + 1] + bq;
              continue;
            }
            const int m = min(kGroup, lend - lj);
            double2 pp[kGroup];
#pragma unroll
            for (int u = 0; u < kGroup; ++u) pp[u] = pol.pos_lds(L, lj + u);
#pragma unroll
            for (int u = 0; u < kGroup; ++u) {
              const int j = lj + u;
              const bool cand = u < m && j != li;
              ncand += cand;
              const double drx = o.r.x - pp[u].x, dry = o.r.y - pp[u].y;
              const double r2 = (drx * drx) + (dry * dry);
              bool in = cand && r2 < c2lo;
              if (cand && !in && r2 <= c2hi) in = sqrt(r2) / k.h <= 2.0;
              my_list[cnt * kWave] = (uint16_t)j;
              cnt += in;
            }
            lj += m;
          }
          npair += cnt;
          for (int e = 0; e < cnt; ++e) pol.pair_lds(acc, o, L, my_list[e * kWave]);
        }
      }
      if (act) pol.finish(acc, o, i);
    }
    it = next;
  }
  tq_mark(tq, -1);
  wave_count(npair, ncand, nact, counters);
}

template <bool DYN, class P>
__device__ __forceinline__ void walk32(const P &pol, const Consts &k, const int *__restrict__ cell_start,
                                       const int *__restrict__ cellof, int rows,
                                       int *__restrict__ scal, const int *__restrict__ tile_list,
                                       const int *__restrict__ bound, int cap, int lds_wave,
                                       unsigned long long *__restrict__ counters) {
#pragma clang fp contract(fast)
  extern __shared__ __align__(16) unsigned char smem[];
  unsigned char *wl = wave_lds<DYN>(smem, lds_wave);
  int *tab = reinterpret_cast<int *>(wl);
  uint16_t *s_list = reinterpret_cast<uint16_t *>(wl + kTabBytes);
  const typename P::Lds L = P::lds(wl + kTabBytes + sizeof(uint16_t) * kList * kWave, cap);
  const int lane = tile_lane();
  const float r2cut = 4.f * (float)k.h * (float)k.h;
  if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&counters[C_PASSES], 1ull);
  unsigned npair = 0, ncand = 0, nact = 0;
  TileQueue tq = tq_begin<DYN>(scal, bound, tile_list, smem);
  for (int it = tq_take<DYN>(tq, tq_ticket<DYN>(tq)); it >= 0;) {
    wave_sync();
    const int ticket = tq_ticket<DYN>(tq);  // read once this tile's loads are waited for
    Tile t;
    tile_begin(k, cell_start, tile_list[it], rows, tab, t);
    tq_mark(tq, tile_list[it]);
    const bool staged = t.lq[3] <= cap;
    const int mid = t.row_lo + (t.nrows >> 1);
    const double ox = k.x0 - k.of + t.cx * k.kh, oy = k.y0 - k.of + mid * k.kh;
    if (staged)
      for (int l = lane; l < t.lq[3]; l += kWave)
        pol.stage(L, l, l - sel3(t.base, slice_of(t, l)), ox, oy);
    wave_sync();
    const int next = tq_take<DYN>(tq, ticket);
    uint16_t *my_list = s_list + lane;
    for (int i0 = t.own0; i0 < t.own1; i0 += kWave) {
      const int i = i0 + lane;
      int li = 0, ri = 0;
      bool act = false;
      if (i < t.own1) {
        li = i + t.base[1];
        act = pol.active(i);
        ri = cellof[i] - t.cx * k.ncy - t.row_lo;
      }
      nact += act;
      const typename P::Own o = pol.own(act ? i : -1, ox, oy);
      typename P::Acc acc{};
      if (!staged) {
        if (act) {
          const int cc = cellof[i];
          const int cy = cc - t.cx * k.ncy;
          for (int dx = -1; dx <= 1; ++dx) {
            if (t.cx + dx < 0 || t.cx + dx >= k.ncx) continue;
            for (int dy = -1; dy <= 1; ++dy) {
              if (cy + dy < 0 || cy + dy >= k.ncy) continue;
              const int cl = cc + dy + dx * k.ncy;
              for (int j = cell_start[cl]; j < cell_start[cl + 1]; ++j) {
                if (j == i) continue;
                ++ncand;
                const float2 pj = pol.pos_glob(j, ox, oy);
                const float drx = o.p.x - pj.x, dry = o.p.y - pj.y;
                const float r2 = drx * drx + dry * dry;
                if (!(r2 < r2cut)) continue;
                ++npair;
                pol.pair_glob(acc, o, j, drx, dry, r2);
              }
            }
          }
        }
      } else {
        const int rlo = max(ri - 1, 0), rhi = min(ri + 2, t.nrows);
        const int a0 = tab[rlo] + t.base[0], n0 = tab[rhi] + t.base[0] - a0;
        const int a1 = tab[t.w + rlo] + t.base[1], n1 = tab[t.w + rhi] + t.base[1] - a1;
        const int a2 = tab[2 * t.w + rlo] + t.base[2], n2 = tab[2 * t.w + rhi] + t.base[2] - a2;
        const int tot = act ? n0 + n1 + n2 : 0;
        // candidates kGroup at a time from one column slice, their LDS loads
        // in flight together, appended without branches; slots past the
        // slice's range are masked (they lie in the staged area or the next
        // array of the wave's slots, positions coming first)
        const int n01 = n0 + n1;
        int cur = 0;
        while (__any(cur < tot)) {
          int cnt = 0;
          while (cur < tot && cnt <= kList - kGroup) {
            const bool in0 = cur < n0, in01 = cur < n01;
            const int end = in0 ? n0 : (in01 ? n01 : tot);
            const int j0 = cur + (in0 ? a0 : (in01 ? a1 - n0 : a2 - n01));
            const int m = min(kGroup, end - cur);
            float2 pp[kGroup];
#pragma unroll
            for (int u = 0; u < kGroup; ++u) pp[u] = pol.pos_lds(L, j0 + u);
#pragma unroll
            for (int u = 0; u < kGroup; ++u) {
              const int j = j0 + u;
              const bool cand = u < m && j != li;
              ncand += cand;
              const float drx = o.p.x - pp[u].x, dry = o.p.y - pp[u].y;
              my_list[cnt * kWave] = (uint16_t)j;
              cnt += cand && drx * drx + dry * dry < r2cut;
            }
            cur += m;
          }
          npair += cnt;
          for (int e = 0; e < cnt; ++e) {
            const int j = my_list[e * kWave];
            const float2 pj = pol.pos_lds(L, j);
            const float drx = o.p.x - pj.x, dry = o.p.y - pj.y;
            pol.pair_lds(acc, o, L, j, drx, dry, drx * drx + dry * dry);
          }
        }
      }
      if (act) pol.finish(acc, o, i);
    }
    it = next;
  }
  tq_mark(tq, -1);
  wave_count(npair, ncand, nact, counters);
}

// a pair-loop launch, bracketed by HIP events while pair timing is enabled
template <typename K>
int launch_timed(tsph_ctx *c, K kernel_launch) {
  const bool timed = c->time_pairs && c->tev_used + 2 <= (int)c->tev.size();
  if (timed) TSPH_HIP(hipEventRecord(c->tev[c->tev_used], c->stream));
  kernel_launch();
  TSPH_HIP(hipGetLastError());
  if (timed) {
    TSPH_HIP(hipEventRecord(c->tev[c->tev_used + 1], c->stream));
    c->tev_used += 2;
  }
  return 0;
}

}  // namespace tiles
}  // namespace tsph
