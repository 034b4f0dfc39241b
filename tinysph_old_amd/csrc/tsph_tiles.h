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
// (pooled walk), which share nothing but the pool counter.  LDS layout:
// kPoolHdr bytes (the counter), then one area of lds_wave bytes per wavefront.
constexpr int kPoolWaves = 4;
constexpr int kPoolHdr = 16;

__device__ __forceinline__ int tile_lane() { return threadIdx.x & (kWave - 1); }

__device__ __forceinline__ unsigned char *wave_lds(unsigned char *smem, int lds_wave) {
  return smem + kPoolHdr + (threadIdx.x >> 6) * lds_wave;
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
//     about equal work (own particles plus a per-tile overhead, chunk bounds
//     written by the list build), the chunks of XCD x again contiguous; the
//     workgroup's wavefronts take the tiles of their chunk in order from a
//     counter in LDS.  Waves of one compute unit thereby balance each other
//     without any device-wide atomic.
struct TileQueue {
  int ntl;   // static: tiles in the list
  int band;  // static: band of the wave
  int nwb;   // static: waves per band
  int wb;    // static: this wave's rank in its band
  int j;     // static: round of the current tile
  int hi;    // pooled: end of the chunk
  int *ctr;  // pooled: the chunk's next tile (LDS)
};

__device__ __forceinline__ int band_lo(int ntl, int b) { return (int)(((long long)ntl * b) >> 3); }

template <bool DYN>
__device__ __forceinline__ TileQueue tq_begin(const int *__restrict__ scal, const int *__restrict__ bound,
                                              unsigned char *smem) {
  TileQueue tq;
  tq.ntl = 0;
  tq.band = blockIdx.x & 7;
  tq.nwb = gridDim.x >> 3;
  tq.wb = blockIdx.x >> 3;
  tq.j = -1;
  tq.hi = 0;
  tq.ctr = reinterpret_cast<int *>(smem);
  if (!DYN) {
    tq.ntl = scal[S_NTILES];
  } else {
    const int chunk = tq.band * tq.nwb + tq.wb;
    tq.hi = bound[chunk + 1];
    if (threadIdx.x == 0) *tq.ctr = bound[chunk];
    __syncthreads();
  }
  return tq;
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
  return v < tq.hi ? v : -1;
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
  unsigned char *wl = wave_lds(smem, lds_wave);
  int *tab = reinterpret_cast<int *>(wl);
  uint16_t *s_list = reinterpret_cast<uint16_t *>(wl + kTabBytes);
  const typename P::Lds L = P::lds(wl + kTabBytes + sizeof(uint16_t) * kList * kWave, cap);
  const int lane = tile_lane();
  if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&counters[C_PASSES], 1ull);
  unsigned npair = 0, ncand = 0, nact = 0;
  TileQueue tq = tq_begin<DYN>(scal, bound, smem);
  for (int it = tq_take<DYN>(tq, tq_ticket<DYN>(tq)); it >= 0;) {
    wave_sync();
    const int ticket = tq_ticket<DYN>(tq);  // read once this tile's loads are waited for
    Tile t;
    tile_begin(k, cell_start, tile_list[it], rows, tab, t);
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
  unsigned char *wl = wave_lds(smem, lds_wave);
  int *tab = reinterpret_cast<int *>(wl);
  uint16_t *s_list = reinterpret_cast<uint16_t *>(wl + kTabBytes);
  const typename P::Lds L = P::lds(wl + kTabBytes + sizeof(uint16_t) * kList * kWave, cap);
  const int lane = tile_lane();
  const float r2cut = 4.f * (float)k.h * (float)k.h;
  if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&counters[C_PASSES], 1ull);
  unsigned npair = 0, ncand = 0, nact = 0;
  TileQueue tq = tq_begin<DYN>(scal, bound, smem);
  for (int it = tq_take<DYN>(tq, tq_ticket<DYN>(tq)); it >= 0;) {
    wave_sync();
    const int ticket = tq_ticket<DYN>(tq);  // read once this tile's loads are waited for
    Tile t;
    tile_begin(k, cell_start, tile_list[it], rows, tab, t);
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
