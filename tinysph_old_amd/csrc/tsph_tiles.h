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
constexpr int kMaxTileRows = 24;   // upper bound of tile_rows (LDS offset table)
constexpr int kCellCap = 11;       // LDS records budgeted per cell of the slices
constexpr int kList = 32;          // per-lane neighbour list entries in LDS

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
    if (a) atomicAdd(&counters[C_PAIRS], a);
    if (b) atomicAdd(&counters[C_CAND], b);
    if (c) atomicAdd(&counters[C_ACTIVE], c);
  }
}

// exact reference inclusion test: qq = |dr|/h <= 2 (src/SPH_kernel.h:63)
__device__ __forceinline__ bool in_support64(double drx, double dry, double h) {
  const double r2 = (drx * drx) + (dry * dry);
  if (r2 > 4.0 * h * h * (1.0 + 1e-12)) return false;
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

// Persistent waves: workgroup b runs on XCD b % 8, so the list of non-empty
// tiles (in column order) is cut into 8 contiguous bands and the waves of
// XCD x stride through band x; neighbouring tiles, which share two of their
// three slices, are then staged from the same L2.
__device__ __forceinline__ void band_of_wave(int ntl, int &it, int &end, int &step) {
  const int x = blockIdx.x & 7;
  it = (int)(((long long)ntl * x) >> 3) + (blockIdx.x >> 3);
  end = (int)(((long long)ntl * (x + 1)) >> 3);
  step = gridDim.x >> 3;
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
  for (int e = threadIdx.x; e < 3 * t.w; e += blockDim.x) {
    const int q = e / t.w, r = e - q * t.w;
    const int col = t.cx + q - 1;
    tab[e] = (col >= 0 && col < k.ncx) ? cell_start[col * k.ncy + t.row_lo + r] : 0;
  }
  __syncthreads();
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

__device__ __forceinline__ int slice_of(const Tile &t, int l) {
  return l >= t.lq[2] ? 2 : (l >= t.lq[1] ? 1 : 0);
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
