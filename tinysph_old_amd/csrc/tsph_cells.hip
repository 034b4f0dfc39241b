// Cell-list build: the device form of Clear_cells + Particle_To_Cell
// (src/SPH_find_pairs.h:45-92).
//
//   1. k_hash     cell id of every particle (reference truncation), and its
//                 arrival rank in the cell from one atomic per particle;
//                 removed particles go to the extra bucket `ncells`
//   2. scan       exclusive scan of the per-cell counts -> cell_start (CSR)
//   3. k_scatter  sorted position = cell_start[cell] + rank
//   4. k_order    per cell, order its slots by `seq` (reference particle
//                 order), which makes the result identical to the reference's
//                 serial push_back in ascending particle index and independent
//                 of the atomic arrival order
//   5. k_reorder  gather every SoA field into (cell, seq) order
//
// Steps 1-4 are a one-digit counting sort on the cell key (the key range is
// the cell count, so a single digit covers it and its scan is the cell-start
// table the pair loops need anyway); step 4 supplies the stability.
// All passes read the particle count from device memory, so the sequence can
// be captured in a HIP graph.
#include "tsph_internal.h"
#include "tsph_physics.h"

namespace tsph {

namespace {

constexpr int kScanBlock = 256;
constexpr int kScanItems = 4;
constexpr int kScanTile = kScanBlock * kScanItems;

__global__ void k_fill_i32(int *__restrict__ p, int v, int count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[i] = v;
}

__global__ void k_hash(const double2 *__restrict__ r, const uint8_t *__restrict__ type,
                       const int *__restrict__ scal, Consts k, int *__restrict__ cell_count,
                       int *__restrict__ hcell, int *__restrict__ hrank,
                       unsigned long long *__restrict__ counters) {
  const int n = scal[S_NP];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int c;
  if (type[i] == kRemoved) {
    c = k.ncells;
  } else {
    const double2 ri = r[i];
    const int cx = cell_coord(ri.x, k.x0, k.of, k.kh);
    const int cy = cell_coord(ri.y, k.y0, k.of, k.kh);
    c = cy + cx * k.ncy;  // POS(x,y) = y + x*Y, src/SPH_defs.h:36
    if (c < 0 || c >= k.ncells) {
      // undefined behaviour in the reference; counted and clamped here
      atomicAdd(&counters[C_OUTGRID], 1ull);
      c = c < 0 ? 0 : k.ncells - 1;
    }
  }
  hcell[i] = c;
  hrank[i] = atomicAdd(&cell_count[c], 1);
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

__global__ __launch_bounds__(kScanBlock) void k_scan_partials(const int *__restrict__ in, int len,
                                                              int *__restrict__ partial) {
  __shared__ int lds[kScanBlock / 64];
  const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  int s = 0;
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) s += (base + j < len) ? in[base + j] : 0;
  int total;
  block_exclusive_scan<kScanBlock>(s, lds, total);
  if (threadIdx.x == 0) partial[blockIdx.x] = total;
}

// one block: exclusive scan of the block partials; writes out[len] = total
__global__ __launch_bounds__(1024) void k_scan_top(int *__restrict__ partial, int nb,
                                                   int *__restrict__ out, int len) {
  __shared__ int lds[1024 / 64];
  int carry = 0;
  for (int base = 0; base < nb; base += 1024) {
    const int i = base + threadIdx.x;
    const int v = i < nb ? partial[i] : 0;
    int total;
    const int ex = block_exclusive_scan<1024>(v, lds, total);
    if (i < nb) partial[i] = carry + ex;
    carry += total;
  }
  if (threadIdx.x == 0) out[len] = carry;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_final(const int *__restrict__ in, int len,
                                                           const int *__restrict__ partial,
                                                           int *__restrict__ out) {
  __shared__ int lds[kScanBlock / 64];
  const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  int v[kScanItems];
  int s = 0;
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    v[j] = (base + j < len) ? in[base + j] : 0;
    s += v[j];
  }
  int total;
  int ex = block_exclusive_scan<kScanBlock>(s, lds, total) + partial[blockIdx.x];
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    if (base + j < len) out[base + j] = ex;
    ex += v[j];
  }
}

__global__ void k_scatter(const int *__restrict__ scal, const int *__restrict__ hcell,
                          const int *__restrict__ hrank, const int *__restrict__ cell_start,
                          int *__restrict__ perm, int *__restrict__ cellof) {
  const int n = scal[S_NP];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int c = hcell[i];
  const int pos = cell_start[c] + hrank[i];
  perm[pos] = i;
  cellof[pos] = c;
}

// insertion sort of each cell's slots by reference order (seq)
__global__ void k_order(const int *__restrict__ cell_start, int ncells,
                        const uint32_t *__restrict__ seq, int *__restrict__ perm) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  const int s = cell_start[c], e = cell_start[c + 1];
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

__global__ void k_reorder(const int *__restrict__ cell_start, int ncells,
                          const int *__restrict__ perm, Soa src, Soa dst) {
  const int n_new = cell_start[ncells];  // removed particles sit after this
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n_new) return;
  const int j = perm[k];
  dst.r[k] = src.r[j];
  dst.v[k] = src.v[j];
  dst.a[k] = src.a[j];
  dst.r_o[k] = src.r_o[j];
  dst.v_o[k] = src.v_o[j];
  dst.rho[k] = src.rho[j];
  dst.drho[k] = src.drho[j];
  dst.p[k] = src.p[j];
  dst.rho_o[k] = src.rho_o[j];
  dst.type[k] = src.type[j];
  dst.seq[k] = src.seq[j];
  dst.index[k] = src.index[j];
  if (src.gnr) dst.gnr[k] = src.gnr[j];
  if (src.n) dst.n[k] = src.n[j];
}

__global__ void k_set_np(const int *__restrict__ cell_start, int ncells, int *__restrict__ scal) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    scal[S_NP] = cell_start[ncells];
    scal[S_NTILES] = 0;  // the pair-loop tile list is rebuilt for the new cell list
  }
}

int scan_exclusive(const int *in, int len, int *out, int *partial, hipStream_t s) {
  const int nb = (len + kScanTile - 1) / kScanTile;
  hipLaunchKernelGGL(k_scan_partials, dim3(nb), dim3(kScanBlock), 0, s, in, len, partial);
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, s, partial, nb, out, len);
  hipLaunchKernelGGL(k_scan_final, dim3(nb), dim3(kScanBlock), 0, s, in, len, partial, out);
  TSPH_HIP(hipGetLastError());
  return 0;
}

}  // namespace

int launch_fill_i32(int *p, int value, int count, hipStream_t s) {
  if (count <= 0) return 0;
  hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(count, 256)), dim3(256), 0, s, p, value, count);
  TSPH_HIP(hipGetLastError());
  return 0;
}

int scan_partials_needed(int len) { return (len + kScanTile - 1) / kScanTile + 1; }

int launch_cell_list(tsph_ctx *c) {
  const Consts &k = c->k;
  const int nb = 256;
  const int g = grid_for(c->n_ub, nb);
  hipStream_t s = c->stream;
  if (int e = launch_fill_i32(c->cell_count, 0, k.ncells + 1, s)) return e;
  hipLaunchKernelGGL(k_hash, dim3(g), dim3(nb), 0, s, c->cur.r, c->cur.type, c->scal, k,
                     c->cell_count, c->hcell, c->hrank, c->counters);
  if (int e = scan_exclusive(c->cell_count, k.ncells + 1, c->cell_start, c->scan_tmp, s)) return e;
  hipLaunchKernelGGL(k_scatter, dim3(g), dim3(nb), 0, s, c->scal, c->hcell, c->hrank,
                     c->cell_start, c->perm, c->cellof);
  hipLaunchKernelGGL(k_order, dim3(grid_for(k.ncells, nb)), dim3(nb), 0, s, c->cell_start,
                     k.ncells, c->cur.seq, c->perm);
  hipLaunchKernelGGL(k_reorder, dim3(g), dim3(nb), 0, s, c->cell_start, k.ncells, c->perm,
                     c->cur, c->alt);
  hipLaunchKernelGGL(k_set_np, dim3(1), dim3(64), 0, s, c->cell_start, k.ncells, c->scal);
  TSPH_HIP(hipGetLastError());
  std::swap(c->cur, c->alt);
  c->cells_valid = true;
  c->tiles_valid = false;
  c->tiles_zeroed = true;
  return 0;
}

}  // namespace tsph
