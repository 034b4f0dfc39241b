// Cell-list build: the device form of Clear_cells + Particle_To_Cell
// (src/SPH_find_pairs.h:45-92).
//
//   1. k_hash     cell id of every particle (reference truncation) and its
//                 arrival rank in the cell; lanes of a wavefront that fall in
//                 the same cell share one atomic increment (the input is in
//                 cell order from the previous build, so a wavefront spans only
//                 a handful of cells); removed particles go to the extra
//                 bucket `ncells`
//   2. scan       exclusive scan of the per-cell counts -> cell_start (CSR)
//   3. k_scatter  sorted position = cell_start[cell] + rank
//   4. k_order    per cell, order its slots by `seq` (reference particle
//                 order), which makes the result identical to the reference's
//                 serial push_back in ascending particle index and independent
//                 of the atomic arrival order
//   5. k_reorder  gather the SoA fields into (cell, seq) order; also lists the
//                 wall particles for the wall-density gathers
//
// Steps 1-4 are a one-digit counting sort on the cell key (the key range is
// the cell count, so a single digit covers it and its scan is the cell-start
// table the pair loops need anyway); step 4 supplies the stability.
// All passes read the particle count from device memory, so the sequence can
// be captured in a HIP graph.  The per-cell counters are cleared by the scan
// that consumes them, so no separate clearing pass is needed.
#include "tsph_internal.h"
#include "tsph_physics.h"

namespace tsph {

namespace {

constexpr int kScanBlock = 256;
constexpr int kScanItems = 4;
constexpr int kScanTile = kScanBlock * kScanItems;

__device__ __forceinline__ unsigned long long lanemask_lt() {
  return (1ull << (threadIdx.x & 63)) - 1ull;
}

__global__ void k_hash(const double2 *__restrict__ r, const uint8_t *__restrict__ type,
                       const uint32_t *__restrict__ seq, const int *__restrict__ scal, Consts k0,
                       const Consts *__restrict__ cases, const uint8_t *__restrict__ cid,
                       int ncells_tot, int *__restrict__ cell_count, int *__restrict__ hcell,
                       int *__restrict__ hrank, uint8_t *__restrict__ seqstate,
                       unsigned long long *__restrict__ counters) {
  const int n = scal[S_NP];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = i < n;
  int c = -1;
  if (live) {
    if (type[i] == kRemoved) {
      c = ncells_tot;
      seqstate[seq[i]] = 0;  // its reference position is vacated
    } else {
      const Consts k = case_consts(k0, cases, cid, i);
      const double2 ri = r[i];
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

// one block: exclusive scan of the block partials; writes out[len] = total and
// clears the list counters that the reorder and pair passes rebuild
__global__ __launch_bounds__(1024) void k_scan_top(int *__restrict__ partial, int nb,
                                                   int *__restrict__ out, int len,
                                                   int *__restrict__ scal) {
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
  if (threadIdx.x == 0) {
    out[len] = carry;
    scal[S_NTILES] = 0;
    scal[S_NWALL] = 0;
  }
}

// final pass; also clears the counts it consumed for the next build
__global__ __launch_bounds__(kScanBlock) void k_scan_final(int *__restrict__ in, int len,
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
    if (base + j < len) {
      out[base + j] = ex;
      in[base + j] = 0;
    }
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

// Orders each cell's slots by reference order (seq).  Cells of up to 16
// particles are sorted in registers (all loads in flight together, then an
// odd-even transposition network); larger cells use an insertion sort.
__global__ void k_order(const int *__restrict__ cell_start, int ncells,
                        const uint32_t *__restrict__ seq, int *__restrict__ perm) {
  constexpr int R = 16;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
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

// Gathers every field into sorted order (the state after the build is the
// reference's state, permuted), lists the wall particles, and sets the new
// particle count (removed particles sit after n_new and are dropped).
// With EOS the pressure is recomputed as Density_to_pressure does
// (src/SPH_pressure.h:65-89) and the FP32 records of the mixed path are
// written, for the step sequences in which that stage follows the cell build.
template <bool EOS>
__global__ void k_reorder(const int *__restrict__ cell_start, int ncells, const int *__restrict__ perm,
                          Soa src, Soa dst, int *__restrict__ scal, int *__restrict__ wall_list,
                          const int *__restrict__ cellof, Consts k0,
                          const Consts *__restrict__ cases, Mixed mx,
                          unsigned long long *__restrict__ counters) {
  const int n_new = cell_start[ncells];
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (blockIdx.x == 0 && threadIdx.x == 0) scal[S_NP] = n_new;
  int wall = 0;
  unsigned bad = 0;
  if (k < n_new) {
    const int j = perm[k];
    const uint8_t t = src.type[j];
    wall = t == TSPH_WALL;
    const double2 r = src.r[j], v = src.v[j];
    const double rho = src.rho[j];
    if (src.cid) dst.cid[k] = src.cid[j];
    dst.r[k] = r;
    dst.v[k] = v;
    dst.a[k] = src.a[j];
    dst.rho[k] = rho;
    dst.drho[k] = src.drho[j];
    dst.type[k] = t;
    dst.seq[k] = src.seq[j];
    dst.index[k] = src.index[j];
    dst.r_o[k] = src.r_o[j];
    dst.v_o[k] = src.v_o[j];
    dst.rho_o[k] = src.rho_o[j];
    double p = src.p[j];
    const Consts kc = EOS ? case_consts(k0, cases, src.cid, j) : k0;
    if (EOS && t != TSPH_OUTLET && t != TSPH_INLET) {
      bad = rho > 1500. || rho < 0.;
      p = tait64(rho, kc.rho0, kc.cs);
    }
    dst.p[k] = p;
    if (src.gnr) dst.gnr[k] = src.gnr[j];
    double2 nv = make_double2(0., 0.);
    if (src.n) {
      nv = src.n[j];
      dst.n[k] = nv;
    }
    if (EOS && mx.rv)
      write_mixed_record(mx, kc, cellof[k] - kc.cell_base, k, r, v, rho, p, src.n != nullptr, nv);
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
  }
  const unsigned long long m = __ballot(wall);
  if (m) {
    const int lane = threadIdx.x & 63;
    int base = 0;
    if (lane == __builtin_ctzll(m)) base = atomicAdd(&scal[S_NWALL], __popcll(m));
    base = __shfl(base, __builtin_ctzll(m), 64);
    if (wall) wall_list[base + __popcll(m & lanemask_lt())] = k;
  }
  if (EOS) {
    const unsigned long long w = __ballot(bad);
    if (w && (threadIdx.x & 63) == 0) atomicAdd(&counters[C_BADRHO], (unsigned long long)__popcll(w));
  }
}

int scan_exclusive(int *in, int len, int *out, int *partial, int *scal, hipStream_t s) {
  const int nb = (len + kScanTile - 1) / kScanTile;
  hipLaunchKernelGGL(k_scan_partials, dim3(nb), dim3(kScanBlock), 0, s, in, len, partial);
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, s, partial, nb, out, len, scal);
  hipLaunchKernelGGL(k_scan_final, dim3(nb), dim3(kScanBlock), 0, s, in, len, partial, out);
  TSPH_HIP(hipGetLastError());
  return 0;
}

}  // namespace

int launch_cell_list(tsph_ctx *c, bool eos) {
  const Consts &k = c->k;
  const int nb = 256;
  const int g = grid_for(c->n_ub, nb);
  hipStream_t s = c->stream;
  const int nct = c->ncells_tot;
  hipLaunchKernelGGL(k_hash, dim3(g), dim3(nb), 0, s, c->cur.r, c->cur.type, c->cur.seq, c->scal, k,
                     c->cases, c->cur.cid, nct, c->cell_count, c->hcell, c->hrank, c->seqstate,
                     c->counters);
  if (int e = scan_exclusive(c->cell_count, nct + 1, c->cell_start, c->scan_tmp, c->scal, s))
    return e;
  hipLaunchKernelGGL(k_scatter, dim3(g), dim3(nb), 0, s, c->scal, c->hcell, c->hrank,
                     c->cell_start, c->perm, c->cellof);
  hipLaunchKernelGGL(k_order, dim3(grid_for(nct, nb)), dim3(nb), 0, s, c->cell_start, nct,
                     c->cur.seq, c->perm);
  const Mixed mx = c->prm.precision == TSPH_PREC_MIXED ? c->mx : Mixed{};
  if (eos)
    hipLaunchKernelGGL(k_reorder<true>, dim3(g), dim3(nb), 0, s, c->cell_start, nct, c->perm,
                       c->cur, c->alt, c->scal, c->wall_list, c->cellof, k, c->cases, mx, c->counters);
  else
    hipLaunchKernelGGL(k_reorder<false>, dim3(g), dim3(nb), 0, s, c->cell_start, nct, c->perm,
                       c->cur, c->alt, c->scal, c->wall_list, c->cellof, k, c->cases, mx, c->counters);
  TSPH_HIP(hipGetLastError());
  std::swap(c->cur, c->alt);
  c->parity ^= 1;
  c->cells_valid = true;
  c->tiles_valid = false;
  c->records_valid = eos && c->prm.precision == TSPH_PREC_MIXED;
  return 0;
}

}  // namespace tsph
