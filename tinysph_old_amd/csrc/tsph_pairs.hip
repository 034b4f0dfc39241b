// Pair loops: the device form of Compute_Forces.
//   TSPH_FORCES_WCSPH  src/SPH_interaction.h:8-149    (active: fluid)
//   TSPH_FORCES_DBC    src/SPH_interactionDBC.h:8-155 (active: all but inlet/outlet,
//                      wall-wall pairs skipped, walls keep dRho but a = 0)
//   TSPH_FORCES_BT     src/SPH_interactionBT.h:8-192  (wall neighbours enter as a
//                      boundary integral with normal n; Shepard gamma normalises)
//
// Work decomposition.  Particles are stored in (cell, seq) order and cell ids
// run along grid columns (POS(x,y) = y + x*ncy), so a run of consecutive cells
// of one column holds a contiguous slice of particles, and its 9-cell
// neighbourhood is three contiguous slices (columns cx-1, cx, cx+1, rows
// cy0-1 .. cy1).  One wavefront takes one such tile of `tile_rows` cells:
//   0. it loads the cell offsets of the three slices (one load per lane) and
//      stages the slices into LDS with coalesced loads;
//   1. each lane takes one particle of the tile and walks its candidates,
//      keeping those inside the 2h support in a lane-private list in LDS
//      (a cheap distance test per candidate);
//   2. each lane then runs the pair physics over its list.
// Splitting the loop this way keeps the lanes busy: only ~30% of 9-cell
// candidates are real pairs, and a single loop would make every lane wait for
// the pair arithmetic of the others at almost every candidate.  A lane whose
// list fills up processes it and resumes its candidate walk.  Sums stay in the
// lane's registers: no atomics, deterministic results.
// Workgroups are dispatched round-robin over the 8 XCDs; the tile index is
// remapped so that every XCD walks one contiguous band of grid columns and the
// neighbour columns a tile shares with the next are found in that XCD's L2.
// A tile whose neighbourhood exceeds the LDS budget gathers from global memory
// instead (same arithmetic).
//
// Summation order.  FP64 (k_tile64) visits the 9 cells in the reference order
// {c, c+1, c+1+ncy, c+ncy, c-1+ncy, c-1, c-1-ncy, c-ncy, c+1-ncy}
// (src/SPH_interaction.h:26-36) and each cell in ascending reference index, and
// its lists keep that order, so every per-particle sum is formed in the
// reference order.  Cells beyond the top/bottom row of a column, which the
// reference reaches by wrapping into the neighbouring column, hold only
// particles far outside the 2h support and add exact zeros there; they are not
// visited.  Candidates beyond 2h likewise add exact zeros in the reference
// (W = F = 0); they are dropped by the same qq <= 2 test.
// The mixed form (k_tile32) visits the three column slices in memory order.
//
// Arithmetic.  k_tile64: IEEE double in the reference operand order (built with
// -ffp-contract=off).  k_tile32: FP32 pair arithmetic with hardware reciprocal
// and square root; positions are staged relative to the tile (x from the
// tile's column origin, y from its middle row), which keeps |x|, |y| < 6 cells
// and their FP32 rounding below 1e-6 of the particle spacing; FP32
// accumulation; FP64 outputs for the FP64 integrator.
#include "tsph_internal.h"
#include "tsph_physics.h"
#include "tsph_tiles.h"

namespace tsph {

namespace {

using namespace tiles;


template <int VAR>
__device__ __forceinline__ bool is_active(int t) {
  return VAR == TSPH_FORCES_DBC ? (t != TSPH_INLET && t != TSPH_OUTLET) : (t == TSPH_FLUID);
}

// --------------------------------------------------------------------------
// FP64 per-pair accumulation (reference operand order)
// --------------------------------------------------------------------------
struct Acc64 {
  double accx = 0., accy = 0., drho_sum = 0., diff_sum = 0., gamma = 0.;
};


template <int VAR>
__device__ __forceinline__ void pair64(const Consts &k, Acc64 &s, double2 ar, double2 av, double arho,
                                       double ap, double2 nr, double2 nv, double nrho, double np,
                                       int nt, double2 nb, unsigned &npair) {
  const double h = k.h;
  const double drx = ar.x - nr.x, dry = ar.y - nr.y;
  const double r2 = (drx * drx) + (dry * dry);
  if (r2 > 4.0 * h * h * (1.0 + 1e-12)) return;
  const double drs = sqrt(r2);
  if (drs / h > 2.0) return;
  if (drs / h < 2.0) ++npair;
  const KW64 kw = wendland64(drs, h, k.awen, k.bwen);
  const double dvx = av.x - nv.x, dvy = av.y - nv.y;
  const double drdv = drx * dvx + dry * dvy;
  const double dWx = drx * kw.f, dWy = dry * kw.f;
  const double dvdW = dvx * dWx + dvy * dWy;
  const double drdW = drx * dWx + dry * dWy;
  const double p_temp = (ap + np) / (arho * nrho);
  const double visco = art_visc64(h, drs, drdv, 0.5 * (arho + nrho), k.cs, k.avisc);
  if (VAR == TSPH_FORCES_BT && nt == TSPH_WALL) {
    s.accx += kw.w * (p_temp + visco) * k.dp * nb.x * nrho;
    s.accy += kw.w * (p_temp + visco) * k.dp * nb.y * nrho;
    const double dvnb = dvx * nb.x + dvy * nb.y;
    s.drho_sum -= dvnb * kw.w * nrho * k.dp;
  } else {
    const double diff =
        diff_fourtakas64(h, drs, dry, drdW, arho, nrho, k.rho0, k.cs, k.delta, k.m, k.gy);
    s.accx -= dWx * (p_temp + visco) * k.m;
    s.accy -= dWy * (p_temp + visco) * k.m;
    s.drho_sum += dvdW * k.m;
    s.diff_sum -= diff;
  }
  if (VAR == TSPH_FORCES_BT && nt == TSPH_FLUID) s.gamma += kw.w * k.m / nrho;
}

// final assignment, src/SPH_interaction.h:134-136, src/SPH_interactionDBC.h:137-142,
// src/SPH_interactionBT.h:166-178
template <int VAR>
__device__ __forceinline__ void finish64(const Consts &k, const Acc64 &s, int at, double2 &a, double &dr) {
  if (VAR == TSPH_FORCES_BT) {
    if (s.gamma > kEps) {
      a = make_double2(s.accx / s.gamma - k.gx, s.accy / s.gamma - k.gy);
      dr = s.drho_sum / s.gamma + s.diff_sum / s.gamma;
    } else {
      a = make_double2(0. - k.gx, 0. - k.gy);
      dr = 0;
    }
  } else if (VAR == TSPH_FORCES_DBC && at == TSPH_WALL) {
    a = make_double2(0., 0.);
    dr = s.drho_sum + s.diff_sum;
  } else {
    a = make_double2(s.accx - k.gx, s.accy - k.gy);
    dr = s.drho_sum + s.diff_sum;
  }
}

// --------------------------------------------------------------------------
// FP32 per-pair accumulation
// --------------------------------------------------------------------------
struct Acc32 {
  float accx = 0.f, accy = 0.f, drho_sum = 0.f, diff_sum = 0.f, gamma = 0.f;
};

struct C32 {
  float h, ih, m, c0, rho0, dp, dpm, kh, r2cut, epsh2, awen, bwen, visc, dcoef, hx;
};

__device__ __forceinline__ C32 make_c32(const Consts &k) {
  C32 c;
  c.h = (float)k.h; c.ih = 1.f / c.h; c.m = (float)k.m; c.c0 = (float)k.cs; c.rho0 = (float)k.rho0;
  c.dp = (float)k.dp; c.dpm = (float)(k.dp / k.m); c.kh = (float)k.kh; c.r2cut = 4.f * c.h * c.h;
  c.epsh2 = (float)kEps * c.h * c.h; c.awen = (float)k.awen; c.bwen = (float)k.bwen;
  // Pi = -alpha c0 (h drdv / (r^2 + eps h^2)) / (0.5 (rho_i + rho_j)) for drdv < 0
  c.visc = -2.f * (float)k.avisc * c.c0 * c.h;
  c.dcoef = 2.f * c.c0 * c.h * (float)k.delta;
  c.hx = (float)(k.rho0 * k.gy / (k.cs * k.cs * k.rho0 / 7.));  // rho0 g / cb
  return c;
}

// one pair of the mixed path; dr = r_i - r_j with r2 = |dr|^2 < (2h)^2; dev is
// rho - rho0, ir is 1/rho.  Every sum is kept divided by the particle mass m
// (the BT wall terms, which carry dp instead, by dp/m); finish32 applies m
// once, or lets it cancel in the BT ratios.
template <int VAR>
__device__ __forceinline__ void pair32(const C32 &c, Acc32 &s, float drx, float dry, float r2,
                                       float2 vi, float adev, float ap, float air, float2 vj,
                                       float ndev, float np, float nir, int nt, float2 nb) {
#pragma clang fp contract(fast)
  const float drs = __builtin_amdgcn_sqrtf(r2);
  const float qq = drs * c.ih;
  const float wq1 = 1.f - 0.5f * qq;
  const float wq2 = wq1 * wq1;
  const float f = c.bwen * wq2 * wq1 * c.ih;  // F = bwen qq wq2 wq1 / r
  const float dvx = vi.x - vj.x, dvy = vi.y - vj.y;
  const float drdv = drx * dvx + dry * dvy;
  const float p_temp = (ap + np) * (air * nir);
  const float rinv = __builtin_amdgcn_rcpf(r2 + c.epsh2);
  float visco = 0.f;
  if (drdv < 0.f) visco = c.visc * drdv * rinv * __builtin_amdgcn_rcpf(2.f * c.rho0 + adev + ndev);
  if (VAR == TSPH_FORCES_BT && nt == TSPH_WALL) {
    const float w = c.awen * (qq + qq + 1.f) * wq2 * wq2;
    const float nrho = c.rho0 + ndev;
    const float t = w * (p_temp + visco) * c.dpm * nrho;
    s.accx += t * nb.x;
    s.accy += t * nb.y;
    s.drho_sum -= (dvx * nb.x + dvy * nb.y) * w * nrho * c.dpm;
  } else {
    const float dWx = drx * f, dWy = dry * f;  // dr.dW = f r^2
    const float dvdW = dvx * dWx + dvy * dWy;
    const float drdW = f * r2;
    const float drhoH = hydro_drho_f32(dry * c.hx, c.rho0);
    const float psi = c.dcoef * ((ndev - adev) - drhoH) * rinv;
    const float t = p_temp + visco;
    s.accx -= dWx * t;
    s.accy -= dWy * t;
    s.drho_sum += dvdW;
    s.diff_sum -= psi * drdW * nir;
    if (VAR == TSPH_FORCES_BT && nt == TSPH_FLUID) {
      const float w = c.awen * (qq + qq + 1.f) * wq2 * wq2;
      s.gamma += w * nir;
    }
  }
}

// the sums of pair32 are divided by m: m is applied here (in the BT ratios it
// cancels; its Shepard sum is compared as m gamma)
template <int VAR>
__device__ __forceinline__ void finish32(const Consts &k, const Acc32 &s, int at, double2 &a, double &dr) {
  if (VAR == TSPH_FORCES_BT) {
    if ((double)s.gamma * k.m > kEps) {
      a = make_double2((double)(s.accx / s.gamma) - k.gx, (double)(s.accy / s.gamma) - k.gy);
      dr = (double)(s.drho_sum / s.gamma + s.diff_sum / s.gamma);
    } else {
      a = make_double2(0. - k.gx, 0. - k.gy);
      dr = 0;
    }
  } else if (VAR == TSPH_FORCES_DBC && at == TSPH_WALL) {
    a = make_double2(0., 0.);
    dr = (double)(s.drho_sum + s.diff_sum) * k.m;
  } else {
    a = make_double2((double)s.accx * k.m - k.gx, (double)s.accy * k.m - k.gy);
    dr = (double)(s.drho_sum + s.diff_sum) * k.m;
  }
}

// --------------------------------------------------------------------------
// global-memory forms (tiles whose neighbourhood overflows the LDS budget)
// --------------------------------------------------------------------------
template <int VAR>
__device__ void particle64_global(const Soa &d, const int *__restrict__ cell_start,
                                  const int *__restrict__ cellof, const Consts &k, int i,
                                  unsigned &npair, unsigned &ncand) {
  const int at = d.type[i];
  const double2 ar = d.r[i], av = d.v[i];
  const double arho = d.rho[i], ap = d.p[i];
  const int c = cellof[i] - k.cell_base;
  const int cx = c / k.ncy, cy = c - cx * k.ncy;
  Acc64 s;
  for (int q = 0; q < 9; ++q) {
    int dx, dy;
    stencil_dxy(q, dx, dy);
    if (cx + dx < 0 || cx + dx >= k.ncx || cy + dy < 0 || cy + dy >= k.ncy) continue;
    const int cl = c + dy + dx * k.ncy;
    const int j1 = cell_start[cl + 1];
    for (int j = cell_start[cl]; j < j1; ++j) {
      if (j == i) continue;
      const int nt = VAR != TSPH_FORCES_WCSPH ? d.type[j] : 0;
      if (VAR == TSPH_FORCES_DBC && at == TSPH_WALL && nt == TSPH_WALL) continue;
      ++ncand;
      const double2 nb = (VAR == TSPH_FORCES_BT) ? d.n[j] : make_double2(0., 0.);
      pair64<VAR>(k, s, ar, av, arho, ap, d.r[j], d.v[j], d.rho[j], d.p[j], nt, nb, npair);
    }
  }
  double2 a;
  double dr;
  finish64<VAR>(k, s, at, a, dr);
  d.a[i] = a;
  d.drho[i] = dr;
}

template <int VAR>
__device__ void particle32_global(const Soa &d, const Mixed &mx, const int *__restrict__ cell_start,
                                  const int *__restrict__ cellof, const Consts &k, const C32 &cc,
                                  int i, unsigned &npair, unsigned &ncand) {
  const int at = d.type[i];
  const float4 ai = mx.rv[i];
  const float2 api = mx.rp[i];
  const float air = __builtin_amdgcn_rcpf(cc.rho0 + api.x);
  const int c = cellof[i] - k.cell_base;
  const int cx = c / k.ncy, cy = c - cx * k.ncy;
  Acc32 s;
  for (int dx = -1; dx <= 1; ++dx) {
    if (cx + dx < 0 || cx + dx >= k.ncx) continue;
    for (int dy = -1; dy <= 1; ++dy) {
      if (cy + dy < 0 || cy + dy >= k.ncy) continue;
      const int cl = c + dy + dx * k.ncy;
      const float bx = ai.x - (float)dx * cc.kh, by = ai.y - (float)dy * cc.kh;
      const int j1 = cell_start[cl + 1];
      for (int j = cell_start[cl]; j < j1; ++j) {
        if (j == i) continue;
        const int nt = VAR != TSPH_FORCES_WCSPH ? d.type[j] : 0;
        if (VAR == TSPH_FORCES_DBC && at == TSPH_WALL && nt == TSPH_WALL) continue;
        ++ncand;
        const float4 nj = mx.rv[j];
        const float drx = bx - nj.x, dry = by - nj.y;
        const float r2 = drx * drx + dry * dry;
        if (!(r2 < cc.r2cut)) continue;
        ++npair;
        const float2 pj = mx.rp[j];
        const float2 nb = (VAR == TSPH_FORCES_BT) ? mx.nrm[j] : make_float2(0.f, 0.f);
        pair32<VAR>(cc, s, drx, dry, r2, make_float2(ai.z, ai.w), api.x, api.y, air,
                    make_float2(nj.z, nj.w), pj.x, pj.y, __builtin_amdgcn_rcpf(cc.rho0 + pj.x), nt,
                    nb);
      }
    }
  }
  double2 a;
  double dr;
  finish32<VAR>(k, s, at, a, dr);
  d.a[i] = a;
  d.drho[i] = dr;
}


// The list of non-empty tiles in tile order, built in two passes over blocks
// of kListBlock tiles: k_tile_count counts each block's non-empty tiles;
// k_tile_emit adds up the counts of the blocks before its own (the block
// count is small), scans its block and writes the list.
constexpr int kListBlock = 256;



__global__ __launch_bounds__(kListBlock) void k_tile_count(const int *__restrict__ cell_start, Consts k0,
                                                           const Consts *__restrict__ cases,
                                                           const int *__restrict__ tile_base,
                                                           int ncases, int rows, int ntiles,
                                                           int *__restrict__ part) {
  __shared__ int lds[kListBlock / 64];
  const int t = blockIdx.x * kListBlock + threadIdx.x;
  const int own = t < ntiles ? tile_own(cell_start, k0, cases, tile_base, ncases, rows, t) : 0;
  int n = 0;
  block_exscan(own > 0 ? 1 : 0, n, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = n;
}

__global__ __launch_bounds__(kListBlock) void k_tile_emit(const int *__restrict__ cell_start, Consts k0,
                                                          const Consts *__restrict__ cases,
                                                          const int *__restrict__ tile_base,
                                                          int ncases, int rows, int ntiles,
                                                          const int *__restrict__ part,
                                                          int *__restrict__ scal,
                                                          int *__restrict__ tile_list) {
  __shared__ int lds[kListBlock / 64];
  const int nb = gridDim.x;
  // tiles of the blocks before this one, and of all blocks
  int bn = 0, tn = 0;
  for (int b = threadIdx.x; b < nb; b += kListBlock) {
    const int n = part[b];
    if (b < (int)blockIdx.x) bn += n;
    tn += n;
  }
  int sum = 0;
  block_exscan(bn, sum, lds);
  bn = sum;
  block_exscan(tn, sum, lds);
  tn = sum;
  const int t = blockIdx.x * kListBlock + threadIdx.x;
  const int own = t < ntiles ? tile_own(cell_start, k0, cases, tile_base, ncases, rows, t) : 0;
  const int flag = own > 0;
  const int idx = bn + block_exscan(flag, sum, lds);
  if (flag) tile_list[idx] = t;
  if (blockIdx.x == 0 && threadIdx.x == 0) scal[S_NTILES] = tn;
}

// the recorded costs (cost table: tile t at cost_in[kCostTables t]) of the
// listed tiles by list position, or their own-particle estimates
__global__ void k_list_costs(const int *__restrict__ cell_start, Consts k0,
                             const Consts *__restrict__ cases, const int *__restrict__ tile_base,
                             int ncases, int rows, const int *__restrict__ scal,
                             const int *__restrict__ cost_in, const int *__restrict__ tile_list,
                             int *__restrict__ list_cost, CostEstimate cm) {
  const int n = scal[S_NTILES];
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
    const int t = tile_list[e];
    const int w = cost_in[kCostTables * t];
    list_cost[e] = w > 0 ? w : cm.own * tile_own(cell_start, k0, cases, tile_base, ncases, rows, t) + cm.tile;
  }
}

// the chunk cut of a pooled walk as a launch of its own, for a tile list that
// its cell-list build did not cut (cut_chunks_block, one workgroup per segment)
__global__ __launch_bounds__(kCutBlock) void k_cut_chunks(const int *__restrict__ scal,
                                                          const int *__restrict__ list_cost,
                                                          int groups, int *__restrict__ bound) {
  __shared__ int wl[kCutSeg];
  __shared__ long long red[kCutBlock / 64];
  cut_chunks_block(scal, list_cost, groups, bound, blockIdx.x, wl, red);
}

// --------------------------------------------------------------------------
// FP64 tile kernel
// --------------------------------------------------------------------------
template <int VAR, bool DYN>
__global__ __launch_bounds__(kWave * kPoolWaves) void k_tile64(Soa d, const int *__restrict__ cell_start,
                                                  const int *__restrict__ cellof, Consts k0,
                                                  const Consts *__restrict__ cases,
                                                  const int *__restrict__ tile_base, int ncases,
                                                  int rows, int *__restrict__ scal,
                                                  const int *__restrict__ tile_list, int cap,
                                                  const int *__restrict__ bound, int lds_wave,
                                                  unsigned long long *__restrict__ counters) {
  extern __shared__ __align__(16) unsigned char smem[];
  unsigned char *wl = wave_lds<DYN>(smem, lds_wave);
  int *tab = reinterpret_cast<int *>(wl);
  double2 *s_r = reinterpret_cast<double2 *>(wl + kTabBytes);
  double2 *s_v = s_r + cap;
  double2 *s_rp = s_v + cap;  // {rho, p}
  double2 *s_n = s_rp + cap;  // BT only
  uint16_t *s_list = reinterpret_cast<uint16_t *>(s_n + (VAR == TSPH_FORCES_BT ? cap : 0));
  uint8_t *s_t = reinterpret_cast<uint8_t *>(s_list + kList * kWave);
  uint8_t *s_row = s_t + cap;  // row of the slot within the tile's rows

  if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&counters[C_PASSES], 1ull);
  unsigned npair = 0, ncand = 0, nact = 0;
  TileQueue tq = tq_begin<DYN>(scal, bound, tile_list, smem);
  for (int it = tq_take<DYN>(tq, tq_ticket<DYN>(tq)); it >= 0;) {
  wave_sync();  // the previous tile's LDS is no longer read
  // the next tile's ticket: its atomic is in flight with this tile's loads,
  // and it is read once they have been waited for
  const int ticket = tq_ticket<DYN>(tq);
  // the tile's case: its constants and its segment of the cell range
  int tile = tile_list[it];
  Consts k = k0;
  if (cases) {
    const int cse = case_of_tile(tile_base, ncases, tile);
    tile -= tile_base[cse];
    k = cases[cse];
  }
  const int *cs = cell_start + k.cell_base;
  Tile t;
  tile_begin(k, cs, tile, rows, tab, t);
  tq_mark(tq, tile_list[it]);
  if (t.lq[3] > cap) {
    const int next = tq_take<DYN>(tq, ticket);
    for (int i = t.own0 + tile_lane(); i < t.own1; i += kWave) {
      if (!is_active<VAR>(d.type[i])) continue;
      ++nact;
      particle64_global<VAR>(d, cs, cellof, k, i, npair, ncand);
    }
    it = next;
    continue;
  }
  for (int l = tile_lane(); l < t.lq[3]; l += kWave) {
    const int q = slice_of(t, l);
    const int j = l - sel3(t.base, q);
    s_r[l] = d.r[j];
    s_v[l] = d.v[j];
    s_rp[l] = make_double2(d.rho[j], d.p[j]);
    s_t[l] = d.type[j];
    s_row[l] = (uint8_t)(cellof[j] - k.cell_base - (t.cx + q - 1) * k.ncy - t.row_lo);
    if (VAR == TSPH_FORCES_BT) s_n[l] = d.n[j];
  }
  wave_sync();
  const int next = tq_take<DYN>(tq, ticket);

  const int lane = tile_lane();
  uint16_t *my_list = s_list + lane;  // entry e at my_list[e * kWave]
  for (int i0 = t.own0; i0 < t.own1; i0 += kWave) {
    const int i = i0 + lane;
    int li = 0, at = 0, cy = 0;
    bool act = false;
    if (i < t.own1) {
      li = i + t.base[1];
      at = s_t[li];
      act = is_active<VAR>(at);
      cy = s_row[li] + t.row_lo;
    }
    nact += act;
    const double2 ar = s_r[li], av = s_v[li], arp = s_rp[li];
    // walk of the 9 stencil cells in reference order: cell q, LDS range [lj, lend).
    // A cell's candidates are tested kGroup at a time (their LDS loads in
    // flight together) and appended without branches: every slot of the group
    // is written, the count advances by the test.  The test admits a superset
    // of the support (r^2 <= 4h^2 (1 + 1e-12)); pair64 applies the reference's
    // exact qq <= 2, and a candidate it drops adds nothing, as in the
    // reference, so the list keeps every pair in the reference order.  Slots
    // past the cell's end are masked (they lie inside the wave's LDS area).
    const double c2 = 4.0 * k.h * k.h * (1.0 + 1e-12);
    const bool dbc_wall = VAR == TSPH_FORCES_DBC && at == TSPH_WALL;
    int q = -1, lj = 0, lend = 0;
    bool more = act;
    Acc64 s;
    while (__any(more)) {
      int cnt = 0;
      while (more && cnt <= kList - kGroup) {
        if (lj >= lend) {
          ++q;
          if (q == 9) { more = false; break; }
          int dx, dy;
          stencil_dxy(q, dx, dy);
          const int row = cy + dy;
          if (row < 0 || row >= k.ncy) { lend = lj; continue; }  // wrapped cell: no visit
          const int e = (dx + 1) * t.w + (row - t.row_lo);
          const int bq = sel3(t.base, dx + 1);
          lj = tab[e] + bq;
          lend = tab[e + 1] + bq;
          continue;
        }
        const int m = min(kGroup, lend - lj);
        double2 pp[kGroup];
#pragma unroll
        for (int u = 0; u < kGroup; ++u) pp[u] = s_r[lj + u];
#pragma unroll
        for (int u = 0; u < kGroup; ++u) {
          const int j = lj + u;
          bool cand = u < m && j != li;
          if (VAR == TSPH_FORCES_DBC) cand = cand && !(dbc_wall && s_t[j] == TSPH_WALL);
          ncand += cand;
          const double drx = ar.x - pp[u].x, dry = ar.y - pp[u].y;
          const bool in = cand && (drx * drx) + (dry * dry) <= c2;
          my_list[cnt * kWave] = (uint16_t)j;
          cnt += in;
        }
        lj += m;
      }
      for (int e = 0; e < cnt; ++e) {
        const int j = my_list[e * kWave];
        const double2 nrp = s_rp[j];
        const int nt = VAR != TSPH_FORCES_WCSPH ? s_t[j] : 0;
        const double2 nb = VAR == TSPH_FORCES_BT ? s_n[j] : make_double2(0., 0.);
        pair64<VAR>(k, s, ar, av, arp.x, arp.y, s_r[j], s_v[j], nrp.x, nrp.y, nt, nb, npair);
      }
    }
    if (act) {
      double2 a;
      double dr;
      finish64<VAR>(k, s, at, a, dr);
      d.a[i] = a;
      d.drho[i] = dr;
    }
  }
  it = next;
  }  // tiles
  tq_mark(tq, -1);
  wave_count(npair, ncand, nact, counters);
}

// --------------------------------------------------------------------------
// mixed tile kernel
// --------------------------------------------------------------------------
// CORR: the symplectic corrector (src/SPH_integration.h:287-309) is applied to
// every particle of the tile as soon as its forces are known (inactive ones
// with the a, drho they hold).  Other lanes never read a particle's FP64 state
// in this kernel (neighbours come from the FP32 records), so the update may
// land while the pass runs; it replaces the corrector pass that follows the
// first pair pass of a symplectic step.
template <int VAR, bool DYN, bool CORR>
__global__ __launch_bounds__(kWave * kPoolWaves) void k_tile32(Soa d, Mixed mx, const int *__restrict__ cell_start,
                                                  const int *__restrict__ cellof, Consts k0,
                                                  const Consts *__restrict__ cases,
                                                  const int *__restrict__ tile_base, int ncases,
                                                  int rows, int *__restrict__ scal,
                                                  const int *__restrict__ tile_list, int cap,
                                                  const int *__restrict__ bound, int lds_wave,
                                                  unsigned long long *__restrict__ counters,
                                                  double dt0, uint8_t *__restrict__ state) {
#pragma clang fp contract(fast)
  extern __shared__ __align__(16) unsigned char smem[];
  unsigned char *wl = wave_lds<DYN>(smem, lds_wave);
  int *tab = reinterpret_cast<int *>(wl);
  float2 *s_pos = reinterpret_cast<float2 *>(wl + kTabBytes);  // tile frame
  float2 *s_vel = s_pos + cap;
  float2 *s_dp = s_vel + cap;   // {rho - rho0, p}
  float2 *s_n = s_dp + cap;     // BT only
  float *s_ir = reinterpret_cast<float *>(s_n + (VAR == TSPH_FORCES_BT ? cap : 0));  // 1/rho
  uint16_t *s_list = reinterpret_cast<uint16_t *>(s_ir + cap);
  uint8_t *s_t = reinterpret_cast<uint8_t *>(s_list + kList * kWave);
  uint8_t *s_row = s_t + cap;  // row of the slot within the tile's rows

  if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&counters[C_PASSES], 1ull);
  unsigned npair = 0, ncand = 0, nact = 0;
  int pend_i = -1;  // particle whose results are still to be stored
  double2 pend_a = make_double2(0., 0.);
  double pend_dr = 0.;
  TileQueue tq = tq_begin<DYN>(scal, bound, tile_list, smem, CORR ? 1 : 0);
  for (int it = tq_take<DYN>(tq, tq_ticket<DYN>(tq)); it >= 0;) {
  wave_sync();  // the previous tile's LDS is no longer read
  // the next tile's ticket: its atomic is in flight with this tile's loads,
  // and it is read once they have been waited for
  const int ticket = tq_ticket<DYN>(tq);
  // the tile's case: its constants and its segment of the cell range
  int tile = tile_list[it];
  Consts k = k0;
  if (cases) {
    const int cse = case_of_tile(tile_base, ncases, tile);
    tile -= tile_base[cse];
    k = cases[cse];
  }
  const int *cs = cell_start + k.cell_base;
  const C32 cc = make_c32(k);
  Tile t;
  tile_begin(k, cs, tile, rows, tab, t);
  tq_mark(tq, tile_list[it]);
  // LDS layout with kGroup sentinels after each column slice's records: a
  // candidate group may run past the end of its slice (the sentinels lie far
  // away and fail every distance test) and need not be masked
  int se[3];  // end of slice q's records
#pragma unroll
  for (int q = 0; q < 3; ++q) se[q] = t.lq[q + 1] + q * kGroup;
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    t.lq[q] += q * kGroup;
    t.base[q] += q * kGroup;
  }
  t.lq[3] += 3 * kGroup;
  const double dtc = cases ? k.dt : dt0;  // the corrector's time step (CORR)
  if (t.lq[3] > cap) {
    const int next = tq_take<DYN>(tq, ticket);
    for (int i = t.own0 + tile_lane(); i < t.own1; i += kWave) {
      if (is_active<VAR>(d.type[i])) {
        ++nact;
        particle32_global<VAR>(d, mx, cs, cellof, k, cc, i, npair, ncand);
      }
      if (CORR) correct_particle(d, k, dtc, i, d.a[i], d.drho[i], state);
    }
    it = next;
    continue;
  }
  const int mid = t.row_lo + (t.nrows >> 1);
  // staging: the records of kStageU slots per lane are loaded together (slots
  // past the end load the last slot again, so every load is consumed on every
  // path and no wait is left for the compiler to place conservatively).  The
  // previous tile's last results are stored after the staging: their write
  // acknowledgements then return while this tile is computed, instead of
  // delaying the wait for the staging loads (one counter covers both).
  for (int l0 = 0; l0 < t.lq[3]; l0 += kStageU * kWave) {
    float4 rv[kStageU];
    float2 rp[kStageU], nr[kStageU];
    int cl[kStageU], ty[kStageU];
#pragma unroll
    for (int u = 0; u < kStageU; ++u) {
      const int l = min(l0 + u * kWave + tile_lane(), t.lq[3] - 1);
      const int q = slice_of(t, l);
      // sentinel slots load the tile's first own particle (never used)
      const int j = l < sel3(se, q) ? l - sel3(t.base, q) : t.own0;
      rv[u] = mx.rv[j];
      rp[u] = mx.rp[j];
      cl[u] = cellof[j];
      ty[u] = d.type[j];
      if (VAR == TSPH_FORCES_BT) nr[u] = mx.nrm[j];
    }
#pragma unroll
    for (int u = 0; u < kStageU; ++u) {
      const int l = l0 + u * kWave + tile_lane();
      const int q = slice_of(t, min(l, t.lq[3] - 1));
      const int row = cl[u] - k.cell_base - (t.cx + q - 1) * k.ncy;
      const bool real = l < sel3(se, q);
      const float2 pos = real ? make_float2(rv[u].x + (float)(q - 1) * cc.kh,
                                            rv[u].y + (float)(row - mid) * cc.kh)
                              : make_float2(kFar32, kFar32);
      const float ir = __builtin_amdgcn_rcpf(cc.rho0 + rp[u].x);
      if (l < t.lq[3]) {
        s_pos[l] = pos;
        s_vel[l] = make_float2(rv[u].z, rv[u].w);
        s_dp[l] = rp[u];
        s_ir[l] = ir;
        s_t[l] = (uint8_t)ty[u];
        s_row[l] = (uint8_t)(row - t.row_lo);
        if (VAR == TSPH_FORCES_BT) s_n[l] = nr[u];
      }
    }
  }
  if (pend_i >= 0) {
    d.a[pend_i] = pend_a;
    d.drho[pend_i] = pend_dr;
    pend_i = -1;
  }
  wave_sync();
  const int next = tq_take<DYN>(tq, ticket);

  const int lane = tile_lane();
  uint16_t *my_list = s_list + lane;
  for (int i0 = t.own0; i0 < t.own1; i0 += kWave) {
    const int i = i0 + lane;
    int li = 0, at = 0, ri = 0;
    bool act = false;
    if (i < t.own1) {
      li = i + t.base[1];
      at = s_t[li];
      act = is_active<VAR>(at);
      ri = s_row[li];
    }
    nact += act;
    const float2 pi = s_pos[li], vi = s_vel[li], dpi = s_dp[li];
    const float iri = s_ir[li];
    // the three column slices, rows ri-1 .. ri+1, as one flat candidate range
    const int rlo = max(ri - 1, 0), rhi = min(ri + 2, t.nrows);
    const int a0 = tab[rlo] + t.base[0], n0 = tab[rhi] + t.base[0] - a0;
    const int a1 = tab[t.w + rlo] + t.base[1], n1 = tab[t.w + rhi] + t.base[1] - a1;
    const int a2 = tab[2 * t.w + rlo] + t.base[2], n2 = tab[2 * t.w + rhi] + t.base[2] - a2;
    const int tot = act ? n0 + n1 + n2 : 0;
    // candidate c of the flattened range lives at LDS slot c + off(c)
    const int n01 = n0 + n1, off1 = a1 - n0, off2 = a2 - n01;
    const bool dbc_wall = VAR == TSPH_FORCES_DBC && at == TSPH_WALL;
    const unsigned c2m1 = __float_as_uint(cc.r2cut) - 1u;
    int cur = 0;
    Acc32 s;
    while (__any(cur < tot)) {
      int cnt = 0;
      // Candidates in groups of kGroup from one column slice (so one slot
      // offset serves the group), while the list has room for a group: their
      // LDS loads are issued together, and the group is appended to the list
      // without branches (every slot is written, the count advances by the
      // test).  Slots past the lane's range in the slice are rows at least
      // one cell (> 2h) away, or the slice's far sentinels, and fail the test.
      // The particle itself is the candidate at distance exactly 0; "0 < r^2"
      // drops it (and only a coincident particle with it, whose reference
      // kernel gradient would be 0/0).  0 < r^2 < (2h)^2 is one unsigned
      // compare of the bit patterns (r^2 >= 0, so they order as the values).
      while (cur < tot && cnt <= kList - kGroup) {
        const bool in0 = cur < n0, in01 = cur < n01;
        const int end = in0 ? n0 : (in01 ? n01 : tot);
        const int j0 = cur + (in0 ? a0 : (in01 ? off1 : off2));
        const int m = min(kGroup, end - cur);
        float2 pp[kGroup];
#pragma unroll
        for (int u = 0; u < kGroup; ++u) pp[u] = s_pos[j0 + u];
        bool kk[kGroup];
#pragma unroll
        for (int u = 0; u < kGroup; ++u) {
          const float x = pi.x - pp[u].x, y = pi.y - pp[u].y;
          const float q = x * x + y * y;
          kk[u] = __float_as_uint(q) - 1u < c2m1;
        }
        if (VAR == TSPH_FORCES_DBC) {
#pragma unroll
          for (int u = 0; u < kGroup; ++u) {
            const bool w = dbc_wall && s_t[j0 + u] == TSPH_WALL;
            ncand += (u < m) && (j0 + u != li) && !w;
            kk[u] = kk[u] && !w;
          }
        }
#pragma unroll
        for (int u = 0; u < kGroup; ++u) {
          my_list[cnt * kWave] = (uint16_t)(j0 + u);
          cnt += kk[u];
        }
        cur += m;
      }
      npair += cnt;
      // pair terms in list order; the next entry's records are loaded while
      // the current one is evaluated
      int jn = my_list[0];
      for (int e = 0; e < cnt; ++e) {
        const int j = jn;
        // read unconditionally (past cnt it is an unused entry of the list)
        jn = my_list[min(e + 1, kList - 1) * kWave];
        const float2 pj = s_pos[j], dpj = s_dp[j];
        const float drx = pi.x - pj.x, dry = pi.y - pj.y;
        const int nt = VAR != TSPH_FORCES_WCSPH ? s_t[j] : 0;
        const float2 nb = VAR == TSPH_FORCES_BT ? s_n[j] : make_float2(0.f, 0.f);
        pair32<VAR>(cc, s, drx, dry, drx * drx + dry * dry, vi, dpi.x, dpi.y, iri, s_vel[j], dpj.x,
                    dpj.y, s_ir[j], nt, nb);
      }
    }
    // candidates: the stencil minus the particle itself (DBC: wall-wall pairs
    // are not candidates either, counted as the reference skips them)
    if (VAR != TSPH_FORCES_DBC) ncand += tot > 0 ? tot - 1 : 0;
    if (CORR) {
      // forces, then the corrector, at once (the corrector's loads make the
      // deferred store of the plain pass pointless)
      if (i < t.own1) {
        double2 a;
        double dr;
        if (act) {
          finish32<VAR>(k, s, at, a, dr);
          d.a[i] = a;
          d.drho[i] = dr;
        } else {
          a = d.a[i];
          dr = d.drho[i];
        }
        correct_particle(d, k, dtc, i, a, dr, state);
      }
    } else {
      // results are stored one chunk later (or after the next tile's staging)
      if (pend_i >= 0) {
        d.a[pend_i] = pend_a;
        d.drho[pend_i] = pend_dr;
      }
      pend_i = -1;
      if (act) {
        finish32<VAR>(k, s, at, pend_a, pend_dr);
        pend_i = i;
      }
    }
  }
  it = next;
  }  // tiles
  if (pend_i >= 0) {
    d.a[pend_i] = pend_a;
    d.drho[pend_i] = pend_dr;
  }
  tq_mark(tq, -1);
  wave_count(npair, ncand, nact, counters);
}

__global__ void k_mixed_records(Soa d, const int *__restrict__ cellof, const int *__restrict__ scal,
                                Consts k0, const Consts *__restrict__ cases, Mixed mx) {
  const int n = scal[S_NP];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  write_mixed_record(d, mx, cellof, case_consts(k0, cases, d.cid, i), i, d.rho[i], d.p[i]);
}


}  // namespace

int launch_mixed_records(tsph_ctx *c) {
  const int nb = 256;
  hipLaunchKernelGGL(k_mixed_records, dim3(grid_for(c->n_ub, nb)), dim3(nb), 0, c->stream, c->cur,
                     c->cellof, c->scal, c->k, c->cases, c->mx);
  TSPH_HIP(hipGetLastError());
  return 0;
}

int launch_cut(tsph_ctx *c, hipStream_t s, int groups, int table) {
  if (groups > c->bound_cap) {
    set_error("tile list: more pooled workgroups than allocated");
    return TSPH_ESTATE;
  }
  if (c->list_cost_table != table) {
    hipLaunchKernelGGL(k_list_costs, dim3(2 * c->num_cu), dim3(256), 0, s, c->cell_start, c->k,
                       c->cases, c->tile_base, c->ncases, c->tile_rows, c->scal,
                       c->scal + S_NSCAL + table, c->tile_list, c->list_cost,
                       cost_estimate(c->prm.precision == TSPH_PREC_MIXED));
    c->list_cost_table = table;
  }
  hipLaunchKernelGGL(k_cut_chunks, dim3(cut_segments(total_tiles(c))), dim3(kCutBlock), 0, s,
                     c->scal, c->list_cost, groups, c->tile_bound);
  TSPH_HIP(hipGetLastError());
  c->list_groups = groups;
  c->list_table = table;
  return 0;
}

int ensure_tile_list(tsph_ctx *c, int groups, int table) {
  const Consts &k = c->k;
  const int rows = c->tile_rows;
  hipStream_t s = c->stream;
  if (!c->tiles_valid) {
    const int ntiles = total_tiles(c);
    const int nb = grid_for(ntiles, kListBlock);
    if (nb > c->list_part_cap) {
      set_error("tile list: more tiles than allocated");
      return TSPH_ESTATE;
    }
    hipLaunchKernelGGL(k_tile_count, dim3(nb), dim3(kListBlock), 0, s, c->cell_start, k, c->cases,
                       c->tile_base, c->ncases, rows, ntiles, c->list_part);
    hipLaunchKernelGGL(k_tile_emit, dim3(nb), dim3(kListBlock), 0, s, c->cell_start, k, c->cases,
                       c->tile_base, c->ncases, rows, ntiles, c->list_part, c->scal, c->tile_list);
    TSPH_HIP(hipGetLastError());
    c->tiles_valid = true;
    c->list_groups = 0;
    c->list_cost_table = -1;
  }
  if (groups > 0 && (groups != c->list_groups || table != c->list_table))
    return launch_cut(c, s, groups, table);
  return 0;
}

int clear_tile_costs(tsph_ctx *c) {
  TSPH_HIP(hipMemsetAsync(c->scal + S_NSCAL, 0,
                          sizeof(int) * kCostTables * ((size_t)c->ncells_tot + 8), c->stream));
  return 0;
}

TileGrid tile_grid(tsph_ctx *c, bool fp32, size_t lds_wave, bool may_pool) {
  TileGrid tg;
  const int waves = c->num_cu * c->waves_per_cu;
  tg.lds_wave = (int)((lds_wave + 15) / 16 * 16);
  tg.pooled = may_pool && walk_balanced(c, fp32, waves);
  int pw = 1;
  if (tg.pooled) {
    // the workgroup size (up to kPoolWaves wavefronts, within one workgroup's
    // LDS) that keeps the most wavefronts of the grid resident on a compute
    // unit at once; ties go to the larger workgroup.  Four 21 KB waves in one
    // workgroup would leave a 160 KB compute unit with one workgroup (4 waves)
    // where 1-wave workgroups give it 7.
    int best = 0;
    for (int p = 1; p <= kPoolWaves; ++p) {
      const int block = kPoolHdr + p * tg.lds_wave;
      if (block > c->max_lds_block) break;
      const int fit = c->lds_cu / block, want = (c->waves_per_cu + p - 1) / p;
      const int resident = (fit < want ? fit : want) * p;
      if (resident >= best) best = resident, pw = p;
    }
  }
  tg.groups = 8 * ((waves + 8 * pw - 1) / (8 * pw));
  tg.grid = dim3(tg.groups);
  tg.block = dim3(kWave * pw);
  tg.lds = pool_hdr(tg.pooled) + (size_t)pw * tg.lds_wave;
  return tg;
}

// LDS bytes per wavefront of the pair pass of `variant` at the current tile rows
static size_t forces_lds_wave(const tsph_ctx *c, int variant) {
  const int cap = 3 * (c->tile_rows + 2) * kCellCap;
  const size_t list = sizeof(uint16_t) * kList * kWave;
  const bool bt = variant == TSPH_FORCES_BT;
  if (c->prm.precision == TSPH_PREC_MIXED)
    return kTabBytes + list +
           (size_t)cap * (3 * sizeof(float2) + sizeof(float) + 2 + (bt ? sizeof(float2) : 0));
  return kTabBytes + list + (size_t)cap * (3 * sizeof(double2) + 2 + (bt ? sizeof(double2) : 0));
}

int pass_groups(tsph_ctx *c, int variant) {
  const TileGrid tg = tile_grid(c, c->prm.precision == TSPH_PREC_MIXED, forces_lds_wave(c, variant));
  return tg.pooled ? tg.groups : 0;
}

namespace {

// one pair pass: the kernel for the variant, the precision and the tile walk
// (pooled or static)
template <int VAR, bool DYN>
void pair_launch(tsph_ctx *c, const TileGrid &tg, int rows, int cap, bool corr, double corr_dt) {
  int *sc = c->scal;
  if (c->prm.precision == TSPH_PREC_MIXED && corr)
    hipLaunchKernelGGL((k_tile32<VAR, DYN, true>), tg.grid, tg.block, tg.lds, c->stream, c->cur,
                       c->mx, c->cell_start, c->cellof, c->k, c->cases, c->tile_base, c->ncases,
                       rows, sc, c->tile_list, cap, c->tile_bound, tg.lds_wave, c->counters,
                       corr_dt, c->seqstate);
  else if (c->prm.precision == TSPH_PREC_MIXED)
    hipLaunchKernelGGL((k_tile32<VAR, DYN, false>), tg.grid, tg.block, tg.lds, c->stream, c->cur,
                       c->mx, c->cell_start, c->cellof, c->k, c->cases, c->tile_base, c->ncases,
                       rows, sc, c->tile_list, cap, c->tile_bound, tg.lds_wave, c->counters, 0.,
                       nullptr);
  else
    hipLaunchKernelGGL((k_tile64<VAR, DYN>), tg.grid, tg.block, tg.lds, c->stream, c->cur,
                       c->cell_start, c->cellof, c->k, c->cases, c->tile_base, c->ncases, rows, sc,
                       c->tile_list, cap, c->tile_bound, tg.lds_wave, c->counters);
}

template <int VAR>
void pair_launch(tsph_ctx *c, const TileGrid &tg, int rows, int cap, bool corr, double corr_dt) {
  if (tg.pooled) pair_launch<VAR, true>(c, tg, rows, cap, corr, corr_dt);
  else pair_launch<VAR, false>(c, tg, rows, cap, corr, corr_dt);
}

}  // namespace

int launch_forces(tsph_ctx *c, int variant, bool corr, double corr_dt) {
  const int rows = c->tile_rows;
  const int cap = 3 * (rows + 2) * kCellCap;
  const bool mixed = c->prm.precision == TSPH_PREC_MIXED;
  if (mixed && !c->records_valid) {
    if (int e = launch_mixed_records(c)) return e;
    c->records_valid = true;
  }
  const TileGrid tg = tile_grid(c, mixed, forces_lds_wave(c, variant));
  if (int e = ensure_tile_list(c, tg.pooled ? tg.groups : 0, mixed && corr ? 1 : 0)) return e;
  const int e = launch_timed(c, [&] {
    switch (variant) {
      case TSPH_FORCES_WCSPH: pair_launch<TSPH_FORCES_WCSPH>(c, tg, rows, cap, corr, corr_dt); break;
      case TSPH_FORCES_DBC: pair_launch<TSPH_FORCES_DBC>(c, tg, rows, cap, corr, corr_dt); break;
      default: pair_launch<TSPH_FORCES_BT>(c, tg, rows, cap, corr, corr_dt);
    }
  });
  if (!e && mixed && corr) {  // as launch_corrector leaves them
    c->records_valid = false;
    c->marks_valid = true;
  }
  return e;
}

}  // namespace tsph
