// Riemann-ALE-SPH, first order: the pipeline the ALE dam-break driver runs
// (cases/ALESPH_dambreakLong/run.sh -> SPH_simulationINT.h:84-132):
//   Integrate_LeapFrog_partOne  src/RALESPH_integration.h:10-36
//   Compute_Forces              src/RALESPH_interaction.h:12-287
//   Integrate_LeapFrog_partTwo  src/RALESPH_integration.h:38-107
// with the particle volume variables of ALESPH_Add_particle
// (src/SPH_particle_system.h:51-108).
//
// The pair loop reuses the tile machinery of the WCSPH loops (persistent
// waves over XCD-banded tiles, neighbour slices staged in LDS, lane-private
// neighbour lists).  Per pair it solves the linearised Riemann problem along
// the pair normal (src/ALESPH_riemann_solver.h:20-34, 86-96), with a
// one-sided wall state for BT wall neighbours (src/RALESPH_interaction.h:217-253).
//
// Reference quirk kept: gamma_sum is declared per cell and never reset between
// the cell's particles (src/RALESPH_interaction.h:48, 257, 277-280), so a fluid
// particle with no neighbour in its support takes the "gamma == 0" (all-zero)
// branch only if no earlier fluid particle of its cell had one.  The pair loop
// records whether each particle's own sum is non-zero; k_rale_isolated then
// settles the particles whose own sum is zero by looking at the earlier
// particles of their cell (reference order inside a cell).
#include "tsph_internal.h"
#include "tsph_physics.h"
#include "tsph_tiles.h"

namespace tsph {

namespace {

using namespace tiles;

struct AccA64 {
  double dom = 0, domrho = 0, dovx = 0, dovy = 0, gamma = 0;
};

// one RALE pair in the reference operand order
__device__ __forceinline__ void rale_pair64(const Consts &k, AccA64 &s, double2 ar, double2 av,
                                            double arho, double2 nr, double2 nv, double nrho,
                                            double no, int nt, double2 n) {
  const double h = k.h, c0 = k.cs, dp = k.dp, m = k.m, rho0 = k.rho0;
  const double drx = ar.x - nr.x, dry = ar.y - nr.y;
  const double drs = sqrt((drx * drx) + (dry * dry));
  const KW64 kw = wendland64(drs, h, k.awen, k.bwen);
  const double W = kw.w;
  const double dWx = drx * kw.f, dWy = dry * kw.f;
  const double dvx = av.x - nv.x, dvy = av.y - nv.y;
  const double dvdW = dvx * dWx + dvy * dWy;
  const double inv = 1 / drs;
  const double drnx = drx * inv * (-1), drny = dry * inv * (-1);
  const double vL = av.x * drnx + av.y * drny;
  const double vR = nv.x * drnx + nv.y * drny;
  const double avgrho = 0.5 * (arho + nrho);
  const double vss = 0.5 * (vR + vL) - 0.5 * (nrho - arho) * c0 / avgrho;
  const double dd = vL - vR;
  const double mx = dd > 0 ? dd : 0;
  const double e = 3. * mx;
  const double beta = e < c0 ? e : c0;
  const double rhos = 0.5 * (arho + nrho) + 0.5 * beta * (vL - vR) * (0.5 * (arho + nrho)) / (c0 * c0);
  const double vsx = drnx * vss + ((av.x + nv.x) * 0.5 - drnx * (vL + vR) * 0.5);
  const double vsy = drny * vss + ((av.y + nv.y) * 0.5 - drny * (vL + vR) * 0.5);
  if (nt == TSPH_WALL) {
    const double nbx = n.x * (-1), nby = n.y * (-1);
    const double rw = arho - ((0 - av.x) * nbx + (0 - av.y) * nby) * arho / c0;
    const double pw = c0 * c0 * (rw - rho0) * 1;
    s.dovx -= 2 * W * dp * (pw * nbx);
    s.dovy -= 2 * W * dp * (pw * nby);
    const double dvn = dvx * nbx + dvy * nby;
    s.dom -= dvn * W * dp;
  } else {
    const double ps = tait64(rhos, rho0, c0);
    s.dovx -= 2 * no * ((rhos * vsx * (vsx - av.x) + ps) * dWx + (rhos * vsx * (vsy - av.y)) * dWy);
    s.dovy -= 2 * no * ((rhos * vsy * (vsx - av.x)) * dWx + (rhos * vsy * (vsy - av.y) + ps) * dWy);
    const double vsdW = (vsx - av.x) * dWx + (vsy - av.y) * dWy;
    s.dom -= dvdW * no;
    s.domrho -= 2 * rhos * no * vsdW;
  }
  s.gamma += W * m / nrho;
}

// outputs for a particle whose own gamma sum is non-zero (:271-273), or the
// all-zero branch (:265-267) pending k_rale_isolated
__device__ __forceinline__ void rale_finish(const Soa &d, const Consts &k, int i, double ao,
                                           double arho, double dom, double domrho, double dovx,
                                           double dovy, bool nonzero, uint8_t *gflag) {
  if (nonzero) {
    d.domega[i] = dom * ao;
    d.domegarho[i] = domrho * ao;
    d.domegav[i] = make_double2(dovx * ao - k.gx * ao * arho, dovy * ao - k.gy * ao * arho);
  } else {
    d.domega[i] = 0.;
    d.domegarho[i] = 0.;
    d.domegav[i] = make_double2(0., 0.);
  }
  gflag[i] = nonzero;
}

// FP64 tile kernel, reference summation order (9 cells, ascending index)
template <bool DYN>
__global__ __launch_bounds__(kWave * kPoolWaves) void k_rale64(Soa d, const int *__restrict__ cell_start,
                                                  const int *__restrict__ cellof, Consts k, int rows,
                                                  int *__restrict__ scal,
                                                  const int *__restrict__ tile_list, int cap,
                                                  const int *__restrict__ bound, int lds_wave,
                                                  uint8_t *__restrict__ gflag,
                                                  unsigned long long *__restrict__ counters) {
  extern __shared__ __align__(16) unsigned char smem[];
  unsigned char *wl = wave_lds<DYN>(smem, lds_wave);
  int *tab = reinterpret_cast<int *>(wl);
  double2 *s_r = reinterpret_cast<double2 *>(wl + kTabBytes);
  double2 *s_v = s_r + cap;
  double2 *s_rw = s_v + cap;  // {rho, omega}
  double2 *s_n = s_rw + cap;
  uint16_t *s_list = reinterpret_cast<uint16_t *>(s_n + cap);
  uint8_t *s_t = reinterpret_cast<uint8_t *>(s_list + kList * kWave);

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
    if (staged) {
      for (int l = tile_lane(); l < t.lq[3]; l += kWave) {
        const int j = l - sel3(t.base, slice_of(t, l));
        s_r[l] = d.r[j];
        s_v[l] = d.v[j];
        s_rw[l] = make_double2(d.rho[j], d.omega[j]);
        s_t[l] = d.type[j];
        s_n[l] = d.n ? d.n[j] : make_double2(0., 0.);
      }
    }
    wave_sync();
    const int next = tq_take<DYN>(tq, ticket);
    const int lane = tile_lane();
    uint16_t *my_list = s_list + lane;
    for (int i0 = t.own0; i0 < t.own1; i0 += kWave) {
      const int i = i0 + lane;
      int li = 0, cy = 0;
      bool act = false;
      if (i < t.own1) {
        li = i + t.base[1];
        act = d.type[i] == TSPH_FLUID;
        cy = cellof[i] - t.cx * k.ncy;
      }
      nact += act;
      const double2 ar = act ? d.r[i] : make_double2(0., 0.);
      const double2 av = act ? d.v[i] : make_double2(0., 0.);
      const double arho = act ? d.rho[i] : 1.;
      const double ao = act ? d.omega[i] : 0.;
      AccA64 s;
      if (!staged) {
        // neighbourhood beyond the LDS budget: gather from global memory
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
              const double2 nr = d.r[j];
              if (!in_support64(ar.x - nr.x, ar.y - nr.y, k.h)) continue;
              ++npair;
              rale_pair64(k, s, ar, av, arho, nr, d.v[j], d.rho[j], d.omega[j], d.type[j],
                          d.n ? d.n[j] : make_double2(0., 0.));
            }
          }
        }
      } else {
        int q = -1, lj = 0, lend = 0;
        bool more = act;
        while (__any(more)) {
          int cnt = 0;
          while (more && cnt < kList) {
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
            const int j = lj++;
            if (j == li) continue;
            ++ncand;
            const double2 nr = s_r[j];
            if (in_support64(ar.x - nr.x, ar.y - nr.y, k.h)) my_list[(cnt++) * kWave] = (uint16_t)j;
          }
          npair += cnt;
          for (int e = 0; e < cnt; ++e) {
            const int j = my_list[e * kWave];
            const double2 rw = s_rw[j];
            rale_pair64(k, s, ar, av, arho, s_r[j], s_v[j], rw.x, rw.y, s_t[j], s_n[j]);
          }
        }
      }
      if (act) rale_finish(d, k, i, ao, arho, s.dom, s.domrho, s.dovx, s.dovy, s.gamma != 0, gflag);
    }
    it = next;
  }
  tq_mark(tq, -1);
  wave_count(npair, ncand, nact, counters);
}

// --------------------------------------------------------------------------
// mixed form: FP32 fluxes (positions staged relative to the tile, computed in
// FP64 and rounded once; density as rho - rho0), FP64 outputs
// --------------------------------------------------------------------------
struct AccA32 {
  float dom = 0, domrho = 0, dovx = 0, dovy = 0, gamma = 0;
};

struct CA32 {
  float h, ih, m, c0, ic0, rho0, dp, r2cut, awen, bwen, btait;
};

__device__ __forceinline__ void rale_pair32(const CA32 &c, AccA32 &s, float drx, float dry, float r2,
                                            float2 vi, float adev, float2 vj, float ndev, float no,
                                            int nt, float2 n) {
#pragma clang fp contract(fast)
  const float drs = __builtin_amdgcn_sqrtf(r2);
  const float inv = __builtin_amdgcn_rcpf(drs);
  const float qq = drs * c.ih;
  const float wq1 = 1.f - 0.5f * qq;
  const float wq2 = wq1 * wq1;
  const float W = c.awen * (qq + qq + 1.f) * wq2 * wq2;
  const float f = c.bwen * wq2 * wq1 * c.ih;
  const float dWx = drx * f, dWy = dry * f;
  const float dvx = vi.x - vj.x, dvy = vi.y - vj.y;
  const float dvdW = dvx * dWx + dvy * dWy;
  const float drnx = -drx * inv, drny = -dry * inv;
  const float vL = vi.x * drnx + vi.y * drny;
  const float vR = vj.x * drnx + vj.y * drny;
  const float arho = c.rho0 + adev, nrho = c.rho0 + ndev;
  const float avgrho = c.rho0 + 0.5f * (adev + ndev);
  const float vss = 0.5f * (vR + vL) - 0.5f * (ndev - adev) * c.c0 * __builtin_amdgcn_rcpf(avgrho);
  const float dd = vL - vR;
  const float beta = fminf(3.f * fmaxf(dd, 0.f), c.c0);
  const float rsdev = 0.5f * (adev + ndev) + 0.5f * beta * dd * avgrho * (c.ic0 * c.ic0);
  const float vsx = drnx * vss + (0.5f * (vi.x + vj.x) - 0.5f * drnx * (vL + vR));
  const float vsy = drny * vss + (0.5f * (vi.y + vj.y) - 0.5f * drny * (vL + vR));
  if (nt == TSPH_WALL) {
    const float nbx = -n.x, nby = -n.y;
    // wall state rho_w - rho0 = (rho_i - rho0) + (v_i . nb) rho_i / c0, linear EOS
    const float pw = c.c0 * c.c0 * (adev + (vi.x * nbx + vi.y * nby) * arho * c.ic0);
    s.dovx -= 2.f * W * c.dp * (pw * nbx);
    s.dovy -= 2.f * W * c.dp * (pw * nby);
    s.dom -= (dvx * nbx + dvy * nby) * W * c.dp;
  } else {
    // Tait from the density deviation: B((1+x)^7 - 1), binomial expansion
    const float x = rsdev * (1.f / c.rho0);
    const float ps = c.btait * x * (7.f + x * (21.f + x * (35.f + x * (35.f + x * (21.f + x * (7.f + x))))));
    const float rhos = c.rho0 + rsdev;
    const float ex = vsx - vi.x, ey = vsy - vi.y;
    s.dovx -= 2.f * no * ((rhos * vsx * ex + ps) * dWx + (rhos * vsx * ey) * dWy);
    s.dovy -= 2.f * no * ((rhos * vsy * ex) * dWx + (rhos * vsy * ey + ps) * dWy);
    s.dom -= dvdW * no;
    s.domrho -= 2.f * rhos * no * (ex * dWx + ey * dWy);
  }
  s.gamma += W * c.m * __builtin_amdgcn_rcpf(nrho);
}

template <bool DYN>
__global__ __launch_bounds__(kWave * kPoolWaves) void k_rale32(Soa d, const int *__restrict__ cell_start,
                                                  const int *__restrict__ cellof, Consts k, int rows,
                                                  int *__restrict__ scal,
                                                  const int *__restrict__ tile_list, int cap,
                                                  const int *__restrict__ bound, int lds_wave,
                                                  uint8_t *__restrict__ gflag,
                                                  unsigned long long *__restrict__ counters) {
#pragma clang fp contract(fast)
  extern __shared__ __align__(16) unsigned char smem[];
  unsigned char *wl = wave_lds<DYN>(smem, lds_wave);
  int *tab = reinterpret_cast<int *>(wl);
  float2 *s_pos = reinterpret_cast<float2 *>(wl + kTabBytes);
  float2 *s_vel = s_pos + cap;
  float2 *s_dw = s_vel + cap;  // {rho - rho0, omega}
  float2 *s_n = s_dw + cap;
  uint16_t *s_list = reinterpret_cast<uint16_t *>(s_n + cap);
  uint8_t *s_t = reinterpret_cast<uint8_t *>(s_list + kList * kWave);

  CA32 c;
  c.h = (float)k.h; c.ih = 1.f / c.h; c.m = (float)k.m; c.c0 = (float)k.cs; c.ic0 = 1.f / c.c0;
  c.rho0 = (float)k.rho0; c.dp = (float)k.dp; c.r2cut = 4.f * c.h * c.h;
  c.awen = (float)k.awen; c.bwen = (float)k.bwen; c.btait = (float)(k.cs * k.cs * k.rho0 / 7.);

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
    if (staged) {
      for (int l = tile_lane(); l < t.lq[3]; l += kWave) {
        const int j = l - sel3(t.base, slice_of(t, l));
        const double2 r = d.r[j], v = d.v[j];
        s_pos[l] = make_float2((float)(r.x - ox), (float)(r.y - oy));
        s_vel[l] = make_float2((float)v.x, (float)v.y);
        s_dw[l] = make_float2((float)(d.rho[j] - k.rho0), (float)d.omega[j]);
        s_t[l] = d.type[j];
        s_n[l] = d.n ? make_float2((float)d.n[j].x, (float)d.n[j].y) : make_float2(0.f, 0.f);
      }
    }
    wave_sync();
    const int next = tq_take<DYN>(tq, ticket);
    const int lane = tile_lane();
    uint16_t *my_list = s_list + lane;
    for (int i0 = t.own0; i0 < t.own1; i0 += kWave) {
      const int i = i0 + lane;
      int li = 0, ri = 0;
      bool act = false;
      if (i < t.own1) {
        li = i + t.base[1];
        act = d.type[i] == TSPH_FLUID;
        ri = cellof[i] - t.cx * k.ncy - t.row_lo;
      }
      nact += act;
      const double2 ar = act ? d.r[i] : make_double2(0., 0.);
      const double2 av = act ? d.v[i] : make_double2(0., 0.);
      const float2 pi = make_float2((float)(ar.x - ox), (float)(ar.y - oy));
      const float2 vi = make_float2((float)av.x, (float)av.y);
      const double arho = act ? d.rho[i] : 1.;
      const float adev = (float)(arho - k.rho0);
      AccA32 s;
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
                const double2 r = d.r[j];
                const float drx = pi.x - (float)(r.x - ox), dry = pi.y - (float)(r.y - oy);
                const float r2 = drx * drx + dry * dry;
                if (!(r2 < c.r2cut)) continue;
                ++npair;
                const double2 v = d.v[j];
                const float2 nn = d.n ? make_float2((float)d.n[j].x, (float)d.n[j].y) : make_float2(0.f, 0.f);
                rale_pair32(c, s, drx, dry, r2, vi, adev, make_float2((float)v.x, (float)v.y),
                            (float)(d.rho[j] - k.rho0), (float)d.omega[j], d.type[j], nn);
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
        int cur = 0;
        while (__any(cur < tot)) {
          int cnt = 0;
          while (cur < tot && cnt < kList) {
            const int j = cur < n0 ? a0 + cur : (cur < n0 + n1 ? a1 + (cur - n0) : a2 + (cur - n0 - n1));
            ++cur;
            if (j == li) continue;
            ++ncand;
            const float2 pj = s_pos[j];
            const float drx = pi.x - pj.x, dry = pi.y - pj.y;
            if (drx * drx + dry * dry < c.r2cut) my_list[(cnt++) * kWave] = (uint16_t)j;
          }
          npair += cnt;
          for (int e = 0; e < cnt; ++e) {
            const int j = my_list[e * kWave];
            const float2 pj = s_pos[j], dw = s_dw[j];
            const float drx = pi.x - pj.x, dry = pi.y - pj.y;
            rale_pair32(c, s, drx, dry, drx * drx + dry * dry, vi, adev, s_vel[j], dw.x, dw.y, s_t[j],
                        s_n[j]);
          }
        }
      }
      if (act) {
        const double ao = d.omega[i];
        rale_finish(d, k, i, ao, arho, (double)s.dom, (double)s.domrho, (double)s.dovx,
                    (double)s.dovy, s.gamma != 0.f, gflag);
      }
    }
    it = next;
  }
  tq_mark(tq, -1);
  wave_count(npair, ncand, nact, counters);
}

// settles fluid particles without neighbours: the reference's cell-wide
// gamma_sum is non-zero for them iff an earlier fluid particle of the same cell
// had neighbours (src/RALESPH_interaction.h:263-274)
__global__ void k_rale_isolated(Soa d, const int *__restrict__ cell_start,
                                const int *__restrict__ cellof, const int *__restrict__ scal,
                                const uint8_t *__restrict__ gflag, Consts k) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= scal[S_NP] || d.type[i] != TSPH_FLUID || gflag[i]) return;
  const int c = cellof[i];
  bool earlier = false;
  for (int j = cell_start[c]; j < i; ++j) earlier |= d.type[j] == TSPH_FLUID && gflag[j];
  if (!earlier) return;
  const double ao = d.omega[i], arho = d.rho[i];
  const double z = 0.;
  d.domega[i] = z * ao;
  d.domegarho[i] = z * ao;
  d.domegav[i] = make_double2(z * ao - k.gx * ao * arho, z * ao - k.gy * ao * arho);
}

// ALESPH_Add_particle's volume variables: omega = dp^2, omega*rho, omega*v
__global__ void k_ale_init(Soa d, const int *__restrict__ scal, double dp) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= scal[S_NP]) return;
  const double w = dp * dp;
  const double2 v = d.v[i];
  const double orho = d.rho[i] * (dp * dp);
  const double2 ov = make_double2(v.x * (dp * dp), v.y * (dp * dp));
  d.omega[i] = w;
  d.omegarho[i] = orho;
  d.omegav[i] = ov;
  d.omega_o[i] = w;
  d.omegarho_o[i] = orho;
  d.omegav_o[i] = ov;
  d.domega[i] = 0.;
  d.domegarho[i] = 0.;
  d.domegav[i] = make_double2(0., 0.);
  d.gradrho[i] = d.gradvx[i] = d.gradvy[i] = make_double2(0., 0.);
}

// Integrate_LeapFrog_partOne, src/RALESPH_integration.h:10-36
__global__ void k_rale_lf1(Soa d, const int *__restrict__ scal, double dt) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= scal[S_NP] || d.type[p] != TSPH_FLUID) return;
  const double om = d.omega[p], orho = d.omegarho[p];
  const double2 ov = d.omegav[p], dov = d.domegav[p];
  d.omega_o[p] = om;
  d.omegarho_o[p] = orho;
  d.omegav_o[p] = ov;
  const double om1 = om + d.domega[p] * dt * 0.5;
  const double orho1 = orho + d.domegarho[p] * dt * 0.5;
  const double2 ov1 = make_double2(ov.x + dov.x * dt * 0.5, ov.y + dov.y * dt * 0.5);
  d.omega[p] = om1;
  d.omegarho[p] = orho1;
  d.omegav[p] = ov1;
  const double rho = orho1 / (om1);
  d.rho[p] = rho;
  d.v[p] = make_double2(ov1.x / (rho * om1), ov1.y / (rho * om1));
}

// Integrate_LeapFrog_partTwo, src/RALESPH_integration.h:38-107
__global__ void k_rale_lf2(Soa d, const int *__restrict__ scal, double dt, int first) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= scal[S_NP] || d.type[p] != TSPH_FLUID) return;
  const double2 dov = d.domegav[p];
  if (first) {
    const double om = d.omega[p], orho = d.omegarho[p];
    const double2 ov = d.omegav[p];
    d.omega_o[p] = om;
    d.omegarho_o[p] = orho;
    d.omegav_o[p] = ov;
    const double om1 = om + d.domega[p] * dt * 0.5;
    const double orho1 = orho + d.domegarho[p] * dt * 0.5;
    const double2 ov1 = make_double2(ov.x + dov.x * dt * 0.5, ov.y + dov.y * dt * 0.5);
    d.omega[p] = om1;
    d.omegarho[p] = orho1;
    d.omegav[p] = ov1;
    const double rho = orho1 / (om1);
    d.rho[p] = rho;
    const double2 v = make_double2(ov1.x / (rho * om1), ov1.y / (rho * om1));
    d.v[p] = v;
    const double2 r = d.r[p];
    d.r[p] = make_double2(r.x + v.x * dt * 0.5, r.y + v.y * dt * 0.5);
  } else {
    const double2 r = d.r[p], v0 = d.v[p];
    d.r[p] = make_double2(r.x + v0.x * dt, r.y + v0.y * dt);
    const double om1 = d.omega_o[p] + d.domega[p] * dt;
    const double orho1 = d.omegarho_o[p] + d.domegarho[p] * dt;
    const double2 ovo = d.omegav_o[p];
    const double2 ov1 = make_double2(ovo.x + dov.x * dt, ovo.y + dov.y * dt);
    d.omega[p] = om1;
    d.omegarho[p] = orho1;
    d.omegav[p] = ov1;
    const double rho = orho1 / (om1);
    d.rho[p] = rho;
    d.v[p] = make_double2(ov1.x / (rho * om1), ov1.y / (rho * om1));
  }
}

}  // namespace

int launch_ale_init(tsph_ctx *c) {
  hipLaunchKernelGGL(k_ale_init, dim3(grid_for(c->n_ub, 256)), dim3(256), 0, c->stream, c->cur,
                     c->scal, c->k.dp);
  TSPH_HIP(hipGetLastError());
  return 0;
}

int launch_rale_forces(tsph_ctx *c) {
  const Consts &k = c->k;
  const int rows = c->tile_rows;
  const int cap = 3 * (rows + 2) * kCellCap;
  hipStream_t s = c->stream;
  const size_t list = sizeof(uint16_t) * kList * kWave;
  int e = 0;
  if (c->prm.precision == TSPH_PREC_MIXED) {
    // the ALE passes stride (see launch_pass, tsph_alemuscl.hip)
    const TileGrid tg =
        tile_grid(c, true, kTabBytes + list + (size_t)cap * (4 * sizeof(float2) + 1), false);
    if ((e = ensure_tile_list(c, tg.pooled ? tg.groups : 0))) return e;
    e = launch_timed(c, [&] {
      if (tg.pooled)
        hipLaunchKernelGGL(k_rale32<true>, tg.grid, tg.block, tg.lds, s, c->cur, c->cell_start,
                           c->cellof, k, rows, c->scal, c->tile_list, cap, c->tile_bound,
                           tg.lds_wave, c->ale_flag, c->counters);
      else
        hipLaunchKernelGGL(k_rale32<false>, tg.grid, tg.block, tg.lds, s, c->cur, c->cell_start,
                           c->cellof, k, rows, c->scal, c->tile_list, cap, c->tile_bound,
                           tg.lds_wave, c->ale_flag, c->counters);
    });
  } else {
    const TileGrid tg =
        tile_grid(c, false, kTabBytes + list + (size_t)cap * (4 * sizeof(double2) + 1), false);
    if ((e = ensure_tile_list(c, tg.pooled ? tg.groups : 0))) return e;
    e = launch_timed(c, [&] {
      if (tg.pooled)
        hipLaunchKernelGGL(k_rale64<true>, tg.grid, tg.block, tg.lds, s, c->cur, c->cell_start,
                           c->cellof, k, rows, c->scal, c->tile_list, cap, c->tile_bound,
                           tg.lds_wave, c->ale_flag, c->counters);
      else
        hipLaunchKernelGGL(k_rale64<false>, tg.grid, tg.block, tg.lds, s, c->cur, c->cell_start,
                           c->cellof, k, rows, c->scal, c->tile_list, cap, c->tile_bound,
                           tg.lds_wave, c->ale_flag, c->counters);
    });
  }
  if (e) return e;
  hipLaunchKernelGGL(k_rale_isolated, dim3(grid_for(c->n_ub, 256)), dim3(256), 0, s, c->cur,
                     c->cell_start, c->cellof, c->scal, c->ale_flag, k);
  TSPH_HIP(hipGetLastError());
  return 0;
}

int launch_rale_leapfrog_one(tsph_ctx *c, double dt) {
  hipLaunchKernelGGL(k_rale_lf1, dim3(grid_for(c->n_ub, 256)), dim3(256), 0, c->stream, c->cur,
                     c->scal, dt);
  TSPH_HIP(hipGetLastError());
  c->records_valid = false;
  c->marks_valid = false;
  return 0;
}

int launch_rale_leapfrog_two(tsph_ctx *c, double dt, int step) {
  hipLaunchKernelGGL(k_rale_lf2, dim3(grid_for(c->n_ub, 256)), dim3(256), 0, c->stream, c->cur,
                     c->scal, dt, (int)(step == 1));
  TSPH_HIP(hipGetLastError());
  c->records_valid = false;
  c->marks_valid = false;
  return 0;
}

}  // namespace tsph
