// Pair loops: the device form of Compute_Forces.
//   TSPH_FORCES_WCSPH  src/SPH_interaction.h:8-149    (active: fluid)
//   TSPH_FORCES_DBC    src/SPH_interactionDBC.h:8-155 (active: all but inlet/outlet,
//                      wall-wall pairs skipped, walls keep dRho but a = 0)
//   TSPH_FORCES_BT     src/SPH_interactionBT.h:8-192  (wall neighbours enter as a
//                      boundary integral with normal n; Shepard gamma normalises)
//
// One thread per particle of the (cell, seq)-sorted arrays, so a 64-wide
// wavefront covers a run of consecutive cells of one grid column and its 9-cell
// stencils overlap almost completely: neighbour records are read coalesced from
// HBM once and then served from L2/L1.  Each thread accumulates its own sums in
// the reference order (stencil cells in the reference sequence, ascending
// reference index inside a cell), so there are no atomics and the FP64 path is
// deterministic.  Candidates beyond the 2h support contribute exact zeros in the
// reference (W = F = 0); they are skipped here after the same qq <= 2 test.
//
// Two arithmetic forms:
//   k_forces64  IEEE double, reference operand order (with -ffp-contract=off)
//   k_forces32  FP32 pair arithmetic on FP32 positions relative to the cell
//               origin (float4 {x, y, vx, vy} + float2 {rho, p} records), FP32
//               accumulation, FP64 outputs for the FP64 integrator.
#include "tsph_internal.h"
#include "tsph_physics.h"

namespace tsph {

namespace {

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

template <int VAR>
__global__ __launch_bounds__(256) void k_forces64(Soa d, const int *__restrict__ cell_start,
                                                  const int *__restrict__ cellof,
                                                  const int *__restrict__ scal, Consts k,
                                                  unsigned long long *__restrict__ counters) {
  const int n = scal[S_NP];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned npair = 0, ncand = 0, nact = 0;
  int at = -1;
  if (i < n) at = d.type[i];
  const bool active = (i < n) && (VAR == TSPH_FORCES_DBC ? (at != TSPH_INLET && at != TSPH_OUTLET)
                                                        : (at == TSPH_FLUID));
  if (active) {
    nact = 1;
    const double h = k.h, m = k.m, c0 = k.cs, delta = k.delta, rho0 = k.rho0, dp = k.dp;
    const double r2cut = 4.0 * h * h * (1.0 + 1e-12);
    const double2 ar = d.r[i], av = d.v[i];
    const double arho = d.rho[i], ap = d.p[i];
    const int c = cellof[i];
    double accx = 0., accy = 0., drho_sum = 0., diff_sum = 0., gamma = 0.;
    for (int s = 0; s < 9; ++s) {
      const int cl = c + stencil_off(s, k.ncy);
      if (cl < 0 || cl >= k.ncells) continue;
      const int j1 = cell_start[cl + 1];
      for (int j = cell_start[cl]; j < j1; ++j) {
        if (j == i) continue;
        int nt = 0;
        if (VAR != TSPH_FORCES_WCSPH) nt = d.type[j];
        if (VAR == TSPH_FORCES_DBC && at == TSPH_WALL && nt == TSPH_WALL) continue;
        ++ncand;
        const double2 nr = d.r[j];
        const double drx = ar.x - nr.x, dry = ar.y - nr.y;
        const double r2 = (drx * drx) + (dry * dry);
        if (r2 > r2cut) continue;
        const double drs = sqrt(r2);
        if (drs / h > 2.0) continue;
        if (drs / h < 2.0) ++npair;
        const KW64 kw = wendland64(drs, h, k.awen, k.bwen);
        const double2 nv = d.v[j];
        const double nrho = d.rho[j], np = d.p[j];
        const double dvx = av.x - nv.x, dvy = av.y - nv.y;
        const double drdv = drx * dvx + dry * dvy;
        const double dWx = drx * kw.f, dWy = dry * kw.f;
        const double dvdW = dvx * dWx + dvy * dWy;
        const double drdW = drx * dWx + dry * dWy;
        const double p_temp = (ap + np) / (arho * nrho);
        const double visco = art_visc64(h, drs, drdv, 0.5 * (arho + nrho), c0, k.avisc);
        if (VAR == TSPH_FORCES_BT && nt == TSPH_WALL) {
          const double2 nb = d.n[j];
          accx += kw.w * (p_temp + visco) * dp * nb.x * nrho;
          accy += kw.w * (p_temp + visco) * dp * nb.y * nrho;
          const double dvnb = dvx * nb.x + dvy * nb.y;
          drho_sum -= dvnb * kw.w * nrho * dp;
        } else {
          const double diff =
              diff_fourtakas64(h, drs, dry, drdW, arho, nrho, rho0, c0, delta, m, k.gy);
          accx -= dWx * (p_temp + visco) * m;
          accy -= dWy * (p_temp + visco) * m;
          drho_sum += dvdW * m;
          diff_sum -= diff;
        }
        if (VAR == TSPH_FORCES_BT && nt == TSPH_FLUID) gamma += kw.w * m / nrho;
      }
    }
    double2 a;
    double dr;
    if (VAR == TSPH_FORCES_BT) {
      if (gamma > kEps) {
        a = make_double2(accx / gamma - k.gx, accy / gamma - k.gy);
        dr = drho_sum / gamma + diff_sum / gamma;
      } else {
        a = make_double2(0. - k.gx, 0. - k.gy);
        dr = 0;
      }
    } else if (VAR == TSPH_FORCES_DBC && at == TSPH_WALL) {
      a = make_double2(0., 0.);
      dr = drho_sum + diff_sum;
    } else {
      a = make_double2(accx - k.gx, accy - k.gy);
      dr = drho_sum + diff_sum;
    }
    d.a[i] = a;
    d.drho[i] = dr;
  }
  wave_count(npair, ncand, nact, counters);
}

// ---------------------------------------------------------------------------
// mixed path
// ---------------------------------------------------------------------------
__global__ void k_mixed_records(Soa d, const int *__restrict__ cellof, const int *__restrict__ scal,
                                Consts k, Mixed mx) {
  const int n = scal[S_NP];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int c = cellof[i];
  const int cx = c / k.ncy, cy = c - cx * k.ncy;
  const double ox = k.x0 - k.of + cx * k.kh, oy = k.y0 - k.of + cy * k.kh;
  const double2 r = d.r[i], v = d.v[i];
  mx.rv[i] = make_float4((float)(r.x - ox), (float)(r.y - oy), (float)v.x, (float)v.y);
  mx.rp[i] = make_float2((float)d.rho[i], (float)d.p[i]);
  if (mx.nrm && d.n) mx.nrm[i] = make_float2((float)d.n[i].x, (float)d.n[i].y);
}

template <int VAR>
__global__ __launch_bounds__(256) void k_forces32(Soa d, Mixed mx, const int *__restrict__ cell_start,
                                                  const int *__restrict__ cellof,
                                                  const int *__restrict__ scal, Consts k,
                                                  unsigned long long *__restrict__ counters) {
#pragma clang fp contract(fast)
  const int n = scal[S_NP];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned npair = 0, ncand = 0, nact = 0;
  int at = -1;
  if (i < n) at = d.type[i];
  const bool active = (i < n) && (VAR == TSPH_FORCES_DBC ? (at != TSPH_INLET && at != TSPH_OUTLET)
                                                        : (at == TSPH_FLUID));
  if (active) {
    nact = 1;
    const float h = (float)k.h, m = (float)k.m, c0 = (float)k.cs, rho0 = (float)k.rho0;
    const float dp = (float)k.dp, kh = (float)k.kh;
    const float ih = 1.f / h;
    const float r2cut = 4.f * h * h;
    const float epsh2 = (float)kEps * h * h;
    const float awen = (float)k.awen, bwen = (float)k.bwen;
    const float avisc = (float)k.avisc;
    const float dcoef = 2.f * c0 * h * (float)k.delta;       // 2 c0 h delta
    const float hx = rho0 * (float)k.gy / (c0 * c0 * rho0 / 7.f);  // rho0 g / cb
    const float4 ai = mx.rv[i];
    const float2 api = mx.rp[i];
    const float arho = api.x, ap = api.y;
    const int c = cellof[i];
    float accx = 0.f, accy = 0.f, drho_sum = 0.f, diff_sum = 0.f, gamma = 0.f;
    for (int s = 0; s < 9; ++s) {
      const int cl = c + stencil_off(s, k.ncy);
      if (cl < 0 || cl >= k.ncells) continue;
      int dcx, dcy;
      stencil_dxy(s, dcx, dcy);
      // position of i relative to the origin of neighbour cell cl
      const float bx = ai.x - (float)dcx * kh, by = ai.y - (float)dcy * kh;
      const int j1 = cell_start[cl + 1];
      for (int j = cell_start[cl]; j < j1; ++j) {
        if (j == i) continue;
        int nt = 0;
        if (VAR != TSPH_FORCES_WCSPH) nt = d.type[j];
        if (VAR == TSPH_FORCES_DBC && at == TSPH_WALL && nt == TSPH_WALL) continue;
        ++ncand;
        const float4 nj = mx.rv[j];
        const float drx = bx - nj.x, dry = by - nj.y;
        const float r2 = drx * drx + dry * dry;
        if (!(r2 < r2cut)) continue;
        ++npair;
        const float2 pj = mx.rp[j];
        const float nrho = pj.x, np = pj.y;
        const float drs = sqrtf(r2);
        const float qq = drs * ih;
        const float wq1 = 1.f - 0.5f * qq;
        const float wq2 = wq1 * wq1;
        // F = bwen qq wq2 wq1 / r = bwen wq2 wq1 / h
        const float f = bwen * wq2 * wq1 * ih;
        const float dvx = ai.z - nj.z, dvy = ai.w - nj.w;
        const float drdv = drx * dvx + dry * dvy;
        const float dWx = drx * f, dWy = dry * f;
        const float inv_rr = __frcp_rn(arho * nrho);
        const float p_temp = (ap + np) * inv_rr;
        const float rinv = __frcp_rn(r2 + epsh2);
        float visco = 0.f;
        if (drdv < 0.f) {
          const float vmax = drdv > avisc ? drdv : avisc;
          // -visco c0 (h drdv / (r^2 + eps h^2)) / (0.5 (rho_i + rho_j))
          visco = -vmax * c0 * (h * drdv * rinv) * (2.f * __frcp_rn(arho + nrho));
        }
        if (VAR == TSPH_FORCES_BT && nt == TSPH_WALL) {
          const float w = awen * (qq + qq + 1.f) * wq2 * wq2;
          const float2 nb = mx.nrm[j];
          const float t = w * (p_temp + visco) * dp * nrho;
          accx += t * nb.x;
          accy += t * nb.y;
          drho_sum -= (dvx * nb.x + dvy * nb.y) * w * nrho * dp;
        } else {
          const float dvdW = dvx * dWx + dvy * dWy;
          const float drdW = drx * dWx + dry * dWy;
          const float drhoH = hydro_drho_f32(dry * hx, rho0);
          const float psi = dcoef * ((nrho - arho) - drhoH) * rinv;
          const float t = (p_temp + visco) * m;
          accx -= dWx * t;
          accy -= dWy * t;
          drho_sum += dvdW * m;
          diff_sum -= psi * drdW * m * __frcp_rn(nrho);
          if (VAR == TSPH_FORCES_BT && nt == TSPH_FLUID) {
            const float w = awen * (qq + qq + 1.f) * wq2 * wq2;
            gamma += w * m * __frcp_rn(nrho);
          }
        }
      }
    }
    double2 a;
    double dr;
    if (VAR == TSPH_FORCES_BT) {
      if (gamma > (float)kEps) {
        a = make_double2((double)(accx / gamma) - k.gx, (double)(accy / gamma) - k.gy);
        dr = (double)(drho_sum / gamma + diff_sum / gamma);
      } else {
        a = make_double2(0. - k.gx, 0. - k.gy);
        dr = 0;
      }
    } else if (VAR == TSPH_FORCES_DBC && at == TSPH_WALL) {
      a = make_double2(0., 0.);
      dr = (double)(drho_sum + diff_sum);
    } else {
      a = make_double2((double)accx - k.gx, (double)accy - k.gy);
      dr = (double)(drho_sum + diff_sum);
    }
    d.a[i] = a;
    d.drho[i] = dr;
  }
  wave_count(npair, ncand, nact, counters);
}

__global__ void k_count_pass(unsigned long long *counters) {
  if (threadIdx.x == 0 && blockIdx.x == 0) counters[C_PASSES] += 1ull;
}

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

}  // namespace

int launch_mixed_records(tsph_ctx *c) {
  const int nb = 256;
  hipLaunchKernelGGL(k_mixed_records, dim3(grid_for(c->n_ub, nb)), dim3(nb), 0, c->stream, c->cur,
                     c->cellof, c->scal, c->k, c->mx);
  TSPH_HIP(hipGetLastError());
  return 0;
}

int launch_forces(tsph_ctx *c, int variant) {
  const int nb = c->block;
  const dim3 g(grid_for(c->n_ub, nb)), b(nb);
  hipStream_t s = c->stream;
  int e = 0;
  if (c->prm.precision == TSPH_PREC_MIXED) {
    if ((e = launch_mixed_records(c))) return e;
    e = launch_timed(c, [&] {
      switch (variant) {
        case TSPH_FORCES_WCSPH:
          hipLaunchKernelGGL(k_forces32<TSPH_FORCES_WCSPH>, g, b, 0, s, c->cur, c->mx, c->cell_start,
                             c->cellof, c->scal, c->k, c->counters);
          break;
        case TSPH_FORCES_DBC:
          hipLaunchKernelGGL(k_forces32<TSPH_FORCES_DBC>, g, b, 0, s, c->cur, c->mx, c->cell_start,
                             c->cellof, c->scal, c->k, c->counters);
          break;
        default:
          hipLaunchKernelGGL(k_forces32<TSPH_FORCES_BT>, g, b, 0, s, c->cur, c->mx, c->cell_start,
                             c->cellof, c->scal, c->k, c->counters);
      }
    });
  } else {
    e = launch_timed(c, [&] {
      switch (variant) {
        case TSPH_FORCES_WCSPH:
          hipLaunchKernelGGL(k_forces64<TSPH_FORCES_WCSPH>, g, b, 0, s, c->cur, c->cell_start,
                             c->cellof, c->scal, c->k, c->counters);
          break;
        case TSPH_FORCES_DBC:
          hipLaunchKernelGGL(k_forces64<TSPH_FORCES_DBC>, g, b, 0, s, c->cur, c->cell_start,
                             c->cellof, c->scal, c->k, c->counters);
          break;
        default:
          hipLaunchKernelGGL(k_forces64<TSPH_FORCES_BT>, g, b, 0, s, c->cur, c->cell_start,
                             c->cellof, c->scal, c->k, c->counters);
      }
    });
  }
  if (e) return e;
  hipLaunchKernelGGL(k_count_pass, dim3(1), dim3(64), 0, s, c->counters);
  TSPH_HIP(hipGetLastError());
  return 0;
}

}  // namespace tsph
