// The reference's other WCSPH time integrators and its density
// reinitialisation, for the drivers that use them:
//   Integrate_Verlet / Integrate_Euler          src/SPH_integration.h:103-158
//     (cases/*/SPH_simulationVerlet.h: Verlet, with an Euler step at step 1
//     and every 20th step)
//   Integrate_LeapFrog_partOne/partTwo_withDensity  src/SPH_integration.h:161-230
//     (cases/*/SPH_simulationLeapFrog.h)
//   Density_reinitialization_sumation / _overwrite  src/SPH_density_reinitialization.h:8-139
//     (cases/damBreakLong_mDBC/SPH_simulationSymplectic_reinit.h: every 20th step)
// Verlet keeps a second time layer (v_oo, rho_oo) and the reinitialisation a
// per-particle buffer (rhoTemp); both move with the particles through every
// reorder once allocated.  All passes are per-particle HBM streams except the
// reinitialisation sum, a 9-cell FP64 gather in the reference order (it runs
// every 20th step).
#include "tsph_internal.h"
#include "tsph_physics.h"

namespace tsph {

namespace {

// Add_particle's second layer and buffer: v_oo = v, rho_oo = rho, rhoTemp = rho
// (src/SPH_particle_system.h:40-47)
__global__ void k_history_init(Soa d, const int *__restrict__ scal) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= scal[S_NP]) return;
  const double rho = d.rho[p];
  d.v_oo[p] = d.v[p];
  d.rho_oo[p] = rho;
  d.rho_tmp[p] = rho;
}

// src/SPH_integration.h:103-128 (all particle types)
__global__ void k_verlet(Soa d, const int *__restrict__ scal, double dt) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= scal[S_NP]) return;
  const double2 a = d.a[p], vo = d.v_o[p], voo = d.v_oo[p], r = d.r[p];
  const double2 v = make_double2(voo.x + a.x * dt * 2, voo.y + a.y * dt * 2);
  d.v[p] = v;
  d.r[p] = make_double2(r.x + vo.x * dt + a.x * dt * dt * 0.5, r.y + vo.y * dt + a.y * dt * dt * 0.5);
  const double rho = d.rho_oo[p] + d.drho[p] * dt * 2;
  d.rho[p] = rho;
  d.v_oo[p] = vo;
  d.v_o[p] = v;
  d.rho_oo[p] = d.rho_o[p];
  d.rho_o[p] = rho;
}

// src/SPH_integration.h:130-158 (all particle types)
__global__ void k_euler(Soa d, const int *__restrict__ scal, double dt) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= scal[S_NP]) return;
  const double2 a = d.a[p], vo = d.v_o[p], v0 = d.v[p], r = d.r[p];
  const double2 v = make_double2(v0.x + a.x * dt, v0.y + a.y * dt);
  d.v[p] = v;
  d.r[p] = make_double2(r.x + vo.x * dt + a.x * dt * dt * 0.5, r.y + vo.y * dt + a.y * dt * dt * 0.5);
  const double rho = d.rho[p] + d.drho[p] * dt;
  d.rho[p] = rho;
  d.v_oo[p] = vo;
  d.v_o[p] = v;
  d.rho_oo[p] = d.rho_o[p];
  d.rho_o[p] = rho;
}

// src/SPH_integration.h:161-183 (all particle types)
__global__ void k_lf1_density(Soa d, const int *__restrict__ scal, double dt) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= scal[S_NP]) return;
  const double2 v = d.v[p], a = d.a[p];
  const double rho = d.rho[p];
  d.v_o[p] = v;
  d.rho_o[p] = rho;
  d.v[p] = make_double2(v.x + a.x * dt * 0.5, v.y + a.y * dt * 0.5);
  d.rho[p] = rho + d.drho[p] * dt * 0.5;
}

// src/SPH_integration.h:185-230 (all particle types)
__global__ void k_lf2_density(Soa d, const int *__restrict__ scal, double dt, int first) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= scal[S_NP]) return;
  const double2 a = d.a[p], r = d.r[p];
  if (first) {
    const double2 v0 = d.v[p];
    const double2 v = make_double2(v0.x + a.x * dt * 0.5, v0.y + a.y * dt * 0.5);
    d.v[p] = v;
    d.r[p] = make_double2(r.x + v.x * dt * 0.5, r.y + v.y * dt * 0.5);
  } else {
    const double2 v0 = d.v[p], vo = d.v_o[p];
    d.r[p] = make_double2(r.x + v0.x * dt, r.y + v0.y * dt);
    d.v[p] = make_double2(vo.x + a.x * dt, vo.y + a.y * dt);
    d.rho[p] = d.rho_o[p] + d.drho[p] * dt;
  }
}

// src/SPH_density_reinitialization.h:8-121: per fluid particle, every other
// particle of the 9 cells (stencil order, ascending index in a cell; cells
// outside the grid skipped) adds m W to the density sum and m W / rho_j to the
// volume sum; rhoTemp = their ratio where the volume sum is non-zero.
__global__ void k_reinit_sum(Soa d, const int *__restrict__ cell_start,
                             const int *__restrict__ cellof, const int *__restrict__ scal,
                             Consts k) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= scal[S_NP] || d.type[i] != TSPH_FLUID) return;
  const double2 ar = d.r[i];
  const int c = cellof[i];
  const double h = k.h, m = k.m;
  double rho_sum = 0, wv_sum = 0;
  for (int s = 0; s < 9; ++s) {
    const int cl = c + stencil_off(s, k.ncy);
    if (cl < 0 || cl >= k.ncells) continue;
    const int j1 = cell_start[cl + 1];
    for (int j = cell_start[cl]; j < j1; ++j) {
      if (j == i) continue;
      const double2 nr = d.r[j];
      const double drx = ar.x - nr.x, dry = ar.y - nr.y;
      const double drs = sqrt((drx * drx) + (dry * dry));
      const double W = wendland64(drs, h, k.awen, k.bwen).w;
      rho_sum += m * W;
      wv_sum += m * W / d.rho[j];
    }
  }
  if (wv_sum != 0) d.rho_tmp[i] = rho_sum / wv_sum;
}

// src/SPH_density_reinitialization.h:123-139: fluid rho = rhoTemp; the
// reference's warning for a prior density outside [990, 1010] is counted
__global__ void k_reinit_overwrite(Soa d, const int *__restrict__ scal,
                                   unsigned long long *__restrict__ counters) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  bool bad = false;
  if (p < scal[S_NP] && d.type[p] == TSPH_FLUID) {
    const double rho = d.rho[p];
    bad = rho < 990 || rho > 1010;
    d.rho[p] = d.rho_tmp[p];
  }
  const unsigned long long w = __ballot(bad);
  if (w && (threadIdx.x & 63) == 0) atomicAdd(&counters[C_REINIT], (unsigned long long)__popcll(w));
}

}  // namespace

int launch_history_init(tsph_ctx *c) {
  if (int e = ensure_olds(c)) return e;
  hipLaunchKernelGGL(k_history_init, dim3(grid_for(c->n_ub, 256)), dim3(256), 0, c->stream, c->cur,
                     c->scal);
  TSPH_HIP(hipGetLastError());
  return 0;
}

int launch_verlet(tsph_ctx *c, double dt, bool euler) {
  if (int e = ensure_olds(c)) return e;
  if (euler)
    hipLaunchKernelGGL(k_euler, dim3(grid_for(c->n_ub, 256)), dim3(256), 0, c->stream, c->cur,
                       c->scal, dt);
  else
    hipLaunchKernelGGL(k_verlet, dim3(grid_for(c->n_ub, 256)), dim3(256), 0, c->stream, c->cur,
                       c->scal, dt);
  TSPH_HIP(hipGetLastError());
  c->records_valid = false;
  c->marks_valid = false;
  return 0;
}

int launch_leapfrog_one(tsph_ctx *c, double dt) {
  if (int e = ensure_olds(c)) return e;
  hipLaunchKernelGGL(k_lf1_density, dim3(grid_for(c->n_ub, 256)), dim3(256), 0, c->stream, c->cur,
                     c->scal, dt);
  TSPH_HIP(hipGetLastError());
  c->records_valid = false;
  c->marks_valid = false;
  return 0;
}

int launch_leapfrog_two(tsph_ctx *c, double dt, int step) {
  if (int e = ensure_olds(c)) return e;
  hipLaunchKernelGGL(k_lf2_density, dim3(grid_for(c->n_ub, 256)), dim3(256), 0, c->stream, c->cur,
                     c->scal, dt, (int)(step == 1));
  TSPH_HIP(hipGetLastError());
  c->records_valid = false;
  c->marks_valid = false;
  return 0;
}

int launch_reinit_sum(tsph_ctx *c) {
  hipLaunchKernelGGL(k_reinit_sum, dim3(grid_for(c->n_ub, 64)), dim3(64), 0, c->stream, c->cur,
                     c->cell_start, c->cellof, c->scal, c->k);
  TSPH_HIP(hipGetLastError());
  return 0;
}

int launch_reinit_overwrite(tsph_ctx *c) {
  hipLaunchKernelGGL(k_reinit_overwrite, dim3(grid_for(c->n_ub, 256)), dim3(256), 0, c->stream,
                     c->cur, c->scal, c->counters);
  TSPH_HIP(hipGetLastError());
  c->records_valid = false;
  return 0;
}

}  // namespace tsph
