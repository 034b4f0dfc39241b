// Per-particle streaming stages and the wall / probe gathers:
//   Density_to_pressure            src/SPH_pressure.h:65-89
//   Integrate_SymplecticPredictor  src/SPH_integration.h:266-287
//   Integrate_SymplecticCorrector  src/SPH_integration.h:289-309
//   mDBC_compute_density_mdbcGeo, BT_compute_density_BTGeo (tsph_walls.h)
//   RemoveParticlesOutOfDomain     src/SPH_btdebug.h:4-31
//   Water_Elevation(_inverse)      src/SPH_water_level.h:6-150
//   Kernel_velocity_approximation_TEST  src/testing.h:6-148
// The streaming stages are HBM-bound (a few flops per 16-100 bytes).
#include <cmath>
#include <vector>

#include "tsph_internal.h"
#include "tsph_physics.h"
#include "tsph_walls.h"

namespace tsph {

namespace {

// EOS; on the mixed path it also writes the FP32 pair records of every
// particle (mx.rv != nullptr), which the pair loop that follows reads
__global__ void k_eos(Soa d, Mixed mx, const int *__restrict__ cellof, const int *__restrict__ scal,
                      Consts k0, const Consts *__restrict__ cases,
                      unsigned long long *__restrict__ counters) {
  const int n = scal[S_NP];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned bad = 0;
  if (i < n) {
    const Consts k = case_consts(k0, cases, d.cid, i);
    const int t = d.type[i];
    const double rho = d.rho[i];
    double p;
    if (t != TSPH_OUTLET && t != TSPH_INLET) {
      if (rho > 1500. || rho < 0.) bad = 1;
      p = tait64(rho, k.rho0, k.cs);
      d.p[i] = p;
    } else {
      p = d.p[i];
    }
    if (mx.rv) write_mixed_record(d, mx, cellof, k, i, rho, p);
  }
  const unsigned long long w = __ballot(bad);
  if (w && (threadIdx.x & 63) == 0) atomicAdd(&counters[C_BADRHO], (unsigned long long)__popcll(w));
}

__global__ void k_predictor(Soa d, const int *__restrict__ scal, double dt0,
                            const Consts *__restrict__ cases) {
  const int n = scal[S_NP];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double dt = cases ? cases[d.cid[i]].dt : dt0;
  double rho = d.rho[i];
  double2 r = d.r[i], v = d.v[i];
  d.rho_o[i] = rho;
  d.r_o[i] = r;
  d.v_o[i] = v;
  predict(r, v, rho, d.a[i], d.drho[i], dt);
  d.rho[i] = rho;
  d.r[i] = r;
  d.v[i] = v;
}

// corrector; also records the removal state for the RemoveParticlesOutOfDomain
// that follows it in every driver
__global__ void k_corrector(Soa d, const int *__restrict__ scal, Consts k0,
                            const Consts *__restrict__ cases, double dt0,
                            uint8_t *__restrict__ state) {
  const int n = scal[S_NP];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Consts k = case_consts(k0, cases, d.cid, i);
  correct_particle(d, k, cases ? k.dt : dt0, i, d.a[i], d.drho[i], state);
}

using namespace walls;

// nwall: the wall count of the cell list in use (one of the two alternating
// slots); workgroups past wall_blocks sweep the other particles' densities
// (counters set)
__global__ __launch_bounds__(64) void k_mdbc(Soa d, const int *__restrict__ cell_start,
                                             const int *__restrict__ nwall,
                                             const int *__restrict__ wall_list, Consts k0,
                                             const Consts *__restrict__ cases,
                                             unsigned long long *__restrict__ counters,
                                             const int *__restrict__ scal, int wall_blocks) {
  mdbc_block(blockIdx.x, wall_blocks, gridDim.x - wall_blocks, d, cell_start, nwall, wall_list, k0,
             cases, counters, scal);
}

__global__ __launch_bounds__(64) void k_bt_density(Soa d, const int *__restrict__ cell_start,
                                                   const int *__restrict__ nwall,
                                                   const int *__restrict__ wall_list, Consts k0,
                                                   const Consts *__restrict__ cases,
                                                   unsigned long long *__restrict__ counters,
                                                   const int *__restrict__ scal, int wall_blocks) {
  bt_block(blockIdx.x, wall_blocks, gridDim.x - wall_blocks, d, cell_start, nwall, wall_list, k0,
           cases, counters, scal);
}

// Removal.  The reference erases inside an ascending loop without stepping
// back, so the particle that follows an erased one in reference order is never
// examined: in a run of consecutive candidates every other one (the 1st, 3rd,
// ...) is removed.  State per reference position (seq): 0 absent, 1 kept,
// 2 candidate.
// marks for a removal that does not directly follow a corrector
__global__ void k_rm_mark(Soa d, const int *__restrict__ scal, Consts k0,
                          const Consts *__restrict__ cases, uint8_t *__restrict__ state) {
  const int n = scal[S_NP];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  state[d.seq[i]] = removal_state(case_consts(k0, cases, d.cid, i), d.type[i], d.r[i]);
}

__global__ void k_rm_decide(Soa d, const int *__restrict__ scal, const uint8_t *__restrict__ state,
                            unsigned long long *__restrict__ counters) {
  const int n = scal[S_NP];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (removal_decided(state, d.seq[i])) {
    d.type[i] = kRemoved;
    atomicAdd(&counters[C_REMOVED], 1ull);
  }
}

// Shepard volume sum, src/SPH_water_level.h:6-82 (all particle types)
__device__ double sum_kernel_at(const Soa &d, const int *__restrict__ cell_start, const Consts &k,
                                double x, double y) {
  const double kh = k.kh, of = k.of;
  const unsigned npx = (unsigned)(int)((x - k.x0 + of) / kh);
  const unsigned npy = (unsigned)(int)((y - k.y0 + of) / kh);
  const int c = (int)(npy + npx * (unsigned)k.ncy);
  const double h = k.h, m = k.m;
  double msum = 0;
  for (int s = 0; s < 9; ++s) {
    const int cl = c + stencil_off(s, k.ncy);
    if (cl < 0 || cl >= k.ncells) continue;
    const int j1 = cell_start[cl + 1];
    for (int j = cell_start[cl]; j < j1; ++j) {
      const double2 nr = d.r[j];
      const double drx = x - nr.x, dry = y - nr.y;
      const double drs = sqrt(drx * drx + dry * dry);
      if (drs / h > 2.0) continue;
      const KW64 kw = wendland64(drs, h, k.awen, k.bwen);
      msum += m * kw.w / d.rho[j];
    }
  }
  return msum;
}

// One block marches the sampling heights in chunks of blockDim: thread 0 forms
// the cumulative heights of the chunk in the reference's sequential order
// (ar.y += i*dp_wl), all threads evaluate their sample, the first sample that
// meets the threshold ends the march.
__global__ __launch_bounds__(256) void k_water_level(Soa d, const int *__restrict__ cell_start,
                                                     Consts k, double x, double y0, int inverse,
                                                     double *__restrict__ out) {
  __shared__ double ys[256];
  __shared__ int found;
  const double dp_wl = k.dp * 0.00005;
  const int nmax = int(k.ym / dp_wl);
  double y = y0;
  if (threadIdx.x == 0) { found = 0x7fffffff; *out = NAN; }
  __syncthreads();
  for (int base = 0; base < nmax; base += blockDim.x) {
    if (threadIdx.x == 0) {
      for (int t = 0; t < (int)blockDim.x; ++t) {
        const int i = base + t;
        if (inverse) y -= i * dp_wl; else y += i * dp_wl;
        ys[t] = y;
      }
    }
    __syncthreads();
    const int i = base + threadIdx.x;
    if (i < nmax) {
      const double msum = sum_kernel_at(d, cell_start, k, x, ys[threadIdx.x]);
      const bool hit = inverse ? (msum >= 0.5) : (msum <= 0.5);
      if (hit) atomicMin(&found, i);
    }
    __syncthreads();
    if (found != 0x7fffffff) {
      if (threadIdx.x == 0) {
        const int t = found - base;
        *out = inverse ? ys[t] + found * dp_wl : ys[t] - found * dp_wl;
      }
      return;
    }
    __syncthreads();
  }
}

// Kernel_velocity_approximation_TEST, src/testing.h:6-148, at one point:
// out4 = {rho, p, vx, vy}
__device__ void interpolate_at(const Soa &d, const int *__restrict__ cell_start, const Consts &k,
                               double x, double y, double *out4) {
  const double kh = k.kh, of = k.of;
  const unsigned npx = (unsigned)(int)((x - k.x0 + of) / kh);
  const unsigned npy = (unsigned)(int)((y - k.y0 + of) / kh);
  const int c = (int)(npy + npx * (unsigned)k.ncy);
  double vwx = 0, vwy = 0, wsum = 0, rhow = 0, pw = 0, mw = 0;
  for (int s = 0; s < 9; ++s) {
    const int cl = c + stencil_off(s, k.ncy);
    if (cl < 0 || cl >= k.ncells) continue;
    const int j1 = cell_start[cl + 1];
    for (int j = cell_start[cl]; j < j1; ++j) {
      const double2 nr = d.r[j];
      const double drx = x - nr.x, dry = y - nr.y;
      const double drs = sqrt(drx * drx + dry * dry);
      if (drs / k.h > 2.0) continue;
      const KW64 kw = wendland64(drs, k.h, k.awen, k.bwen);
      const double2 nv = d.v[j];
      wsum += kw.w;
      vwx = vwx + nv.x * kw.w;
      vwy = vwy + nv.y * kw.w;
      rhow = rhow + d.rho[j] * kw.w;
      pw = pw + d.p[j] * kw.w;
      mw += k.m * kw.w / d.rho[j];
    }
  }
  double rho = 0, p = 0, vx = 0, vy = 0;
  if (mw > 0.5) {
    vx = 0. + vwx / wsum;
    vy = 0. + vwy / wsum;
    rho = 0 + rhow / wsum;
    p = 0 + pw / wsum;
  }
  if (wsum == 0) { vx = 0.; vy = 0.; }
  out4[0] = rho; out4[1] = p; out4[2] = vx; out4[3] = vy;
}

__global__ void k_interp_point(Soa d, const int *__restrict__ cell_start, Consts k, double x,
                               double y, double *__restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  interpolate_at(d, cell_start, k, x, y, out);
}

// GenerateInterpol's grid, src/SPH_dataInterpolation.h:13-91: point (gx, gy) at
// (x0 + gx dh, y0 + gy dh), stored at POS(gx, gy) = gy + gx ncy
// (src/SPH_defs.h:36), one lane per point
__global__ void k_interp_grid(Soa d, const int *__restrict__ cell_start, Consts k, double x0,
                              double y0, double dh, int ncx, int ncy, double *__restrict__ out,
                              bool soa) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = ncx * ncy;
  if (t >= n) return;
  const int gx = t / ncy, gy = t - gx * ncy;
  double v[4];
  interpolate_at(d, cell_start, k, x0 + gx * dh, y0 + gy * dh, v);
  if (soa) {  // rho[n], p[n], v[2n]
    out[t] = v[0];
    out[n + t] = v[1];
    out[2 * n + 2 * t] = v[2];
    out[2 * n + 2 * t + 1] = v[3];
  } else {
    out[4 * t] = v[0];
    out[4 * t + 1] = v[1];
    out[4 * t + 2] = v[2];
    out[4 * t + 3] = v[3];
  }
}

}  // namespace

int launch_eos(tsph_ctx *c) {
  Mixed mx = c->prm.precision == TSPH_PREC_MIXED ? c->mx : Mixed{};
  hipLaunchKernelGGL(k_eos, dim3(grid_for(c->n_ub, 256)), dim3(256), 0, c->stream, c->cur, mx,
                     c->cellof, c->scal, c->k, c->cases, c->counters);
  TSPH_HIP(hipGetLastError());
  // the records are only meaningful for the cell list they were built on
  c->records_valid = c->prm.precision == TSPH_PREC_MIXED && c->cells_valid;
  return 0;
}

int launch_predictor(tsph_ctx *c, double dt) {
  c->olds_stale = false;  // it rewrites r_o, v_o, rho_o of every particle
  hipLaunchKernelGGL(k_predictor, dim3(grid_for(c->n_ub, 256)), dim3(256), 0, c->stream, c->cur,
                     c->scal, dt, c->cases);
  TSPH_HIP(hipGetLastError());
  c->records_valid = false;
  c->marks_valid = false;
  return 0;
}

int launch_corrector(tsph_ctx *c, double dt) {
  if (int e = ensure_olds(c)) return e;
  hipLaunchKernelGGL(k_corrector, dim3(grid_for(c->n_ub, 256)), dim3(256), 0, c->stream, c->cur,
                     c->scal, c->k, c->cases, dt, c->seqstate);
  TSPH_HIP(hipGetLastError());
  c->records_valid = false;
  c->marks_valid = true;
  return 0;
}

int launch_mdbc(tsph_ctx *c, bool count_density) {
  const int wb = grid_for(c->n_wall, kWallsPerWave), sb = count_density ? sweep_blocks(c->n_ub) : 0;
  hipLaunchKernelGGL(k_mdbc, dim3(wb + sb), dim3(64), 0, c->stream, c->cur, c->cell_start,
                     c->scal + c->nwall_idx, c->wall_list, c->k, c->cases,
                     count_density ? c->counters : nullptr, c->scal, wb);
  TSPH_HIP(hipGetLastError());
  c->records_valid = false;
  return 0;
}

int launch_bt_density(tsph_ctx *c, bool count_density) {
  const int wb = grid_for(c->n_wall, kWallsPerWave), sb = count_density ? sweep_blocks(c->n_ub) : 0;
  hipLaunchKernelGGL(k_bt_density, dim3(wb + sb), dim3(64), 0, c->stream, c->cur,
                     c->cell_start, c->scal + c->nwall_idx, c->wall_list, c->k, c->cases,
                     count_density ? c->counters : nullptr, c->scal, wb);
  TSPH_HIP(hipGetLastError());
  c->records_valid = false;
  return 0;
}

int launch_remove(tsph_ctx *c) {
  if (!c->marks_valid) {
    hipLaunchKernelGGL(k_rm_mark, dim3(grid_for(c->n_ub, 256)), dim3(256), 0, c->stream, c->cur,
                       c->scal, c->k, c->cases, c->seqstate);
  }
  hipLaunchKernelGGL(k_rm_decide, dim3(grid_for(c->n_ub, 256)), dim3(256), 0, c->stream, c->cur,
                     c->scal, c->seqstate, c->counters);
  TSPH_HIP(hipGetLastError());
  c->marks_valid = false;  // removed particles are re-marked as absent next time
  return 0;
}

int run_water_elevation(tsph_ctx *c, double x, double y0, int inverse, double *out) {
  double *dout = c->probe_out;
  hipLaunchKernelGGL(k_water_level, dim3(1), dim3(256), 0, c->stream, c->cur, c->cell_start, c->k,
                     x, y0, inverse, dout);
  TSPH_HIP(hipGetLastError());
  TSPH_HIP(hipMemcpyAsync(out, dout, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  TSPH_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

// out: 4 doubles per grid point, device memory
int launch_interp_grid(tsph_ctx *c, double x0, double y0, double dh, int ncx, int ncy, double *out,
                       bool soa) {
  hipLaunchKernelGGL(k_interp_grid, dim3(grid_for(ncx * ncy, 64)), dim3(64), 0, c->stream, c->cur,
                     c->cell_start, c->k, x0, y0, dh, ncx, ncy, out, soa);
  TSPH_HIP(hipGetLastError());
  return 0;
}

int run_interpolate_point(tsph_ctx *c, double x, double y, double *out4) {
  double *dout = c->probe_out;
  hipLaunchKernelGGL(k_interp_point, dim3(1), dim3(64), 0, c->stream, c->cur, c->cell_start, c->k,
                     x, y, dout);
  TSPH_HIP(hipGetLastError());
  TSPH_HIP(hipMemcpyAsync(out4, dout, 4 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  TSPH_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

}  // namespace tsph
