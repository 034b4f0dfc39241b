// Per-particle streaming stages and the wall / probe gathers:
//   Density_to_pressure            src/SPH_pressure.h:65-89
//   Integrate_SymplecticPredictor  src/SPH_integration.h:266-287
//   Integrate_SymplecticCorrector  src/SPH_integration.h:289-309
//   mDBC_compute_density_mdbcGeo   src/SPH_mDBC.h:9-50 with
//                                  Kernel_density_approximation_MATRIX_mDBC,
//                                  src/SPH_kernel_approx.h:1141-1342
//   BT_compute_density_BTGeo       src/SPH_BT.h:9-50 with
//                                  Kernel_density_approximation_BT_update,
//                                  src/SPH_kernel_approx.h:1838-1953
//   RemoveParticlesOutOfDomain     src/SPH_btdebug.h:4-31
//   Water_Elevation(_inverse)      src/SPH_water_level.h:6-150
//   Kernel_velocity_approximation_TEST  src/testing.h:6-148
// The streaming stages are HBM-bound (a few flops per 16-100 bytes).
#include <cmath>
#include <vector>

#include "tsph_internal.h"
#include "tsph_physics.h"

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

// cell of a sampling point, computed like Particle_To_Cell
__device__ __forceinline__ int point_cell(const Consts &k, double x, double y) {
  return cell_coord(y, k.y0, k.of, k.kh) + cell_coord(x, k.x0, k.of, k.kh) * k.ncy;
}

// Wall densities: kWG lanes per wall particle.  The 9 stencil cells (cells
// outside the grid skipped) form one candidate sequence in the reference order
// (stencil order, ascending index within a cell); lane u of the group takes
// candidates u, u + kWG, u + 2 kWG, ...  The records of kWPre such rounds are
// loaded together (all in flight at once); then, round by round, the group's
// lanes evaluate their candidates' terms at once, put them in LDS, and lane u
// adds sums q = u, u + kWG, ... over the round's candidates in candidate
// order, so every sum is formed in the reference order.  A candidate that is
// not a fluid neighbour within the support adds nothing, as in the reference.
constexpr int kWG = 16;
constexpr int kWPre = 3;
constexpr int kWallsPerWave = 64 / kWG;

template <int NS, typename Term>
__device__ __forceinline__ void wall_sums(const Soa &d, const int *__restrict__ cell_start,
                                          const Consts &k, double2 ar, bool live,
                                          double (&acc)[NS], double *__restrict__ terms,
                                          Term term) {
  constexpr int QPL = (NS + kWG - 1) / kWG;  // sums per lane
  const int u = threadIdx.x & (kWG - 1);
  const int g0 = threadIdx.x & 63 & ~(kWG - 1);  // the group's first lane
  double *grp = terms + g0 * NS;                  // kWG x NS doubles of the group
  const double r2cut = 4.0 * k.h * k.h * (1.0 + 1e-12);
  // candidate f lies in the last cell s with pre[s] <= f, at index f + dj[s]
  int pre[9], dj[9], tot = 0;
#pragma unroll
  for (int s = 0; s < 9; ++s) {
    pre[s] = 0;
    dj[s] = 0;
  }
  if (live) {
    const int c = point_cell(k, ar.x, ar.y);
#pragma unroll
    for (int s = 0; s < 9; ++s) {
      const int cl = c + stencil_off(s, k.ncy);
      int lo = 0, n = 0;
      if (cl >= 0 && cl < k.ncells) {
        lo = cell_start[cl];
        n = cell_start[cl + 1] - lo;
      }
      pre[s] = tot;
      dj[s] = lo - tot;
      tot += n;
    }
  }
  double mine[QPL];  // sums u + kWG x
#pragma unroll
  for (int x = 0; x < QPL; ++x) mine[x] = 0.;
  for (int base = 0; base < tot; base += kWG * kWPre) {  // uniform within the group
    int tt[kWPre];
    double2 rr[kWPre];
    double hh[kWPre];
#pragma unroll
    for (int v = 0; v < kWPre; ++v) {
      const int f = base + v * kWG + u;
      int j = f + dj[0];
#pragma unroll
      for (int s = 1; s < 9; ++s) j = f >= pre[s] ? f + dj[s] : j;
      tt[v] = -1;
      rr[v] = make_double2(0., 0.);
      hh[v] = 1.;
      if (f < tot) {
        tt[v] = d.type[j];
        rr[v] = d.r[j];
        hh[v] = d.rho[j];
      }
    }
#pragma unroll
    for (int v = 0; v < kWPre; ++v) {
      if (base + v * kWG >= tot) break;
      bool ok = false;
      if (tt[v] == TSPH_FLUID) {
        const double drx = ar.x - rr[v].x, dry = ar.y - rr[v].y;
        const double r2 = drx * drx + dry * dry;
        if (r2 <= r2cut) {
          const double drs = sqrt(r2);
          if (drs / k.h <= 2.0) {
            // the candidate's terms, straight into the group's LDS slots (the
            // reference adds each to a running sum that starts at +0, which
            // never turns -0, so the term x itself stands for 0 + x)
            term(hh[v], drx, dry, drs, grp + u * NS);
            ok = true;
          }
        }
      }
      const unsigned gm = (unsigned)(__ballot(ok) >> g0) & ((1u << kWG) - 1u);
      // LDS accesses of one wavefront complete in issue order; the fences keep
      // the compiler from moving them across the hand-over
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
      // the round's terms of sum q = u + kWG x, read for all kWG candidates at
      // once (the LDS reads in flight together), then added in candidate
      // order; a candidate that is not a neighbour adds +0, which leaves the
      // sum unchanged (a sum that starts at +0 and adds terms is never -0).
      // Lanes past the last sum read sum NS - 1 and discard it.
#pragma unroll
      for (int x = 0; x < QPL; ++x) {
        const int q = min(u + kWG * x, NS - 1);
#pragma unroll
        for (int h = 0; h < kWG; h += kWG / 2) {
          double tv[kWG / 2];
#pragma unroll
          for (int w = 0; w < kWG / 2; ++w) tv[w] = grp[(h + w) * NS + q];
#pragma unroll
          for (int w = 0; w < kWG / 2; ++w) mine[x] += ((gm >> (h + w)) & 1u) ? tv[w] : 0.;
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
    }
  }
  // every lane of the group gets all NS sums
#pragma unroll
  for (int x = 0; x < QPL; ++x)
    if (u + kWG * x < NS) grp[u + kWG * x] = mine[x];
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
#pragma unroll
  for (int q = 0; q < NS; ++q) acc[q] = grp[q];
}

// The out-of-range density count of the EOS pass that follows the wall
// densities in the symplectic step (k_eos's count; the pass itself is not
// needed there): the wall densities just computed are counted by the lanes
// that computed them (count_walls), every other particle by a sweep that
// runs in workgroups of its own at the end of the same launch, beside the
// wall workgroups rather than after them.  A sweep thread loads kSweep
// particles at a time, type and density together.
constexpr int kSweep = 4;

__device__ __forceinline__ void count_walls(bool wall_bad, unsigned long long *__restrict__ counters) {
  const unsigned long long w = __ballot(wall_bad);
  if (w && (threadIdx.x & 63) == 0) atomicAdd(&counters[C_BADRHO], (unsigned long long)__popcll(w));
}

// sweep workgroup sb of nsb (64 threads each)
__device__ __forceinline__ void sweep_density(const Soa &d, const int *__restrict__ scal, int sb,
                                              int nsb, unsigned long long *__restrict__ counters) {
  const int n = scal[S_NP];
  const int stride = nsb * 64;
  unsigned bad = 0;
  for (int p0 = sb * 64 + (int)threadIdx.x; p0 < n; p0 += kSweep * stride) {
    int t[kSweep];
    double rho[kSweep];
#pragma unroll
    for (int q = 0; q < kSweep; ++q) {
      const int p = p0 + q * stride;
      t[q] = p < n ? d.type[p] : TSPH_WALL;
      rho[q] = p < n ? d.rho[p] : 0.;
    }
#pragma unroll
    for (int q = 0; q < kSweep; ++q)
      if (t[q] != TSPH_WALL && t[q] != TSPH_INLET && t[q] != TSPH_OUTLET)
        bad += rho[q] > 1500. || rho[q] < 0.;
  }
  for (int o = 32; o > 0; o >>= 1) bad += __shfl_xor(bad, o, 64);
  if (bad && threadIdx.x == 0) atomicAdd(&counters[C_BADRHO], (unsigned long long)bad);
}

// sweep workgroups that follow wall_blocks wall workgroups for n particles
inline int sweep_blocks(int n) { return grid_for(n, 64 * kSweep); }

// nwall: the wall count of the cell list in use (one of the two alternating slots)
// blocks past wall_blocks sweep the other particles' densities (counters set)
__global__ __launch_bounds__(64) void k_mdbc(Soa d, const int *__restrict__ cell_start, const int *__restrict__ nwall,
                       const int *__restrict__ wall_list, Consts k0,
                       const Consts *__restrict__ cases, unsigned long long *__restrict__ counters,
                       const int *__restrict__ scal, int wall_blocks) {
  if ((int)blockIdx.x >= wall_blocks) {
    sweep_density(d, scal, blockIdx.x - wall_blocks, gridDim.x - wall_blocks, counters);
    return;
  }
  __shared__ __align__(16) double s_terms[64 * 12];
  const int w = blockIdx.x * kWallsPerWave + (int)(threadIdx.x / kWG);
  const bool live = w < *nwall;
  const int i = live ? wall_list[w] : 0;
  const Consts k = live ? case_consts(k0, cases, d.cid, i) : k0;
  const double2 gn = live ? d.gnr[i] : make_double2(0., 0.);
  const double2 r = live ? d.r[i] : make_double2(0., 0.);  // loaded with the ghost node
  const double m = k.m;
  // sums of src/SPH_kernel_approx.h:1141-1342 in its order:
  // rhowsum, drhodx, drhody, a11, a12, a13, a21, a22, a23, a31, a32, a33
  double acc[12];
  wall_sums<12>(d, cell_start + k.cell_base, k, gn, live, acc, s_terms,
                [&](double nrho, double drx, double dry, double drs, double *a) {
                  const KW64 kw = wendland64(drs, k.h, k.awen, k.bwen);
                  const double mr = m / nrho;
                  a[0] = m * kw.w;
                  a[1] = drx * kw.f * m;
                  a[2] = dry * kw.f * m;
                  a[3] = kw.w * mr;
                  a[4] = drx * kw.w * mr;
                  a[5] = dry * kw.w * mr;
                  a[6] = drx * kw.f * mr;
                  a[7] = drx * drx * kw.f * mr;
                  a[8] = drx * dry * kw.f * mr;
                  a[9] = dry * kw.f * mr;
                  a[10] = dry * drx * kw.f * mr;
                  a[11] = dry * dry * kw.f * mr;
                });
  bool wall_bad = false;
  if (live && (threadIdx.x & (kWG - 1)) == 0) {
  const double rhowsum = acc[0], drhodx = acc[1], drhody = acc[2];
  const double a11 = acc[3], a12 = acc[4], a13 = acc[5], a21 = acc[6], a22 = acc[7], a23 = acc[8];
  const double a31 = acc[9], a32 = acc[10], a33 = acc[11];
  const double2 gnd = make_double2(r.x - gn.x, r.y - gn.y);
  double rho = 1000.;
  // determinant, src/SPH_kernel_approx.h:663-675
  const double det = (a11 * a22 * a33 + a12 * a23 * a31 + a13 * a21 * a32 - a31 * a22 * a13 -
                      a32 * a23 * a11 - a33 * a21 * a12);
  if (det > 0.001) {
    const double i11 = (a22 * a33 - a23 * a32) / det;
    const double i12 = -(a12 * a33 - a13 * a32) / det;
    const double i13 = (a12 * a23 - a13 * a22) / det;
    const double i21 = -(a21 * a33 - a23 * a31) / det;
    const double i22 = (a11 * a33 - a13 * a31) / det;
    const double i23 = -(a11 * a23 - a13 * a21) / det;
    const double i31 = (a21 * a32 - a22 * a31) / det;
    const double i32 = -(a11 * a32 - a12 * a31) / det;
    const double i33 = (a11 * a22 - a12 * a21) / det;
    const double dtemp = rhowsum * i11 + drhodx * i12 + drhody * i13;
    const double grx = -(rhowsum * i21 + drhodx * i22 + drhody * i23);
    const double gry = -(rhowsum * i31 + drhodx * i32 + drhody * i33);
    rho = dtemp + grx * gnd.x + gry * gnd.y;
  } else if (a11 > 0) {
    rho = rhowsum / a11;
  }
  if (rho < 1000.) rho = 1000.;  // src/SPH_mDBC.h:44
  d.rho[i] = rho;
  wall_bad = rho > 1500. || rho < 0.;
  }
  if (counters) count_walls(wall_bad, counters);
}

__global__ __launch_bounds__(64) void k_bt_density(Soa d, const int *__restrict__ cell_start, const int *__restrict__ nwall,
                             const int *__restrict__ wall_list, Consts k0,
                             const Consts *__restrict__ cases, unsigned long long *__restrict__ counters,
                             const int *__restrict__ scal, int wall_blocks) {
  if ((int)blockIdx.x >= wall_blocks) {
    sweep_density(d, scal, blockIdx.x - wall_blocks, gridDim.x - wall_blocks, counters);
    return;
  }
  __shared__ __align__(16) double s_terms[64 * 2];
  const int w = blockIdx.x * kWallsPerWave + (int)(threadIdx.x / kWG);
  const bool live = w < *nwall;
  const int i = live ? wall_list[w] : 0;
  const Consts k = live ? case_consts(k0, cases, d.cid, i) : k0;
  const double m = k.m;
  double acc[2];  // wsum, rhowsum (src/SPH_kernel_approx.h:1838-1953)
  wall_sums<2>(d, cell_start + k.cell_base, k, live ? d.r[i] : make_double2(0., 0.), live, acc,
               s_terms, [&](double nrho, double, double, double drs, double *a) {
                 const KW64 kw = wendland64(drs, k.h, k.awen, k.bwen);
                 a[0] = kw.w * m / nrho;
                 a[1] = m * kw.w;
               });
  bool wall_bad = false;
  if (live && (threadIdx.x & (kWG - 1)) == 0) {
    double rho = acc[0] > 0. ? acc[1] / acc[0] : 1000.;
    if (rho < 1000.) rho = 1000.;
    d.rho[i] = rho;
    wall_bad = rho > 1500. || rho < 0.;
  }
  if (counters) count_walls(wall_bad, counters);
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
