// Wall densities (mDBC, BT) by workgroups of 64 threads, shared by their own
// launches (tsph_stages.hip) and the first cell build of a symplectic step,
// whose hash pass runs them in workgroups of its own (tsph_cells.hip):
//   mDBC_compute_density_mdbcGeo   src/SPH_mDBC.h:9-50 with
//                                  Kernel_density_approximation_MATRIX_mDBC,
//                                  src/SPH_kernel_approx.h:1141-1342
//   BT_compute_density_BTGeo       src/SPH_BT.h:9-50 with
//                                  Kernel_density_approximation_BT_update,
//                                  src/SPH_kernel_approx.h:1838-1953
#pragma once

#include "tsph_internal.h"
#include "tsph_physics.h"

namespace tsph {
namespace walls {

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

// The wall-density passes by workgroups of 64 threads: workgroup b < wall_blocks
// takes walls kWallsPerWave b ..., workgroup wall_blocks + s (s < nsweep) the
// density sweep's share s (counters set).  nwall: the wall count of the cell
// list in use (one of the two alternating slots).
__device__ __forceinline__ void mdbc_block(int b, int wall_blocks, int nsweep, const Soa &d,
                                           const int *__restrict__ cell_start,
                                           const int *__restrict__ nwall,
                                           const int *__restrict__ wall_list, const Consts &k0,
                                           const Consts *__restrict__ cases,
                                           unsigned long long *__restrict__ counters,
                                           const int *__restrict__ scal) {
  if (b >= wall_blocks) {
    sweep_density(d, scal, b - wall_blocks, nsweep, counters);
    return;
  }
  __shared__ __align__(16) double s_terms[64 * 12];
  const int w = b * kWallsPerWave + (int)(threadIdx.x / kWG);
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

__device__ __forceinline__ void bt_block(int b, int wall_blocks, int nsweep, const Soa &d,
                                         const int *__restrict__ cell_start,
                                         const int *__restrict__ nwall,
                                         const int *__restrict__ wall_list, const Consts &k0,
                                         const Consts *__restrict__ cases,
                                         unsigned long long *__restrict__ counters,
                                         const int *__restrict__ scal) {
  if (b >= wall_blocks) {
    sweep_density(d, scal, b - wall_blocks, nsweep, counters);
    return;
  }
  __shared__ __align__(16) double s_terms[64 * 2];
  const int w = b * kWallsPerWave + (int)(threadIdx.x / kWG);
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

// workgroups of a wall pass: wall workgroups, then sweep workgroups
inline int wall_blocks(const tsph_ctx *c) { return grid_for(c->n_wall, kWallsPerWave); }

}  // namespace walls
}  // namespace tsph
