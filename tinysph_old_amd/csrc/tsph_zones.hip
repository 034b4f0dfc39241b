// Buffer zones: the device form of BZ_Integrate_inlet / BZ_Integrate_outlet
// (src/SPH_buffer_zones.h:10-206) for the inflow driver
// (cases/inflowSlab_mDBC/SPH_simulationSymplectic.h:57-123).
//
// Inlet.  The reference walks the particles in order; a zone particle past the
// zone start becomes fluid and a new zone particle is appended (one buffer
// length behind it), so appended particles follow their parents' order, and the
// growing loop bound visits them in the same pass (three of the four branches
// advance every zone particle by v dt inside the loop, the new ones included).
//   k_bz_inlet_scan    advance the zone particles (those branches), list the
//                      crossing ones as {reference position, slot}
//   k_bz_inlet_append  one workgroup: sort the list by reference position,
//                      append the new particles at the end of the slot range
//                      with consecutive reference positions and indices, turn
//                      the parents into fluid
// New particles land unsorted after the live range; the next cell-list build
// sorts them in.  A new particle that would itself cross in the same pass
// (impossible for |v| dt below a buffer length) is counted, not re-examined.
//
// Outlet.  The reference erases inside its ascending loop, so the particle
// that slides into the erased slot is not examined: in a run of consecutive
// removal candidates every other one is erased, and a fluid particle right
// after an erased one is not converted.  With the removal states of every live
// particle written by reference position,
//   k_bz_outlet_mark   state 2 = removal candidate, 1 = other live particle
//   k_bz_outlet_apply  L = candidates immediately before (absent positions
//                      skipped); a candidate is erased when L is even, a fluid
//                      particle past the zone start converts when L is even
#include "tsph_internal.h"

namespace tsph {

namespace {

__device__ __forceinline__ bool inlet_cross(const Zone &z, double2 r) {
  if (z.ox != 0) return z.ox < 0 ? r.x > z.x0 : r.x < z.x0;
  return z.oy < 0 ? r.y > z.y0 : r.y < z.y0;
}

// src/SPH_buffer_zones.h:28-101
__global__ void k_bz_inlet_scan(Soa d, int *__restrict__ scal, Zone z, double dt,
                                int *__restrict__ list) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= scal[S_NP] || d.type[i] != z.btype) return;
  double2 r = d.r[i];
  if (z.ox > 0) {
    r.x = r.x + d.v[i].x * dt;
    d.r[i] = r;
  } else if (z.oy != 0) {
    r.y = r.y + d.v[i].y * dt;
    d.r[i] = r;
  }
  if (!inlet_cross(z, r)) return;
  const int k = atomicAdd(&scal[S_BZCNT], 1);
  if (k < kBzMax) {
    list[2 * k] = (int)d.seq[i];
    list[2 * k + 1] = i;
  }
}

__global__ __launch_bounds__(1024) void k_bz_inlet_append(Soa d, int *__restrict__ scal, Zone z,
                                                          double dt, const int *__restrict__ list,
                                                          int capacity, int seq_cap, double dp,
                                                          uint8_t *__restrict__ seqstate,
                                                          unsigned long long *__restrict__ counters) {
  __shared__ unsigned long long key[kBzMax];
  const int n = scal[S_BZCNT];
  const int np = scal[S_NP], ms = scal[S_MAXSEQ];
  const int tid = threadIdx.x;
  if (n == 0) return;
  if (n > kBzMax || np + n > capacity || ms + n > seq_cap) {
    // refused as a whole: nothing converted, nothing appended
    if (tid == 0) {
      atomicAdd(&counters[C_BZREJECT], (unsigned long long)n);
      scal[S_BZCNT] = 0;
    }
    return;
  }
  int P = 1;
  while (P < n) P <<= 1;
  for (int t = tid; t < P; t += blockDim.x)
    key[t] = t < n ? ((unsigned long long)(unsigned)list[2 * t] << 32) | (unsigned)list[2 * t + 1]
                   : ~0ull;
  __syncthreads();
  // bitonic sort by reference position (the slot rides in the low bits)
  for (int k = 2; k <= P; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int t = tid; t < P; t += blockDim.x) {
        const int u = t ^ j;
        if (u > t) {
          const unsigned long long a = key[t], b = key[u];
          const bool up = (t & k) == 0;
          if ((a > b) == up) {
            key[t] = b;
            key[u] = a;
          }
        }
      }
      __syncthreads();
    }
  }
  const double2 vin = make_double2(z.vx, z.vy);
  unsigned chained = 0;
  for (int q = tid; q < n; q += blockDim.x) {
    const int parent = (int)(unsigned)(key[q] & 0xffffffffull);
    const int ns = np + q;
    const double2 pr = d.r[parent];
    double2 r0;
    if (z.ox != 0) r0 = make_double2(z.ox < 0 ? pr.x - z.bl : pr.x + z.bl, pr.y);
    else r0 = make_double2(pr.x, z.oy < 0 ? pr.y - z.bl : pr.y + z.bl);
    // Add_particle (src/SPH_particle_system.h:25-50): p 0, rho 1000, a 0, old
    // values = the new ones
    d.type[ns] = (uint8_t)z.btype;
    d.index[ns] = np + q;
    d.seq[ns] = (uint32_t)(ms + q);
    d.p[ns] = 0.;
    d.rho[ns] = 1000.;
    d.drho[ns] = 0.;
    d.v[ns] = vin;
    d.a[ns] = make_double2(0., 0.);
    d.r_o[ns] = r0;
    d.v_o[ns] = vin;
    d.rho_o[ns] = 1000.;
    if (d.gnr) d.gnr[ns] = make_double2(0., 0.);
    if (d.rc) {
      d.rc[ns] = make_double2(0., 0.);
      d.rhoc[ns] = 0.;
    }
    if (d.n) d.n[ns] = make_double2(0., 0.);
    if (d.omega) {
      const double w = dp * dp;
      d.omega[ns] = w; d.omega_o[ns] = w;
      d.omegarho[ns] = 1000. * w; d.omegarho_o[ns] = 1000. * w;
      d.omegav[ns] = make_double2(vin.x * w, vin.y * w); d.omegav_o[ns] = d.omegav[ns];
      d.domega[ns] = 0.; d.domegarho[ns] = 0.; d.domegav[ns] = make_double2(0., 0.);
      d.gradrho[ns] = d.gradvx[ns] = d.gradvy[ns] = make_double2(0., 0.);
    }
    // the pass reaches the new particle later: the advancing branches move it
    double2 r = r0;
    if (z.ox > 0) r.x = r.x + vin.x * dt;
    else if (z.oy != 0) r.y = r.y + vin.y * dt;
    d.r[ns] = r;
    chained += inlet_cross(z, r);
    seqstate[ms + q] = 1;
    d.type[parent] = TSPH_FLUID;
  }
  if (chained) atomicAdd(&counters[C_BZCHAIN], (unsigned long long)chained);
  __syncthreads();
  if (tid == 0) {
    scal[S_NP] = np + n;
    scal[S_MAXSEQ] = ms + n;
    scal[S_BZCNT] = 0;
    atomicAdd(&counters[C_APPENDED], (unsigned long long)n);
  }
}

__device__ __forceinline__ bool outlet_past_start(const Zone &z, double2 r) {
  return z.ox > 0 ? r.x > z.x0 : r.x < z.x0;
}
__device__ __forceinline__ bool outlet_out(const Zone &z, double2 r, double dp) {
  return z.ox > 0 ? (r.x > z.x0 + z.bl || r.y < dp / 2) : (r.x < z.x0 - z.bl || r.y < dp / 2);
}

// src/SPH_buffer_zones.h:136-188 (x-oriented zones; y-oriented ones do nothing)
__global__ void k_bz_outlet_mark(Soa d, const int *__restrict__ scal, Zone z, double dp,
                                 uint8_t *__restrict__ state) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= scal[S_NP]) return;
  const int t = d.type[i];
  const uint32_t s = d.seq[i];
  if (t == kRemoved) {
    state[s] = 0;
    return;
  }
  state[s] = (t == z.btype && outlet_out(z, d.r[i], dp)) ? 2 : 1;
}

__global__ void k_bz_outlet_apply(Soa d, const int *__restrict__ scal, Zone z,
                                  const uint8_t *__restrict__ state,
                                  unsigned long long *__restrict__ counters) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= scal[S_NP]) return;
  const int t = d.type[i];
  const uint32_t s = d.seq[i];
  const bool cand = state[s] == 2;
  const bool conv = !cand && t == TSPH_FLUID && outlet_past_start(z, d.r[i]);
  if (!cand && !conv) return;
  int run = 0;
  for (int64_t j = (int64_t)s - 1; j >= 0; --j) {
    const uint8_t st = state[j];
    if (st == 0) continue;
    if (st != 2) break;
    ++run;
  }
  if (run & 1) return;  // the previous particle was erased: this one is not examined
  if (cand) {
    d.type[i] = kRemoved;
    atomicAdd(&counters[C_REMOVED], 1ull);
  } else {
    d.type[i] = (uint8_t)z.btype;
    d.a[i] = make_double2(0., 0.);
    d.drho[i] = 0.;
  }
}

}  // namespace

int launch_bz_inlet(tsph_ctx *c, int zone, double dt) {
  // appended particles take slots past the last build's permutation
  if (int e = ensure_olds(c)) return e;
  const Zone &z = c->zones[zone];
  // appended particles extend the slot range: launches cover the capacity
  c->n_ub = c->capacity;
  hipLaunchKernelGGL(k_bz_inlet_scan, dim3(grid_for(c->n_ub, 256)), dim3(256), 0, c->stream, c->cur,
                     c->scal, z, dt, c->bz_list);
  hipLaunchKernelGGL(k_bz_inlet_append, dim3(1), dim3(1024), 0, c->stream, c->cur, c->scal, z, dt,
                     c->bz_list, c->capacity, c->seq_cap, c->k.dp, c->seqstate, c->counters);
  TSPH_HIP(hipGetLastError());
  c->cells_valid = false;
  c->records_valid = false;
  c->marks_valid = false;
  return 0;
}

int launch_bz_outlet(tsph_ctx *c, int zone, double dt) {
  (void)dt;
  const Zone &z = c->zones[zone];
  if (z.ox == 0) return 0;  // src/SPH_buffer_zones.h:190-204: y-oriented outlets are empty
  hipLaunchKernelGGL(k_bz_outlet_mark, dim3(grid_for(c->n_ub, 256)), dim3(256), 0, c->stream,
                     c->cur, c->scal, z, c->k.dp, c->seqstate);
  hipLaunchKernelGGL(k_bz_outlet_apply, dim3(grid_for(c->n_ub, 256)), dim3(256), 0, c->stream,
                     c->cur, c->scal, z, c->seqstate, c->counters);
  TSPH_HIP(hipGetLastError());
  c->cells_valid = false;
  c->records_valid = false;
  c->marks_valid = false;
  return 0;
}

}  // namespace tsph
