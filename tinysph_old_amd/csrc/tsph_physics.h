// Device closures of the pair physics.  The FP64 forms keep the reference's
// operand order and rounding points so that, compiled with -ffp-contract=off,
// they produce the same bits as the reference for every IEEE operation; the
// two powers of the reference (pow(rho/rho0, 7) of the EOS and pow(dpH, 1/7)
// of the diffusion term) come from tsph_pow7.h and tsph_root7.h.  The FP32
// forms are the mixed-precision path.
#pragma once

#include <hip/hip_runtime.h>
#include "tsph_internal.h"
#include "tsph_pow7.h"
#include "tsph_root7.h"

namespace tsph {

constexpr double kEps = 0.01;  // src/SPH_defs.h:69

// Particle_To_Cell cell coordinates: truncation toward zero (src/SPH_find_pairs.h:64-65)
__device__ __forceinline__ int cell_coord(double x, double lo, double of, double kh) {
  return (int)((x - lo + of) / kh);
}

// stencil offset s of the reference neighbour order {c, c+1, c+1+ncy, c+ncy,
// c-1+ncy, c-1, c-1-ncy, c-ncy, c+1-ncy} (src/SPH_interaction.h:26-36)
__device__ __forceinline__ int stencil_off(int s, int ncy) {
  switch (s) {
    case 0: return 0;
    case 1: return 1;
    case 2: return 1 + ncy;
    case 3: return ncy;
    case 4: return -1 + ncy;
    case 5: return -1;
    case 6: return -1 - ncy;
    case 7: return -ncy;
    default: return 1 - ncy;
  }
}
// the same cell expressed as (dx, dy) of cell coordinates
__device__ __forceinline__ void stencil_dxy(int s, int &dx, int &dy) {
  // dx: {0, 0, 1, 1, 1, 0, -1, -1, -1}, dy: {0, 1, 1, 0, -1, -1, -1, 0, 1}
  dx = (s >= 2 && s <= 4) ? 1 : (s >= 6 ? -1 : 0);
  dy = (s == 1 || s == 2 || s == 8) ? 1 : ((s >= 4 && s <= 6) ? -1 : 0);
}

// Wendland C2 in 2-D, src/SPH_kernel.h:52-77.  awen/bwen are the float-rounded
// constants; qq, W and F follow the reference's casts exactly.
struct KW64 { double w, f; };
__device__ __forceinline__ KW64 wendland64(double r, double h, double awen, double bwen) {
  const double qq = r / h;
  const float wqq1 = (float)(1.0 - 0.5 * qq);
  const float wqq2 = wqq1 * wqq1;
  const float wqq = (float)(qq + qq + 1.0);
  KW64 k{0.0, 0.0};
  if (qq <= 2.0) {
    k.w = awen * (double)wqq * (double)wqq2 * (double)wqq2;
    k.f = bwen * qq * (double)wqq2 * (double)wqq1 / r;
  }
  if (r == 0.0) { k.w = awen; k.f = 0.0; }
  return k;
}

// Artificial_Viscosity, src/SPH_viscouse_forces.h:7-25 (rhobar = 0.5*(rho_i+rho_j))
__device__ __forceinline__ double art_visc64(double h, double drs, double drdv, double rhobar,
                                             double c0, double alpha) {
  double av = 0.0;
  const double visco = drdv > alpha ? drdv : alpha;
  if (drdv < 0) {
    const double mu = h * drdv / (drs * drs + kEps * h * h);
    av = -visco * c0 * mu / rhobar;
  }
  return av;
}

// Density_Diffusion_term_FOURTAKAS, src/SPH_viscouse_forces.h:45-61
__device__ __forceinline__ double diff_fourtakas64(double h, double drs, double dr_z, double drdW,
                                                   double arho, double nrho, double rho0,
                                                   double c0, double delta, double m,
                                                   double gravity) {
  const double cb = c0 * c0 * rho0 / 7.;
  const double dpH = 1. + dr_z * rho0 * gravity / cb;
  const double drhoH = rho0 * root7(dpH) - rho0;  // pow(dpH, 1/7), tsph_root7.h
  const double psi = 2. * c0 * h * delta * ((nrho - arho) - drhoH) / (drs * drs + kEps * h * h);
  return psi * drdW * m / nrho;
}

// Compute_Pressure (Tait, gamma = 7), src/SPH_pressure.h:7-19
// RemoveParticlesOutOfDomain's decision for the particle at reference
// position s (src/SPH_btdebug.h:4-31): the reference erases inside an
// ascending loop without stepping back, so in a run of consecutive candidates
// (state 2; absent positions, state 0, are particles removed earlier and are
// skipped) every other one, the 1st, 3rd, ..., is erased.
__device__ __forceinline__ bool removal_decided(const uint8_t *__restrict__ state, uint32_t s) {
  if (state[s] != 2) return false;
  int run = 0;  // candidates immediately before s in reference order
  for (int64_t j = (int64_t)s - 1; j >= 0; --j) {
    const uint8_t st = state[j];
    if (st == 0) continue;
    if (st != 2) break;
    ++run;
  }
  return (run & 1) == 0;
}

// x + inc with Kahan's compensation (Kahan 1965): c holds the low-order part
// lost by the previous sums, is subtracted from this increment, and receives
// what this sum loses.  Exact as written only without contraction into FMAs
// (the library is built with -ffp-contract=off).
__device__ __forceinline__ double kahan_add(double x, double inc, double &c) {
  const double y = inc - c;
  const double t = x + y;
  c = (t - x) - y;
  return t;
}

// removal state of a particle at reference position seq: 0 absent, 1 kept,
// 2 candidate (fluid outside the domain by more than eps, src/SPH_btdebug.h:17-22)
__device__ __forceinline__ uint8_t removal_state(const Consts &k, int type, double2 r) {
  if (type == kRemoved) return 0;
  const bool out = type == TSPH_FLUID && ((r.y < (k.y0 - kEps)) || (r.y > k.ym + kEps) ||
                                          (r.x < k.x0 - kEps) || (r.x > k.xm - kEps));
  return out ? 2 : 1;
}

// Integrate_SymplecticCorrector for particle i, src/SPH_integration.h:287-309
// (a, drho: the particle's forces), and its removal state for the
// RemoveParticlesOutOfDomain that follows the corrector in every driver
__device__ __forceinline__ void correct_particle(const Soa &d, const Consts &k, double dt, int i,
                                                 double2 a, double drho, uint8_t *__restrict__ state) {
  const double e = (-drho / d.rho[i]) * dt;
  const double2 vo = d.v_o[i], ro = d.r_o[i];
  const double2 v = make_double2(vo.x + a.x * dt, vo.y + a.y * dt);
  d.v[i] = v;
  double2 r;
  if (!d.rc) {
    d.rho[i] = d.rho_o[i] * ((2. - e) / (2. + e));
    r = make_double2(ro.x + (v.x + vo.x) * dt * 0.5, ro.y + (v.y + vo.y) * dt * 0.5);
  } else {
    // Kahan-compensated form of the same updates: each is written as old value
    // + increment (rho_o (2-e)/(2+e) = rho_o + rho_o (-2e/(2+e))), and the part
    // of the increment that the rounded sum loses is carried to the next step
    double2 c = d.rc[i];
    double rc = d.rhoc[i];
    const double rho_o = d.rho_o[i];
    d.rho[i] = kahan_add(rho_o, rho_o * ((-2. * e) / (2. + e)), rc);
    r.x = kahan_add(ro.x, (v.x + vo.x) * dt * 0.5, c.x);
    r.y = kahan_add(ro.y, (v.y + vo.y) * dt * 0.5, c.y);
    d.rc[i] = c;
    d.rhoc[i] = rc;
  }
  d.r[i] = r;
  state[d.seq[i]] = removal_state(k, d.type[i], r);
}

// Integrate_SymplecticPredictor's half step, src/SPH_integration.h:266-287
// (the expressions in the reference's order; no contraction)
__device__ __forceinline__ double2 predicted_r(double2 r, double2 v, double dt) {
  return make_double2(r.x + v.x * dt * 0.5, r.y + v.y * dt * 0.5);
}
__device__ __forceinline__ void predict(double2 &r, double2 &v, double &rho, double2 a, double drho,
                                        double dt) {
  rho = rho + drho * dt * 0.5;
  r = predicted_r(r, v, dt);
  v = make_double2(v.x + a.x * dt * 0.5, v.y + a.y * dt * 0.5);
}

__device__ __forceinline__ double tait64(double rho, double rho0, double c0) {
  const double b = c0 * c0 * rho0 / 7.;
  const double rr = rho / rho0;
  return b * (pow7(rr) - 1);  // pow(rr, 7.), tsph_pow7.h
}

// ---------------------------------------------------------------------------
// FP32 forms for the mixed path.  Same formulas; divisions become reciprocal
// multiplies and the hydrostatic (1+x)^(1/7)-1 of the Fourtakas term is
// evaluated as a short series in x (|x| < 1e-2 for any case whose kernel
// support is below ~0.1 m of water), which avoids the cancellation of
// computing pow(1+x, 1/7) - 1 in single precision.
// ---------------------------------------------------------------------------
// FP32 pair records of particle i on the mixed path: position relative to its
// cell origin, velocity, and density as its deviation from rho0 (FP32
// resolves rho - rho0 to ~1e-7 of its size, while FP32 rho itself, ulp 6e-5 at
// 1000, would lose the small density differences the diffusion term is built
// from)
// (has_n: the particle carries a wall normal n)
__device__ __forceinline__ void write_mixed_record(const Mixed &mx, const Consts &k, int c, int i,
                                                   double2 r, double2 v, double rho, double p,
                                                   bool has_n, double2 n) {
  const int cx = c / k.ncy, cy = c - cx * k.ncy;
  const double ox = k.x0 - k.of + cx * k.kh, oy = k.y0 - k.of + cy * k.kh;
  mx.rv[i] = make_float4((float)(r.x - ox), (float)(r.y - oy), (float)v.x, (float)v.y);
  mx.rp[i] = make_float2((float)(rho - k.rho0), (float)p);
  if (mx.nrm && has_n) mx.nrm[i] = make_float2((float)n.x, (float)n.y);
}

__device__ __forceinline__ void write_mixed_record(const Soa &d, const Mixed &mx,
                                                   const int *__restrict__ cellof, const Consts &k,
                                                   int i, double rho, double p) {
  const bool has_n = d.n != nullptr;
  write_mixed_record(mx, k, cellof[i] - k.cell_base, i, d.r[i], d.v[i], rho, p, has_n,
                     has_n ? d.n[i] : make_double2(0., 0.));
}

__device__ __forceinline__ float hydro_drho_f32(float x, float rho0) {
  // (1+x)^(1/7) - 1 = x/7 - 3x^2/49 + 13x^3/343 - 65x^4/2401 + ...  (binomial series)
  const float t = x * (1.f / 7.f);
  return rho0 * t * (1.f + x * (-3.f / 7.f + x * (13.f / 49.f + x * (-65.f / 343.f))));
}

}  // namespace tsph
