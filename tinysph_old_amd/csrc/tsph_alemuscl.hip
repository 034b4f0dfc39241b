// ALE-SPH with MUSCL reconstruction: the pipeline of the ALE dam-break driver
// cases/ALESPH_dambreakLong/SPH_simulation.h:78-125,
//   Integrate_LeapFrog_partOne_ALESPH  src/ALESPH_integration.h:9-31
//   Compute_Density                    src/ALESPH_density_and_omega_bt.h:9-281
//   Kernel_gradient_approx             src/ALESPH_kernel_gradient.h:10-190
//   Integrate_density_compute_pressure src/ALESPH_integration.h:86-135
//   Compute_Acceleration_BT            src/ALESPH_acceleration_bt.h:12-295
//   Integrate_LeapFrog_partTwo_ALESPH  src/ALESPH_integration.h:33-83
//
// The three pair passes are policies over the generic tile walks of
// tsph_tiles.h: FP64 in the reference's neighbour and operand order, and a
// mixed form with FP32 fluxes on tile-frame positions and densities carried as
// deviations from rho0 (the linear EOS makes the pressure of a deviation exact).
//
// MUSCL states (src/ALESPH_riemann_solver.h:195-235): A_L = A_i + a/2 gA_i.dr,
// A_R = A_j - a/2 gA_j.dr with dr = r_i - r_j; a = 1 as the reference calls
// them, or the minmod weight of src/ALESPH_riemann_solver.h:180-191 per variable
// when limited (r_CD from the reconstructing particle to the other one; phi = 0
// where the two values are equal).
//
// Reference quirk kept: Compute_Density reads `dvdW` for fluid neighbours but
// never assigns it (src/ALESPH_density_and_omega_bt.h:135,237); the reference's
// own build evaluates that volume-flux term as zero, so only wall neighbours
// change omega in that pass.
#include <algorithm>

#include "tsph_internal.h"
#include "tsph_physics.h"
#include "tsph_tiles.h"

namespace tsph {

namespace {

using namespace tiles;

constexpr int kOutletF = 17;  // outletf, src/SPH_defs.h:44-65

// ---------------------------------------------------------------------------
// shared pieces
// ---------------------------------------------------------------------------
// velRiemannLinearized, src/ALESPH_riemann_solver.h:86-96
__device__ __forceinline__ double vel_riemann64(double rhoL, double rhoR, double vL, double vR,
                                                double rhoAvg, double c) {
  return 0.5 * (vR + vL) - 0.5 * (rhoR - rhoL) * c / rhoAvg;
}
// densRiemannLinearizedWithLimiter, src/ALESPH_riemann_solver.h:20-34
__device__ __forceinline__ double dens_riemann_lim64(double rhoL, double rhoR, double vL, double vR,
                                                     double rhoAvg, double c) {
  const double d = vL - vR;
  const double mx = d > 0 ? d : 0;
  const double e = 3. * mx;
  const double beta = e < c ? e : c;
  return 0.5 * (rhoL + rhoR) + 0.5 * beta * (vL - vR) * rhoAvg / (c * c);
}
template <typename T, typename V>
__device__ __forceinline__ T minmod_phi(T aC, T aD, V g, V rCD) {
  if (aD - aC == T(0)) return T(0);
  const T r = T(2) * (g.x * rCD.x + g.y * rCD.y) / (aD - aC) - T(1);
  const T mn = T(1) < r ? T(1) : r;
  return T(0) > mn ? T(0) : mn;
}

// LDS slots of the FP64 passes: r, v, {rho, omega}, n [, grad rho, grad vx, grad vy], type
template <bool G>
struct Lds64 {
  double2 *r, *v, *ro, *n, *gr, *gx, *gy;
  uint8_t *t;
};
template <bool G>
__device__ __forceinline__ Lds64<G> carve64(unsigned char *b, int cap) {
  double2 *p = reinterpret_cast<double2 *>(b);
  Lds64<G> L;
  L.r = p; L.v = p + cap; L.ro = p + 2 * cap; L.n = p + 3 * cap;
  L.gr = G ? p + 4 * cap : nullptr;
  L.gx = G ? p + 5 * cap : nullptr;
  L.gy = G ? p + 6 * cap : nullptr;
  L.t = reinterpret_cast<uint8_t *>(p + (G ? 7 : 4) * cap);
  return L;
}
template <bool G>
constexpr int slot_bytes64() { return (G ? 7 : 4) * 16 + 1; }

template <bool G>
__device__ __forceinline__ void stage64(const Soa &d, const Lds64<G> &L, int l, int j) {
  L.r[l] = d.r[j];
  L.v[l] = d.v[j];
  L.ro[l] = make_double2(d.rho[j], d.omega[j]);
  L.n[l] = d.n ? d.n[j] : make_double2(0., 0.);
  L.t[l] = d.type[j];
  if (G) {
    L.gr[l] = d.gradrho[j];
    L.gx[l] = d.gradvx[j];
    L.gy[l] = d.gradvy[j];
  }
}

// one neighbour's data
struct Nb64 {
  double2 r, v, n, gr, gx, gy;
  double rho, omega;
  int t;
};
template <bool G>
__device__ __forceinline__ Nb64 nb_lds64(const Lds64<G> &L, int l) {
  Nb64 b;
  b.r = L.r[l]; b.v = L.v[l]; b.n = L.n[l];
  const double2 ro = L.ro[l];
  b.rho = ro.x; b.omega = ro.y; b.t = L.t[l];
  if (G) { b.gr = L.gr[l]; b.gx = L.gx[l]; b.gy = L.gy[l]; }
  return b;
}
template <bool G>
__device__ __forceinline__ Nb64 nb_glob64(const Soa &d, int j) {
  Nb64 b;
  b.r = d.r[j]; b.v = d.v[j]; b.n = d.n ? d.n[j] : make_double2(0., 0.);
  b.rho = d.rho[j]; b.omega = d.omega[j]; b.t = d.type[j];
  if (G) { b.gr = d.gradrho[j]; b.gx = d.gradvx[j]; b.gy = d.gradvy[j]; }
  return b;
}

struct Own64 {
  double2 r, v, gr, gx, gy;
  double rho, omega;
};
template <bool G>
__device__ __forceinline__ Own64 own64(const Soa &d, int i) {
  Own64 o;
  if (i >= 0) {
    o.r = d.r[i]; o.v = d.v[i]; o.rho = d.rho[i]; o.omega = d.omega[i];
    if (G) { o.gr = d.gradrho[i]; o.gx = d.gradvx[i]; o.gy = d.gradvy[i]; }
  } else {
    o.r = o.v = make_double2(0., 0.);
    o.rho = 1.; o.omega = 0.;
    o.gr = o.gx = o.gy = make_double2(0., 0.);
  }
  return o;
}

// reconstructed left/right states (FP64, reference operand order)
struct LR64 {
  double rhoL, rhoR;
  double2 vL, vR;
};
template <bool G, bool LIM>
__device__ __forceinline__ LR64 states64(const Own64 &o, const Nb64 &b, double2 dr) {
  LR64 s;
  if (!G) {
    s.rhoL = o.rho; s.rhoR = b.rho; s.vL = o.v; s.vR = b.v;
    return s;
  }
  double a1 = 1, a2 = 1, a3 = 1, a4 = 1, a5 = 1, a6 = 1;
  if (LIM) {
    const double2 rij = make_double2(0. - dr.x, 0. - dr.y);
    a1 = minmod_phi(o.rho, b.rho, o.gr, rij);
    a2 = minmod_phi(b.rho, o.rho, b.gr, dr);
    a3 = minmod_phi(o.v.x, b.v.x, o.gx, rij);
    a4 = minmod_phi(o.v.y, b.v.y, o.gy, rij);
    a5 = minmod_phi(b.v.x, o.v.x, b.gx, dr);
    a6 = minmod_phi(b.v.y, o.v.y, b.gy, dr);
  }
  s.rhoL = o.rho + a1 * 0.5 * (o.gr.x * dr.x + o.gr.y * dr.y);
  s.rhoR = b.rho - a2 * 0.5 * (b.gr.x * dr.x + b.gr.y * dr.y);
  s.vL.x = o.v.x + a3 * 0.5 * (o.gx.x * dr.x + o.gx.y * dr.y);
  s.vL.y = o.v.y + a4 * 0.5 * (o.gy.x * dr.x + o.gy.y * dr.y);
  s.vR.x = b.v.x - a5 * 0.5 * (b.gx.x * dr.x + b.gx.y * dr.y);
  s.vR.y = b.v.y - a6 * 0.5 * (b.gy.x * dr.x + b.gy.y * dr.y);
  return s;
}

// ---------------------------------------------------------------------------
// FP64 passes
// ---------------------------------------------------------------------------
// Kernel_gradient_approx (fluid only): gradients of rho, vx, vy
struct Grad64 {
  Soa d;
  Consts k;
  using Lds = Lds64<false>;
  using Own = Own64;
  struct Acc {
    double grx = 0, gry = 0, gxx = 0, gxy = 0, gyx = 0, gyy = 0, gamma = 0;
  };
  static __device__ Lds lds(unsigned char *b, int cap) { return carve64<false>(b, cap); }
  __device__ void stage(const Lds &L, int l, int j) const { stage64<false>(d, L, l, j); }
  __device__ bool active(int i) const { return d.type[i] == TSPH_FLUID; }
  __device__ Own own(int i) const { return own64<false>(d, i); }
  __device__ double2 pos_lds(const Lds &L, int l) const { return L.r[l]; }
  __device__ double2 pos_glob(int j) const { return d.r[j]; }
  __device__ void pair(Acc &a, const Own &o, const Nb64 &b) const {
    const double2 dr = make_double2(o.r.x - b.r.x, o.r.y - b.r.y);
    const double drs = sqrt((dr.x * dr.x) + (dr.y * dr.y));
    const KW64 kw = wendland64(drs, k.h, k.awen, k.bwen);
    const double W = kw.w;
    const double dWx = dr.x * kw.f, dWy = dr.y * kw.f;
    const double drho = b.rho - o.rho, dvx = b.v.x - o.v.x, dvy = b.v.y - o.v.y;
    if (b.t == TSPH_WALL) {
      const double nbx = b.n.x * (-1), nby = b.n.y * (-1);
      a.grx += drho * nbx * W * k.dp;
      a.gry += drho * nby * W * k.dp;
      a.gxx += dvx * nbx * W * k.dp;
      a.gxy += dvx * nby * W * k.dp;
      a.gyx += dvy * nbx * W * k.dp;
      a.gyy += dvy * nby * W * k.dp;
    } else {
      a.grx += drho * b.omega * dWx;
      a.gry += drho * b.omega * dWy;
      a.gxx += dvx * b.omega * dWx;
      a.gxy += dvx * b.omega * dWy;
      a.gyx += dvy * b.omega * dWx;
      a.gyy += dvy * b.omega * dWy;
    }
    a.gamma += W * k.m / b.rho;
  }
  __device__ void pair_lds(Acc &a, const Own &o, const Lds &L, int l) const { pair(a, o, nb_lds64<false>(L, l)); }
  __device__ void pair_glob(Acc &a, const Own &o, int j) const { pair(a, o, nb_glob64<false>(d, j)); }
  __device__ void finish(const Acc &a, const Own &, int i) const {
    if (a.gamma == 0) {
      d.gradrho[i] = d.gradvx[i] = d.gradvy[i] = make_double2(0., 0.);
    } else {
      d.gradrho[i] = make_double2(a.grx * 2., a.gry * 2.);
      d.gradvx[i] = make_double2(a.gxx * 2., a.gxy * 2.);
      d.gradvy[i] = make_double2(a.gyx * 2., a.gyy * 2.);
    }
  }
};

// Compute_Density: d(omega), d(omega rho)
template <bool G, bool LIM>
struct Dens64 {
  Soa d;
  Consts k;
  using Lds = Lds64<G>;
  using Own = Own64;
  struct Acc {
    double dom = 0, domrho = 0, gamma = 0;
  };
  static __device__ Lds lds(unsigned char *b, int cap) { return carve64<G>(b, cap); }
  __device__ void stage(const Lds &L, int l, int j) const { stage64<G>(d, L, l, j); }
  __device__ bool active(int i) const {
    const int t = d.type[i];
    return t != kOutletF && t != TSPH_OUTLET && t != TSPH_INLET && t != TSPH_WALL && t != kRemoved;
  }
  __device__ Own own(int i) const { return own64<G>(d, i); }
  __device__ double2 pos_lds(const Lds &L, int l) const { return L.r[l]; }
  __device__ double2 pos_glob(int j) const { return d.r[j]; }
  __device__ void pair(Acc &a, const Own &o, const Nb64 &b) const {
    const double c0 = k.cs;
    const double2 dr = make_double2(o.r.x - b.r.x, o.r.y - b.r.y);
    const double dvx = o.v.x - b.v.x, dvy = o.v.y - b.v.y;
    const double drs = sqrt((dr.x * dr.x) + (dr.y * dr.y));
    const double inv = 1 / drs;
    const double drnx = dr.x * inv * (-1), drny = dr.y * inv * (-1);
    const KW64 kw = wendland64(drs, k.h, k.awen, k.bwen);
    const double W = kw.w;
    const double dWx = dr.x * kw.f, dWy = dr.y * kw.f;
    const LR64 s = states64<G, LIM>(o, b, dr);
    const double vl = s.vL.x * drnx + s.vL.y * drny;
    const double vr = s.vR.x * drnx + s.vR.y * drny;
    const double rhos = dens_riemann_lim64(s.rhoL, s.rhoR, vl, vr, 0.5 * (s.rhoL + s.rhoR), c0);
    const double vss = vel_riemann64(s.rhoL, s.rhoR, vl, vr, 0.5 * (s.rhoL + s.rhoR), c0);
    const double vsx = drnx * vss + ((o.v.x + b.v.x) * 0.5 - drnx * (vl + vr) * 0.5);
    const double vsy = drny * vss + ((o.v.y + b.v.y) * 0.5 - drny * (vl + vr) * 0.5);
    const double vsdW = (vsx - o.v.x) * dWx + (vsy - o.v.y) * dWy;
    if (b.t == TSPH_WALL) {
      const double nbx = b.n.x * (-1), nby = b.n.y * (-1);
      const double dvn = dvx * nbx + dvy * nby;
      a.dom -= dvn * W * k.dp;
    } else {
      // the fluid volume flux term is zero as the reference build evaluates it
      a.domrho -= 2 * rhos * b.omega * vsdW;
    }
    a.gamma += W * k.m / b.rho;
  }
  __device__ void pair_lds(Acc &a, const Own &o, const Lds &L, int l) const { pair(a, o, nb_lds64<G>(L, l)); }
  __device__ void pair_glob(Acc &a, const Own &o, int j) const { pair(a, o, nb_glob64<G>(d, j)); }
  __device__ void finish(const Acc &a, const Own &o, int i) const {
    if (a.gamma == 0) {
      d.domega[i] = 0.;
      d.domegarho[i] = 0.;
    } else {
      d.domega[i] = a.dom * o.omega;
      d.domegarho[i] = a.domrho * o.omega;
    }
  }
};

// Compute_Acceleration_BT: d(omega v)
template <bool G, bool LIM>
struct Acc64P {
  Soa d;
  Consts k;
  using Lds = Lds64<G>;
  using Own = Own64;
  struct Acc {
    double ax = 0, ay = 0, gamma = 0;
  };
  static __device__ Lds lds(unsigned char *b, int cap) { return carve64<G>(b, cap); }
  __device__ void stage(const Lds &L, int l, int j) const { stage64<G>(d, L, l, j); }
  __device__ bool active(int i) const { return d.type[i] == TSPH_FLUID; }
  __device__ Own own(int i) const { return own64<G>(d, i); }
  __device__ double2 pos_lds(const Lds &L, int l) const { return L.r[l]; }
  __device__ double2 pos_glob(int j) const { return d.r[j]; }
  __device__ void pair(Acc &a, const Own &o, const Nb64 &b) const {
    const double c0 = k.cs, rho0 = k.rho0;
    const double2 dr = make_double2(o.r.x - b.r.x, o.r.y - b.r.y);
    const double drs = sqrt((dr.x * dr.x) + (dr.y * dr.y));
    const double inv = 1 / drs;
    const double drnx = dr.x * inv * (-1), drny = dr.y * inv * (-1);
    const KW64 kw = wendland64(drs, k.h, k.awen, k.bwen);
    const double W = kw.w;
    const double dWx = dr.x * kw.f, dWy = dr.y * kw.f;
    const LR64 s = states64<G, LIM>(o, b, dr);
    const double vl = s.vL.x * drnx + s.vL.y * drny;
    const double vr = s.vR.x * drnx + s.vR.y * drny;
    // cs(rho) = c0 (rho/rho0)^(1/2) (src/ALESPH_riemann_solver.h:148-159), summed (:178)
    const double avgc = c0 * sqrt(o.rho / rho0) + c0 * sqrt(b.rho / rho0);
    const double rhos = dens_riemann_lim64(s.rhoL, s.rhoR, vl, vr, 0.5 * (s.rhoL + s.rhoR), avgc);
    const double vss = vel_riemann64(o.rho, b.rho, vl, vr, 0.5 * (s.rhoL + s.rhoR), avgc);
    const double vsx = drnx * vss + ((o.v.x + b.v.x) * 0.5 - drnx * (vl + vr) * 0.5);
    const double vsy = drny * vss + ((o.v.y + b.v.y) * 0.5 - drny * (vl + vr) * 0.5);
    if (b.t == TSPH_WALL) {
      const double nbx = b.n.x * (-1), nby = b.n.y * (-1);
      const double rw = s.rhoL - ((0 - s.vL.x) * nbx + (0 - s.vL.y) * nby) * s.rhoL / c0;
      const double pw = c0 * c0 * (rw - 1000.) * 1;
      a.ax -= 2 * W * k.dp * (pw * nbx);
      a.ay -= 2 * W * k.dp * (pw * nby);
    } else {
      const double ps = c0 * c0 * (rhos - rho0);
      a.ax -= 2 * b.omega * ((rhos * vsx * (vsx - o.v.x) + ps) * dWx + (rhos * vsx * (vsy - o.v.y)) * dWy);
      a.ay -= 2 * b.omega * ((rhos * vsy * (vsx - o.v.x)) * dWx + (rhos * vsy * (vsy - o.v.y) + ps) * dWy);
    }
    a.gamma += W * k.m / b.rho;
  }
  __device__ void pair_lds(Acc &a, const Own &o, const Lds &L, int l) const { pair(a, o, nb_lds64<G>(L, l)); }
  __device__ void pair_glob(Acc &a, const Own &o, int j) const { pair(a, o, nb_glob64<G>(d, j)); }
  __device__ void finish(const Acc &a, const Own &o, int i) const {
    const double gx = k.gx * o.omega * o.rho, gy = k.gy * o.omega * o.rho;
    if (a.gamma == 0)
      d.domegav[i] = make_double2(0. - gx, 0. - gy);
    else
      d.domegav[i] = make_double2(a.ax * o.omega - gx, a.ay * o.omega - gy);
  }
};

// ---------------------------------------------------------------------------
// mixed passes: FP32 fluxes, tile-frame positions, densities as rho - rho0
// ---------------------------------------------------------------------------
template <bool G>
struct Lds32 {
  float2 *p, *v, *dw, *n, *gr, *gx, *gy;  // dw = {rho - rho0, omega}
  uint8_t *t;
};
template <bool G>
__device__ __forceinline__ Lds32<G> carve32(unsigned char *b, int cap) {
  float2 *p = reinterpret_cast<float2 *>(b);
  Lds32<G> L;
  L.p = p; L.v = p + cap; L.dw = p + 2 * cap; L.n = p + 3 * cap;
  L.gr = G ? p + 4 * cap : nullptr;
  L.gx = G ? p + 5 * cap : nullptr;
  L.gy = G ? p + 6 * cap : nullptr;
  L.t = reinterpret_cast<uint8_t *>(p + (G ? 7 : 4) * cap);
  return L;
}
template <bool G>
constexpr int slot_bytes32() { return (G ? 7 : 4) * 8 + 1; }

__device__ __forceinline__ float2 f2(double2 a) { return make_float2((float)a.x, (float)a.y); }

struct Nb32 {
  float2 v, n, gr, gx, gy;
  float dev, omega;
  int t;
};
template <bool G>
__device__ __forceinline__ void stage32(const Soa &d, const Consts &k, const Lds32<G> &L, int l, int j,
                                        double ox, double oy) {
  const double2 r = d.r[j];
  L.p[l] = make_float2((float)(r.x - ox), (float)(r.y - oy));
  L.v[l] = f2(d.v[j]);
  L.dw[l] = make_float2((float)(d.rho[j] - k.rho0), (float)d.omega[j]);
  L.n[l] = d.n ? f2(d.n[j]) : make_float2(0.f, 0.f);
  L.t[l] = d.type[j];
  if (G) {
    L.gr[l] = f2(d.gradrho[j]);
    L.gx[l] = f2(d.gradvx[j]);
    L.gy[l] = f2(d.gradvy[j]);
  }
}
template <bool G>
__device__ __forceinline__ Nb32 nb_lds32(const Lds32<G> &L, int l) {
  Nb32 b;
  b.v = L.v[l]; b.n = L.n[l];
  const float2 dw = L.dw[l];
  b.dev = dw.x; b.omega = dw.y; b.t = L.t[l];
  if (G) { b.gr = L.gr[l]; b.gx = L.gx[l]; b.gy = L.gy[l]; }
  return b;
}
template <bool G>
__device__ __forceinline__ Nb32 nb_glob32(const Soa &d, const Consts &k, int j) {
  Nb32 b;
  b.v = f2(d.v[j]);
  b.n = d.n ? f2(d.n[j]) : make_float2(0.f, 0.f);
  b.dev = (float)(d.rho[j] - k.rho0); b.omega = (float)d.omega[j]; b.t = d.type[j];
  if (G) { b.gr = f2(d.gradrho[j]); b.gx = f2(d.gradvx[j]); b.gy = f2(d.gradvy[j]); }
  return b;
}

struct Own32 {
  float2 p, v, gr, gx, gy;
  float dev;
  double rho, omega;
};
template <bool G>
__device__ __forceinline__ Own32 own32(const Soa &d, const Consts &k, int i, double ox, double oy) {
  Own32 o;
  if (i >= 0) {
    const double2 r = d.r[i];
    o.p = make_float2((float)(r.x - ox), (float)(r.y - oy));
    o.v = f2(d.v[i]);
    o.rho = d.rho[i];
    o.omega = d.omega[i];
    o.dev = (float)(o.rho - k.rho0);
    if (G) { o.gr = f2(d.gradrho[i]); o.gx = f2(d.gradvx[i]); o.gy = f2(d.gradvy[i]); }
  } else {
    o.p = o.v = make_float2(0.f, 0.f);
    o.rho = 1.; o.omega = 0.; o.dev = 0.f;
    o.gr = o.gx = o.gy = make_float2(0.f, 0.f);
  }
  return o;
}

struct C32A {
  float ih, m, c0, ic0, rho0, dp, awen, bwen, pw0;
};
__host__ __device__ __forceinline__ C32A make_c32a(const Consts &k) {
  C32A c;
  c.ih = (float)(1. / k.h); c.m = (float)k.m; c.c0 = (float)k.cs; c.ic0 = (float)(1. / k.cs);
  c.rho0 = (float)k.rho0; c.dp = (float)k.dp; c.awen = (float)k.awen; c.bwen = (float)k.bwen;
  c.pw0 = (float)(k.rho0 - 1000.);  // the wall pressure is taken against 1000 (:242)
  return c;
}

struct KW32 {
  float w, f, inv;
};
__device__ __forceinline__ KW32 wendland32(const C32A &c, float r2) {
#pragma clang fp contract(fast)
  const float drs = __builtin_amdgcn_sqrtf(r2);
  const float qq = drs * c.ih;
  const float wq1 = 1.f - 0.5f * qq;
  const float wq2 = wq1 * wq1;
  KW32 k;
  k.w = c.awen * (qq + qq + 1.f) * wq2 * wq2;
  k.f = c.bwen * wq2 * wq1 * c.ih;
  k.inv = __builtin_amdgcn_rcpf(drs);
  return k;
}

// reconstructed states as deviations (rho - rho0) and velocities
struct LR32 {
  float aL, nR;
  float2 vL, vR;
};
template <bool G, bool LIM>
__device__ __forceinline__ LR32 states32(const Own32 &o, const Nb32 &b, float drx, float dry) {
#pragma clang fp contract(fast)
  LR32 s;
  if (!G) {
    s.aL = o.dev; s.nR = b.dev; s.vL = o.v; s.vR = b.v;
    return s;
  }
  float a1 = 1.f, a2 = 1.f, a3 = 1.f, a4 = 1.f, a5 = 1.f, a6 = 1.f;
  const float2 dr = make_float2(drx, dry), rij = make_float2(-drx, -dry);
  if (LIM) {
    a1 = minmod_phi(o.dev, b.dev, o.gr, rij);
    a2 = minmod_phi(b.dev, o.dev, b.gr, dr);
    a3 = minmod_phi(o.v.x, b.v.x, o.gx, rij);
    a4 = minmod_phi(o.v.y, b.v.y, o.gy, rij);
    a5 = minmod_phi(b.v.x, o.v.x, b.gx, dr);
    a6 = minmod_phi(b.v.y, o.v.y, b.gy, dr);
  }
  s.aL = o.dev + a1 * 0.5f * (o.gr.x * drx + o.gr.y * dry);
  s.nR = b.dev - a2 * 0.5f * (b.gr.x * drx + b.gr.y * dry);
  s.vL.x = o.v.x + a3 * 0.5f * (o.gx.x * drx + o.gx.y * dry);
  s.vL.y = o.v.y + a4 * 0.5f * (o.gy.x * drx + o.gy.y * dry);
  s.vR.x = b.v.x - a5 * 0.5f * (b.gx.x * drx + b.gx.y * dry);
  s.vR.y = b.v.y - a6 * 0.5f * (b.gy.x * drx + b.gy.y * dry);
  return s;
}

struct Grad32 {
  Soa d;
  Consts k;
  C32A c;
  using Lds = Lds32<false>;
  using Own = Own32;
  struct Acc {
    float grx = 0, gry = 0, gxx = 0, gxy = 0, gyx = 0, gyy = 0, gamma = 0;
  };
  static __device__ Lds lds(unsigned char *b, int cap) { return carve32<false>(b, cap); }
  __device__ void stage(const Lds &L, int l, int j, double ox, double oy) const { stage32<false>(d, k, L, l, j, ox, oy); }
  __device__ bool active(int i) const { return d.type[i] == TSPH_FLUID; }
  __device__ Own own(int i, double ox, double oy) const { return own32<false>(d, k, i, ox, oy); }
  __device__ float2 pos_lds(const Lds &L, int l) const { return L.p[l]; }
  __device__ float2 pos_glob(int j, double ox, double oy) const {
    const double2 r = d.r[j];
    return make_float2((float)(r.x - ox), (float)(r.y - oy));
  }
  __device__ void pair(Acc &a, const Own &o, const Nb32 &b, float drx, float dry, float r2) const {
#pragma clang fp contract(fast)
    const KW32 kw = wendland32(c, r2);
    const float dWx = drx * kw.f, dWy = dry * kw.f;
    const float drho = b.dev - o.dev, dvx = b.v.x - o.v.x, dvy = b.v.y - o.v.y;
    if (b.t == TSPH_WALL) {
      const float s = kw.w * c.dp;
      a.grx -= drho * b.n.x * s;
      a.gry -= drho * b.n.y * s;
      a.gxx -= dvx * b.n.x * s;
      a.gxy -= dvx * b.n.y * s;
      a.gyx -= dvy * b.n.x * s;
      a.gyy -= dvy * b.n.y * s;
    } else {
      a.grx += drho * b.omega * dWx;
      a.gry += drho * b.omega * dWy;
      a.gxx += dvx * b.omega * dWx;
      a.gxy += dvx * b.omega * dWy;
      a.gyx += dvy * b.omega * dWx;
      a.gyy += dvy * b.omega * dWy;
    }
    a.gamma += kw.w * c.m * __builtin_amdgcn_rcpf(c.rho0 + b.dev);
  }
  __device__ void pair_lds(Acc &a, const Own &o, const Lds &L, int l, float drx, float dry, float r2) const {
    pair(a, o, nb_lds32<false>(L, l), drx, dry, r2);
  }
  __device__ void pair_glob(Acc &a, const Own &o, int j, float drx, float dry, float r2) const {
    pair(a, o, nb_glob32<false>(d, k, j), drx, dry, r2);
  }
  __device__ void finish(const Acc &a, const Own &, int i) const {
    if (a.gamma == 0.f) {
      d.gradrho[i] = d.gradvx[i] = d.gradvy[i] = make_double2(0., 0.);
    } else {
      d.gradrho[i] = make_double2(2. * a.grx, 2. * a.gry);
      d.gradvx[i] = make_double2(2. * a.gxx, 2. * a.gxy);
      d.gradvy[i] = make_double2(2. * a.gyx, 2. * a.gyy);
    }
  }
};

template <bool G, bool LIM>
struct Dens32 {
  Soa d;
  Consts k;
  C32A c;
  using Lds = Lds32<G>;
  using Own = Own32;
  struct Acc {
    float dom = 0, domrho = 0, gamma = 0;
  };
  static __device__ Lds lds(unsigned char *b, int cap) { return carve32<G>(b, cap); }
  __device__ void stage(const Lds &L, int l, int j, double ox, double oy) const { stage32<G>(d, k, L, l, j, ox, oy); }
  __device__ bool active(int i) const {
    const int t = d.type[i];
    return t != kOutletF && t != TSPH_OUTLET && t != TSPH_INLET && t != TSPH_WALL && t != kRemoved;
  }
  __device__ Own own(int i, double ox, double oy) const { return own32<G>(d, k, i, ox, oy); }
  __device__ float2 pos_lds(const Lds &L, int l) const { return L.p[l]; }
  __device__ float2 pos_glob(int j, double ox, double oy) const {
    const double2 r = d.r[j];
    return make_float2((float)(r.x - ox), (float)(r.y - oy));
  }
  __device__ void pair(Acc &a, const Own &o, const Nb32 &b, float drx, float dry, float r2) const {
#pragma clang fp contract(fast)
    const KW32 kw = wendland32(c, r2);
    const float dWx = drx * kw.f, dWy = dry * kw.f;
    const float drnx = -drx * kw.inv, drny = -dry * kw.inv;
    const LR32 s = states32<G, LIM>(o, b, drx, dry);
    const float vl = s.vL.x * drnx + s.vL.y * drny;
    const float vr = s.vR.x * drnx + s.vR.y * drny;
    const float avg = c.rho0 + 0.5f * (s.aL + s.nR);
    const float dd = vl - vr;
    const float beta = fminf(3.f * fmaxf(dd, 0.f), c.c0);
    const float rsdev = 0.5f * (s.aL + s.nR) + 0.5f * beta * dd * avg * (c.ic0 * c.ic0);
    const float vss = 0.5f * (vr + vl) - 0.5f * (s.nR - s.aL) * c.c0 * __builtin_amdgcn_rcpf(avg);
    const float vsx = drnx * vss + (0.5f * (o.v.x + b.v.x) - 0.5f * drnx * (vl + vr));
    const float vsy = drny * vss + (0.5f * (o.v.y + b.v.y) - 0.5f * drny * (vl + vr));
    if (b.t == TSPH_WALL) {
      const float dvn = -((o.v.x - b.v.x) * b.n.x + (o.v.y - b.v.y) * b.n.y);
      a.dom -= dvn * kw.w * c.dp;
    } else {
      const float vsdW = (vsx - o.v.x) * dWx + (vsy - o.v.y) * dWy;
      a.domrho -= 2.f * (c.rho0 + rsdev) * b.omega * vsdW;
    }
    a.gamma += kw.w * c.m * __builtin_amdgcn_rcpf(c.rho0 + b.dev);
  }
  __device__ void pair_lds(Acc &a, const Own &o, const Lds &L, int l, float drx, float dry, float r2) const {
    pair(a, o, nb_lds32<G>(L, l), drx, dry, r2);
  }
  __device__ void pair_glob(Acc &a, const Own &o, int j, float drx, float dry, float r2) const {
    pair(a, o, nb_glob32<G>(d, k, j), drx, dry, r2);
  }
  __device__ void finish(const Acc &a, const Own &o, int i) const {
    if (a.gamma == 0.f) {
      d.domega[i] = 0.;
      d.domegarho[i] = 0.;
    } else {
      d.domega[i] = (double)a.dom * o.omega;
      d.domegarho[i] = (double)a.domrho * o.omega;
    }
  }
};

template <bool G, bool LIM>
struct Acc32P {
  Soa d;
  Consts k;
  C32A c;
  using Lds = Lds32<G>;
  using Own = Own32;
  struct Acc {
    float ax = 0, ay = 0, gamma = 0;
  };
  static __device__ Lds lds(unsigned char *b, int cap) { return carve32<G>(b, cap); }
  __device__ void stage(const Lds &L, int l, int j, double ox, double oy) const { stage32<G>(d, k, L, l, j, ox, oy); }
  __device__ bool active(int i) const { return d.type[i] == TSPH_FLUID; }
  __device__ Own own(int i, double ox, double oy) const { return own32<G>(d, k, i, ox, oy); }
  __device__ float2 pos_lds(const Lds &L, int l) const { return L.p[l]; }
  __device__ float2 pos_glob(int j, double ox, double oy) const {
    const double2 r = d.r[j];
    return make_float2((float)(r.x - ox), (float)(r.y - oy));
  }
  __device__ void pair(Acc &a, const Own &o, const Nb32 &b, float drx, float dry, float r2) const {
#pragma clang fp contract(fast)
    const KW32 kw = wendland32(c, r2);
    const float dWx = drx * kw.f, dWy = dry * kw.f;
    const float drnx = -drx * kw.inv, drny = -dry * kw.inv;
    const LR32 s = states32<G, LIM>(o, b, drx, dry);
    const float vl = s.vL.x * drnx + s.vL.y * drny;
    const float vr = s.vR.x * drnx + s.vR.y * drny;
    const float ir0 = 1.f / c.rho0;
    const float avgc = c.c0 * (__builtin_amdgcn_sqrtf(1.f + o.dev * ir0) + __builtin_amdgcn_sqrtf(1.f + b.dev * ir0));
    const float avg = c.rho0 + 0.5f * (s.aL + s.nR);
    const float dd = vl - vr;
    const float beta = fminf(3.f * fmaxf(dd, 0.f), avgc);
    const float rsdev = 0.5f * (s.aL + s.nR) + 0.5f * beta * dd * avg * __builtin_amdgcn_rcpf(avgc * avgc);
    const float vss = 0.5f * (vr + vl) - 0.5f * (b.dev - o.dev) * avgc * __builtin_amdgcn_rcpf(avg);
    const float vsx = drnx * vss + (0.5f * (o.v.x + b.v.x) - 0.5f * drnx * (vl + vr));
    const float vsy = drny * vss + (0.5f * (o.v.y + b.v.y) - 0.5f * drny * (vl + vr));
    if (b.t == TSPH_WALL) {
      const float nbx = -b.n.x, nby = -b.n.y;
      // wall state rho_w - rho0 = aL + (vL . nb) (rho0 + aL) / c0
      const float rwdev = s.aL + (s.vL.x * nbx + s.vL.y * nby) * (c.rho0 + s.aL) * c.ic0;
      const float pw = c.c0 * c.c0 * (c.pw0 + rwdev);
      a.ax -= 2.f * kw.w * c.dp * (pw * nbx);
      a.ay -= 2.f * kw.w * c.dp * (pw * nby);
    } else {
      const float ps = c.c0 * c.c0 * rsdev;
      const float rhos = c.rho0 + rsdev;
      const float ex = vsx - o.v.x, ey = vsy - o.v.y;
      a.ax -= 2.f * b.omega * ((rhos * vsx * ex + ps) * dWx + (rhos * vsx * ey) * dWy);
      a.ay -= 2.f * b.omega * ((rhos * vsy * ex) * dWx + (rhos * vsy * ey + ps) * dWy);
    }
    a.gamma += kw.w * c.m * __builtin_amdgcn_rcpf(c.rho0 + b.dev);
  }
  __device__ void pair_lds(Acc &a, const Own &o, const Lds &L, int l, float drx, float dry, float r2) const {
    pair(a, o, nb_lds32<G>(L, l), drx, dry, r2);
  }
  __device__ void pair_glob(Acc &a, const Own &o, int j, float drx, float dry, float r2) const {
    pair(a, o, nb_glob32<G>(d, k, j), drx, dry, r2);
  }
  __device__ void finish(const Acc &a, const Own &o, int i) const {
    const double gx = k.gx * o.omega * o.rho, gy = k.gy * o.omega * o.rho;
    if (a.gamma == 0.f)
      d.domegav[i] = make_double2(0. - gx, 0. - gy);
    else
      d.domegav[i] = make_double2((double)a.ax * o.omega - gx, (double)a.ay * o.omega - gy);
  }
};

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
template <class P, bool DYN>
__global__ __launch_bounds__(kWave * kPoolWaves) void k_ale_pass64(P pol, const int *__restrict__ cell_start,
                                                      const int *__restrict__ cellof, int rows,
                                                      int *__restrict__ scal,
                                                      const int *__restrict__ tile_list,
                                                      const int *__restrict__ bound, int cap,
                                                      int lds_wave,
                                                      unsigned long long *__restrict__ counters) {
  walk64<DYN>(pol, pol.k, cell_start, cellof, rows, scal, tile_list, bound, cap, lds_wave, counters);
}

template <class P, bool DYN>
__global__ __launch_bounds__(kWave * kPoolWaves) void k_ale_pass32(P pol, const int *__restrict__ cell_start,
                                                      const int *__restrict__ cellof, int rows,
                                                      int *__restrict__ scal,
                                                      const int *__restrict__ tile_list,
                                                      const int *__restrict__ bound, int cap,
                                                      int lds_wave,
                                                      unsigned long long *__restrict__ counters) {
  walk32<DYN>(pol, pol.k, cell_start, cellof, rows, scal, tile_list, bound, cap, lds_wave, counters);
}

// Integrate_density_compute_pressure, src/ALESPH_integration.h:86-135 (fluid
// only; linear EOS src/SPH_pressure.h:24-34, pressure clamped at zero)
__global__ void k_ale_density_pressure(Soa d, const int *__restrict__ scal, Consts k, double dt,
                                       unsigned long long *__restrict__ counters) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= scal[S_NP] || d.type[p] != TSPH_FLUID) return;
  const double om = d.omega[p] + d.domega[p] * dt;
  const double orho = d.omegarho[p] + d.domegarho[p] * dt;
  d.omega[p] = om;
  d.omegarho[p] = orho;
  const double rho = orho / (om);
  d.rho[p] = rho;
  if (rho > 1500. || rho < 0.) atomicAdd(&counters[C_BADRHO], 1ull);
  double pr = k.cs * k.cs * (rho - k.rho0);
  if (pr < 0) pr = 0;
  d.p[p] = pr;
}

// Integrate_LeapFrog_partOne_ALESPH, src/ALESPH_integration.h:9-31
__global__ void k_ale_lf1(Soa d, const int *__restrict__ scal, double dt) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= scal[S_NP] || d.type[p] != TSPH_FLUID) return;
  const double2 ov = d.omegav[p], dov = d.domegav[p];
  const double2 ov1 = make_double2(ov.x + dov.x * dt * 0.5, ov.y + dov.y * dt * 0.5);
  d.omegav[p] = ov1;
  const double ro = d.rho[p] * d.omega[p];
  d.v[p] = make_double2(ov1.x / ro, ov1.y / ro);
}

// Integrate_LeapFrog_partTwo_ALESPH, src/ALESPH_integration.h:33-83
__global__ void k_ale_lf2(Soa d, const int *__restrict__ scal, double dt, int first) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= scal[S_NP] || d.type[p] != TSPH_FLUID) return;
  const double2 ov = d.omegav[p], dov = d.domegav[p];
  const double ro = d.rho[p] * d.omega[p];
  const double2 r = d.r[p];
  if (first) {
    const double2 ov1 = make_double2(ov.x + dov.x * dt * 0.5, ov.y + dov.y * dt * 0.5);
    d.omegav[p] = ov1;
    const double2 v = make_double2(ov1.x / ro, ov1.y / ro);
    d.v[p] = v;
    d.r[p] = make_double2(r.x + v.x * dt * 0.5, r.y + v.y * dt * 0.5);
  } else {
    const double2 ov1 = make_double2(ov.x + dov.x * dt, ov.y + dov.y * dt);
    d.omegav[p] = ov1;
    const double2 v = make_double2(ov1.x / ro, ov1.y / ro);
    d.v[p] = v;
    d.r[p] = make_double2(r.x + v.x * dt, r.y + v.y * dt);
  }
}

// launch of one pair pass on the persistent tile grid
template <class P, int SLOT, bool FP32>
int launch_pass(tsph_ctx *c, const P &pol) {
  const int rows = c->tile_rows;
  // slots budgeted per tile, within 64 KiB of LDS per wavefront; a tile whose
  // neighbourhood does not fit is gathered from global memory instead.  The
  // ALE records are large (57 B mixed, 113 B FP64 with the MUSCL gradients),
  // so the budget is kAleCellCap records per cell, about a full cell's count
  // at the configured resolutions: the fullest tiles then take the global
  // path, and the smaller LDS area lets a compute unit hold more waves, which
  // pays more than staging every tile does
  constexpr int kAleCellCap = 9;
  constexpr int kFixed = kTabBytes + sizeof(uint16_t) * kList * kWave;
  const int cap = std::min(3 * (rows + 2) * kAleCellCap, (65536 - kPoolHdr - kFixed) / SLOT);
  // the static walk: a cell build serves several passes of different cost
  // profiles, and the dam breaks these passes run on give each wave about ten
  // tiles, too few for the pooled walk to pay for its chunk cut (183k
  // particles, FP64: 0.56 ms per step pooled against 0.51 static)
  const TileGrid tg = tile_grid(c, FP32, kFixed + (size_t)cap * SLOT, false);
  if (int e = ensure_tile_list(c, tg.pooled ? tg.groups : 0)) return e;
  return launch_timed(c, [&] {
    if constexpr (FP32) {
      if (tg.pooled)
        hipLaunchKernelGGL((k_ale_pass32<P, true>), tg.grid, tg.block, tg.lds, c->stream, pol,
                           c->cell_start, c->cellof, rows, c->scal, c->tile_list, c->tile_bound,
                           cap, tg.lds_wave, c->counters);
      else
        hipLaunchKernelGGL((k_ale_pass32<P, false>), tg.grid, tg.block, tg.lds, c->stream, pol,
                           c->cell_start, c->cellof, rows, c->scal, c->tile_list, c->tile_bound,
                           cap, tg.lds_wave, c->counters);
    } else {
      if (tg.pooled)
        hipLaunchKernelGGL((k_ale_pass64<P, true>), tg.grid, tg.block, tg.lds, c->stream, pol,
                           c->cell_start, c->cellof, rows, c->scal, c->tile_list, c->tile_bound,
                           cap, tg.lds_wave, c->counters);
      else
        hipLaunchKernelGGL((k_ale_pass64<P, false>), tg.grid, tg.block, tg.lds, c->stream, pol,
                           c->cell_start, c->cellof, rows, c->scal, c->tile_list, c->tile_bound,
                           cap, tg.lds_wave, c->counters);
    }
  });
}

}  // namespace

// muscl: 0 no gradients (first order), 1 gradients with alpha = 1, 2 minmod-limited
int launch_ale_density(tsph_ctx *c, int muscl) {
  const bool mixed = c->prm.precision == TSPH_PREC_MIXED;
  if (mixed) {
    const C32A cc = make_c32a(c->k);
    if (muscl == 0) return launch_pass<Dens32<false, false>, slot_bytes32<false>(), true>(c, {c->cur, c->k, cc});
    if (muscl == 1) return launch_pass<Dens32<true, false>, slot_bytes32<true>(), true>(c, {c->cur, c->k, cc});
    return launch_pass<Dens32<true, true>, slot_bytes32<true>(), true>(c, {c->cur, c->k, cc});
  }
  if (muscl == 0) return launch_pass<Dens64<false, false>, slot_bytes64<false>(), false>(c, {c->cur, c->k});
  if (muscl == 1) return launch_pass<Dens64<true, false>, slot_bytes64<true>(), false>(c, {c->cur, c->k});
  return launch_pass<Dens64<true, true>, slot_bytes64<true>(), false>(c, {c->cur, c->k});
}

int launch_ale_acceleration(tsph_ctx *c, int muscl) {
  const bool mixed = c->prm.precision == TSPH_PREC_MIXED;
  if (mixed) {
    const C32A cc = make_c32a(c->k);
    if (muscl == 0) return launch_pass<Acc32P<false, false>, slot_bytes32<false>(), true>(c, {c->cur, c->k, cc});
    if (muscl == 1) return launch_pass<Acc32P<true, false>, slot_bytes32<true>(), true>(c, {c->cur, c->k, cc});
    return launch_pass<Acc32P<true, true>, slot_bytes32<true>(), true>(c, {c->cur, c->k, cc});
  }
  if (muscl == 0) return launch_pass<Acc64P<false, false>, slot_bytes64<false>(), false>(c, {c->cur, c->k});
  if (muscl == 1) return launch_pass<Acc64P<true, false>, slot_bytes64<true>(), false>(c, {c->cur, c->k});
  return launch_pass<Acc64P<true, true>, slot_bytes64<true>(), false>(c, {c->cur, c->k});
}

int launch_ale_gradient(tsph_ctx *c) {
  if (c->prm.precision == TSPH_PREC_MIXED)
    return launch_pass<Grad32, slot_bytes32<false>(), true>(c, {c->cur, c->k, make_c32a(c->k)});
  return launch_pass<Grad64, slot_bytes64<false>(), false>(c, {c->cur, c->k});
}

int launch_ale_density_pressure(tsph_ctx *c, double dt) {
  hipLaunchKernelGGL(k_ale_density_pressure, dim3(grid_for(c->n_ub, 256)), dim3(256), 0, c->stream,
                     c->cur, c->scal, c->k, dt, c->counters);
  TSPH_HIP(hipGetLastError());
  c->records_valid = false;
  return 0;
}

int launch_ale_leapfrog_one(tsph_ctx *c, double dt) {
  hipLaunchKernelGGL(k_ale_lf1, dim3(grid_for(c->n_ub, 256)), dim3(256), 0, c->stream, c->cur,
                     c->scal, dt);
  TSPH_HIP(hipGetLastError());
  c->records_valid = false;
  return 0;
}

int launch_ale_leapfrog_two(tsph_ctx *c, double dt, int step) {
  hipLaunchKernelGGL(k_ale_lf2, dim3(grid_for(c->n_ub, 256)), dim3(256), 0, c->stream, c->cur,
                     c->scal, dt, (int)(step == 1));
  TSPH_HIP(hipGetLastError());
  c->records_valid = false;
  c->marks_valid = false;
  return 0;
}

}  // namespace tsph
