// Case presets: the reference's case setup (constants + geometry + domain
// padding of PREP_SIMULATION_DATA) built with the drop-in geometry routines,
// exported as flat arrays through tsph_case_build.  Host code only.
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <memory>

#include "../../include/tinysph_dropin.hpp"

namespace {

// SplitMix64 (Steele, Lea, Flood 2014): 64-bit state, golden-ratio increment
struct SplitMix64 {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
};

struct Spec {
  double dp, hh, kap, cs, visco, delta, dt;
  double x_0, x_1, y_0, y_1;  // unpadded domain (constants.h)
  int steps, wall;
  int scheme = TSPH_SCHEME_SYMPLECTIC;
};

// PREP_SIMULATION_DATA: pad the domain by nvl*dp on every side
// (cases/damBreakLong_mDBC/SPH_simulationSymplectic.h:206-221, constants.h:23)
Simulation_data padded(const Spec &s) {
  Simulation_data sd(s.x_0, s.x_1, s.y_0, s.y_1);
  const int nvl = std::ceil(s.kap * s.hh / s.dp);
  sd.hh = s.hh;
  sd.kh = s.hh * s.kap;
  sd.nvl = nvl;
  sd.x_0 -= nvl * s.dp;
  sd.x_m += nvl * s.dp;
  sd.y_0 -= nvl * s.dp;
  sd.y_m += nvl * s.dp;
  return sd;
}

template <typename T>
T *dup(const std::vector<T> &v) {
  T *p = (T *)std::malloc(sizeof(T) * (v.empty() ? 1 : v.size()));
  if (!p) return nullptr;
  if (!v.empty()) std::memcpy(p, v.data(), sizeof(T) * v.size());
  return p;
}

std::vector<double> flat(const std::vector<realvec> &v) {
  std::vector<double> o(2 * v.size());
  for (size_t i = 0; i < v.size(); ++i) { o[2 * i] = v[i].x; o[2 * i + 1] = v[i].y; }
  return o;
}

int export_case(const Spec &s, Particle_system &ps, tsph_case **out) {
  const Simulation_data sd = padded(s);
  tsph_case *c = (tsph_case *)std::calloc(1, sizeof(tsph_case));
  if (!c) return TSPH_EINVAL;
  c->prm.h = s.hh; c->prm.kap = s.kap; c->prm.cs = s.cs; c->prm.rho0 = ps.data_const.rho0;
  c->prm.avisc = s.visco; c->prm.delta = s.delta; c->prm.dp = s.dp;
  c->prm.x_0 = sd.x_0; c->prm.x_m = sd.x_m; c->prm.y_0 = sd.y_0; c->prm.y_m = sd.y_m;
  c->prm.precision = TSPH_PREC_FP64;
  // room for the particles an inlet appends: one buffer layer per pass is
  // the most a stable time step converts
  c->prm.capacity = (int)ps.np + (ps.zones.empty() ? 0 : std::max(4096, (int)ps.np / 4));
  c->dt = s.dt;
  c->wall = s.wall;
  c->steps = s.steps;
  c->scheme = s.scheme;
  const int n = (int)ps.np;
  std::vector<int32_t> type(n), index(n);
  for (int i = 0; i < n; ++i) {
    type[i] = ps.data.part_type[i];
    index[i] = ps.data.index[i];
    if (type[i] == fluid) c->n_fluid++;
    if (type[i] == wall) c->n_wall++;
  }
  tsph_particles &h = c->particles;
  h.np = n;
  h.type = dup(type);
  h.index = dup(index);
  h.p = dup(ps.data.p);
  h.rho = dup(ps.data.rho);
  h.drho = dup(ps.data.drho);
  h.r = dup(flat(ps.data.r));
  h.v = dup(flat(ps.data.v));
  h.a = dup(flat(ps.data.a));
  h.r_o = dup(flat(ps.data.r_o));
  h.v_o = dup(flat(ps.data.v_o));
  h.rho_o = dup(ps.data.rho_o);
  if (!ps.special.gnr.empty()) h.gnr = dup(tsph_expand_extra(ps, ps.special.gnr));
  if (!ps.special.n.empty()) h.n = dup(tsph_expand_extra(ps, ps.special.n));
  for (const BufferZone &b : ps.zones) {
    if (c->n_zones >= 4) break;
    tsph_zone &z = c->zones[c->n_zones++];
    z.x0 = b.x0; z.xm = b.xm; z.y0 = b.y0; z.ym = b.ym; z.vx = b.v_b.x; z.vy = b.v_b.y;
    z.ox = (int)b.orientation.x; z.oy = (int)b.orientation.y; z.btype = (int)b.btype;
  }
  *out = c;
  return 0;
}

void jitter_fluid(Particle_system &ps, double amp, uint64_t seed) {
  if (!(amp > 0)) return;
  SplitMix64 g{seed};
  for (unsigned i = 0; i < ps.np; ++i) {
    if (ps.data.part_type[i] != fluid) continue;
    const double jx = (2.0 * g.uniform() - 1.0) * amp;
    const double jy = (2.0 * g.uniform() - 1.0) * amp;
    ps.data.r[i].x += jx;
    ps.data.r[i].y += jy;
    ps.data.r_o[i] = ps.data.r[i];
  }
}

}  // namespace

extern "C" int tsph_case_build(int preset, double dx, double jitter, uint64_t seed, tsph_case **out) {
  if (!out) return TSPH_EINVAL;
  *out = nullptr;
  const double g = 9.81;
  Spec s{};
  switch (preset) {
    case TSPH_CASE_STILL_WATER_DBC: {
      // cases/stillWater_DBC/{constants,geometry}.h with a 1.0 m tank, 0.5 m of
      // water, h = 1.5 dx, kap = 2, c0 = 10 sqrt(g H), dt = 0.2 h / c0
      const double W = 1.0, H = 0.5;
      s.dp = dx > 0 ? dx : 0.01;
      s.hh = 1.5 * s.dp; s.kap = 2.0; s.cs = 10. * std::sqrt(g * H);
      s.visco = 0.01; s.delta = 0.1; s.dt = 0.2 * s.hh / s.cs;
      s.x_0 = 0.; s.x_1 = W + 0.2; s.y_0 = 0.; s.y_1 = 1.;
      s.steps = 2000; s.wall = TSPH_WALL_DBC;
      Particle_system ps(s.hh, s.kap, s.cs, 1000., s.visco, s.delta, s.dp);
      const double dp = s.dp;
      Create_Wall_x_mdbc(ps, 0. + dp, W - dp, 0., -1, dp / 2);
      Create_Wall_y_mdbc(ps, 0., 0 + dp, 1., -1, dp / 2);
      Create_Wall_y_mdbc(ps, W, 0 + dp, 1., +1, W - dp / 2);
      Create_corner_mdbc(ps, -1, -1, {0. + dp / 2, 0. + dp / 2});
      Create_corner_mdbc(ps, +1, -1, {W - dp / 2, 0. + dp / 2});
      Create_Block_of_FluidWP(ps, 0 + dp, 0 + dp, W - dp, H);
      jitter_fluid(ps, (jitter > 0 ? jitter : 0.0) * s.dp, seed);
      return export_case(s, ps, out);
    }
    case TSPH_CASE_DAM_BREAK_MDBC: {
      // cases/damBreakLong_mDBC/{constants,geometry}.h scaled to a 3.22 x 1.8 m
      // tank and a 1.2 x 0.6 m column; h = sqrt(2) dx, c0 = 10 sqrt(g 0.6)
      const double L = 3.22, Ht = 1.8, CW = 1.2, CH = 0.6;
      s.dp = dx > 0 ? dx : 0.001;
      s.hh = 1.0 * std::sqrt(2 * s.dp * s.dp); s.kap = 2.1; s.cs = 10. * std::sqrt(g * CH);
      s.visco = 0.02; s.delta = 0.1; s.dt = 0.2 * s.hh / s.cs;
      s.x_0 = 0.; s.x_1 = L + 0.1; s.y_0 = 0.; s.y_1 = Ht;
      s.steps = 1000; s.wall = TSPH_WALL_MDBC;
      Particle_system ps(s.hh, s.kap, s.cs, 1000., s.visco, s.delta, s.dp);
      const double dp = s.dp;
      Create_Wall_x_mdbc(ps, 0. + dp, L, 0., -1, dp / 2);
      Create_Wall_y_mdbc(ps, 0., 0 + dp, Ht, -1, dp / 2);
      Create_Wall_y_mdbc(ps, L, 0 + dp, Ht, +1, L - dp / 2);
      Create_corner_mdbc(ps, -1, -1, {0. + dp / 2, 0. + dp / 2});
      Create_corner_mdbc(ps, +1, -1, {L - dp / 2, 0. + dp / 2});
      Create_Block_of_Fluid(ps, 0 + dp, 0 + dp, CW, CH);
      jitter_fluid(ps, (jitter >= 0 ? jitter : 0.01) * s.dp, seed);
      return export_case(s, ps, out);
    }
    case TSPH_CASE_DAM_BREAK_MDBC_REF: {
      // cases/damBreakLong_mDBC/constants.h and geometry.h as they stand (the
      // tank of the water-level gauges and the wall pressure sensors,
      // cases/damBreakLong_mDBC/constants.h:58-94); dt = 5e-5 s at dx = 5 mm,
      // scaled with dx
      s.dp = dx > 0 ? dx : 0.005;
      s.hh = 1.0 * std::sqrt(2 * s.dp * s.dp); s.kap = 2.1; s.cs = 34.3;
      s.visco = 0.02; s.delta = 0.1; s.dt = 0.00005 * (s.dp / 0.005);
      s.x_0 = 0.; s.x_1 = 2.; s.y_0 = 0.; s.y_1 = 1.;
      s.steps = 45001; s.wall = TSPH_WALL_MDBC;
      Particle_system ps(s.hh, s.kap, s.cs, 1000., s.visco, s.delta, s.dp);
      const double dp = s.dp;
      Create_Wall_x_mdbc(ps, 0. + dp, 1.61, 0., -1, dp / 2);
      Create_Wall_y_mdbc(ps, 0., 0 + dp, 1., -1, dp / 2);
      Create_Wall_y_mdbc(ps, 1.61, 0 + dp, 1., +1, 1.61 - dp / 2);
      Create_corner_mdbc(ps, -1, -1, {0. + dp / 2, 0. + dp / 2});
      Create_corner_mdbc(ps, +1, -1, {1.61 - dp / 2, 0. + dp / 2});
      Create_Block_of_Fluid(ps, 0 + dp, 0 + dp, 0.6, 0.3);
      jitter_fluid(ps, (jitter > 0 ? jitter : 0.0) * s.dp, seed);
      return export_case(s, ps, out);
    }
    case TSPH_CASE_DAM_BREAK_BT: {
      // cases/damBreakLong_BT/{constants,geometry}.h
      s.dp = dx > 0 ? dx : 0.005;
      s.hh = 1.0 * std::sqrt(2 * s.dp * s.dp); s.kap = 2.1; s.cs = 34.3;
      s.visco = 0.02; s.delta = 0.1; s.dt = 0.00004 * (s.dp / 0.005);
      s.x_0 = 0.; s.x_1 = 2.; s.y_0 = 0.; s.y_1 = 1.;
      s.steps = 1000; s.wall = TSPH_WALL_BT;
      Particle_system ps(s.hh, s.kap, s.cs, 1000., s.visco, s.delta, s.dp);
      const double dp = s.dp;
      Create_Wall_x_bt(ps, 0. + dp, 1.61 - dp, 0., -1, {0., 1.});
      Create_Wall_y_bt(ps, 0., 0. + dp, 1., -1, {1., 0.});
      Create_Wall_y_bt(ps, 1.61, 0. + dp, 1., +1, {-1., 0.});
      Create_corner_bt(ps, -1, -1, {0., 0.}, {std::sqrt(2), std::sqrt(2)});
      Create_corner_bt(ps, +1, -1, {1.61, 0.}, {-std::sqrt(2), std::sqrt(2)});
      Create_Block_of_FluidWP(ps, 0.0 + dp, 0 + dp, 0.6, 0.3);
      jitter_fluid(ps, (jitter > 0 ? jitter : 0.0) * s.dp, seed);
      return export_case(s, ps, out);
    }
    case TSPH_CASE_ALE_DAM_BREAK: {
      // cases/ALESPH_dambreakLong/geometry.h (BT walls, fluid block at rho0; the
      // ALE variables follow from ALESPH_Add_particle, see tsph_ale_enable) in
      // the 3.22 x 1.8 m tank with the 1.2 x 0.6 m column; c0 = 10 sqrt(g 0.6)
      const double L = 3.22, Ht = 1.8, CW = 1.2, CH = 0.6;
      s.dp = dx > 0 ? dx : 0.002;
      s.hh = 1.0 * std::sqrt(2 * s.dp * s.dp); s.kap = 2.1; s.cs = 10. * std::sqrt(g * CH);
      s.visco = 0.02; s.delta = 0.1; s.dt = 0.2 * s.hh / s.cs;
      s.x_0 = 0.; s.x_1 = L + 0.1; s.y_0 = 0.; s.y_1 = Ht;
      s.steps = 1000; s.wall = TSPH_WALL_BT; s.scheme = TSPH_SCHEME_RALE;
      Particle_system ps(s.hh, s.kap, s.cs, 1000., s.visco, s.delta, s.dp);
      const double dp = s.dp;
      Create_Wall_x_bt(ps, 0. + dp, L - dp, 0., -1, {0., 1.});
      Create_Wall_y_bt(ps, 0., 0. + dp, Ht, -1, {1., 0.});
      Create_Wall_y_bt(ps, L, 0. + dp, Ht, +1, {-1., 0.});
      Create_corner_bt(ps, -1, -1, {0., 0.}, {std::sqrt(2), std::sqrt(2)});
      Create_corner_bt(ps, +1, -1, {L, 0.}, {-std::sqrt(2), std::sqrt(2)});
      Create_Block_of_Fluid(ps, 0.0 + dp, 0 + dp, CW, CH);
      jitter_fluid(ps, (jitter > 0 ? jitter : 0.0) * s.dp, seed);
      return export_case(s, ps, out);
    }
    case TSPH_CASE_INFLOW_SLAB: {
      // cases/inflowSlab_mDBC/geometry.h with cases/inflowSlab_mDBC/constants.h
      // (dt scaled with dx from 5e-5 at dx = 5 mm); inlet zone 0, outlet zone 1
      s.dp = dx > 0 ? dx : 0.005;
      s.hh = 1.0 * std::sqrt(2 * s.dp * s.dp); s.kap = 2.1; s.cs = 42.48;
      s.visco = 0.1; s.delta = 0.1; s.dt = 0.00005 * s.dp / 0.005;
      s.x_0 = 0.; s.x_1 = 2.; s.y_0 = 0.; s.y_1 = 1.;
      s.steps = 50000; s.wall = TSPH_WALL_MDBC;
      Particle_system ps(s.hh, s.kap, s.cs, 1000., s.visco, s.delta, s.dp);
      const double dp = s.dp;
      Create_Wall_x_mdbc(ps, 0.0, 0.4 - 3 * dp, 0.1, -1, 0.1 + dp / 2);
      Create_Wall_y_mdbc(ps, 0.4, 0 + dp, 0.1 - 3 * dp, -1, 0.4 + dp / 2);
      Create_Wall_x_mdbc(ps, 0.4 + dp, 1.2, 0., -1, 0 + dp / 2);
      Create_corner_mdbc(ps, -1, -1, {0.4 + dp / 2, 0. + dp / 2});
      Create_corner_mdbc(ps, -1, -1, {0.4 + dp / 2, 0.1 + dp / 2});
      DRAW_CreateInletY(ps, 0.1, 0.1 + dp, 0.2, -1, 4, {0.75, 0.}, inlet1);
      DRAW_CreateOutletY(ps, 1.0, 0.0 + dp, 0.5, +1, 4, {0.5, 0.}, outlet1, true);
      return export_case(s, ps, out);
    }
    case TSPH_CASE_INFLOW_CHANNEL: {
      // the inflow driver on an 8 x 0.5 m channel: an mDBC floor, 0.125 m of
      // water (hydrostatic density) moving at 1 m/s, a 4-layer inlet at x = 0
      // feeding at 1 m/s (zone 0) and an empty outlet from x = 7.9 (zone 1);
      // c0 = 10 sqrt(g H), dt = 0.2 h / c0
      const double L = 8.0, H = 0.125, Hc = 0.5, xo = 7.9, U = 1.0;
      s.dp = dx > 0 ? dx : 0.0005;
      s.hh = 1.0 * std::sqrt(2 * s.dp * s.dp); s.kap = 2.1; s.cs = 10. * std::sqrt(g * H);
      s.visco = 0.02; s.delta = 0.1; s.dt = 0.2 * s.hh / s.cs;
      const double dp = s.dp;
      const int nvl = std::ceil(s.kap * s.hh / s.dp);
      s.x_0 = -(4 + nvl) * dp; s.x_1 = L + 0.1; s.y_0 = 0.; s.y_1 = Hc;
      s.steps = 1000; s.wall = TSPH_WALL_MDBC;
      Particle_system ps(s.hh, s.kap, s.cs, 1000., s.visco, s.delta, s.dp);
      Create_Wall_x_mdbc(ps, -(4 + nvl) * dp, L, 0., -1, 0 + dp / 2);
      DRAW_CreateInletY(ps, 0.0, 0.0 + dp, H, -1, 4, {U, 0.}, inlet1);
      const unsigned first_fluid = ps.np;
      Create_Block_of_FluidWP(ps, dp, dp, xo - dp, H);
      for (unsigned i = first_fluid; i < ps.np; ++i) {
        ps.data.v[i] = {U, 0.};
        ps.data.v_o[i] = ps.data.v[i];
        ps.data.v_oo[i] = ps.data.v[i];
      }
      DRAW_CreateOutletY(ps, xo, 0.0 + dp, Hc, +1, 4, {U, 0.}, outlet1, true);
      jitter_fluid(ps, (jitter > 0 ? jitter : 0.0) * s.dp, seed);
      return export_case(s, ps, out);
    }
    default:
      return TSPH_EINVAL;
  }
}

namespace {

// One tank of the batched ensemble: the dam-break set-up of
// cases/damBreakLong_mDBC/geometry.h (mDBC floor, side walls and corners) in a
// W x Ht tank, with either still water over the whole width or a dam-break
// column half the tank wide.  Constants as the mDBC dam break: h = sqrt(2) dx,
// kap = 2.1, c0 = 10 sqrt(g depth), dt = 0.2 h / c0.
int build_tank(double W, double Ht, double depth, bool dam, double dx, uint64_t seed,
               tsph_case **out) {
  const double g = 9.81;
  Spec s{};
  s.dp = dx;
  s.hh = 1.0 * std::sqrt(2 * s.dp * s.dp); s.kap = 2.1; s.cs = 10. * std::sqrt(g * depth);
  s.visco = 0.02; s.delta = 0.1; s.dt = 0.2 * s.hh / s.cs;
  s.x_0 = 0.; s.x_1 = W + 0.1; s.y_0 = 0.; s.y_1 = Ht;
  s.steps = 500; s.wall = TSPH_WALL_MDBC;
  Particle_system ps(s.hh, s.kap, s.cs, 1000., s.visco, s.delta, s.dp);
  const double dp = s.dp;
  Create_Wall_x_mdbc(ps, 0. + dp, W, 0., -1, dp / 2);
  Create_Wall_y_mdbc(ps, 0., 0 + dp, Ht, -1, dp / 2);
  Create_Wall_y_mdbc(ps, W, 0 + dp, Ht, +1, W - dp / 2);
  Create_corner_mdbc(ps, -1, -1, {0. + dp / 2, 0. + dp / 2});
  Create_corner_mdbc(ps, +1, -1, {W - dp / 2, 0. + dp / 2});
  if (dam) Create_Block_of_Fluid(ps, 0 + dp, 0 + dp, 0.5 * W, depth);
  else Create_Block_of_Fluid(ps, 0 + dp, 0 + dp, W - dp, depth);
  jitter_fluid(ps, 0.01 * s.dp, seed);
  return export_case(s, ps, out);
}

}  // namespace

extern "C" int tsph_ensemble_build(int ncases, uint64_t seed, int nmin, int nmax,
                                   tsph_case **cases) {
  if (!cases || ncases < 1 || nmin < 1 || nmax < nmin) return TSPH_EINVAL;
  for (int q = 0; q < ncases; ++q) cases[q] = nullptr;
  SplitMix64 g{seed};
  for (int q = 0; q < ncases; ++q) {
    const double W = 0.5 + 2.5 * g.uniform();
    const double fill = 0.3 + 0.5 * g.uniform();
    double dx = 0.002 + 0.008 * g.uniform();
    const bool dam = g.uniform() < 0.5;
    const uint64_t jseed = g.next();
    const double Ht = 2.0 * W;  // tank height rule: twice the width
    const double depth = fill * Ht;
    // fluid particles of the block on a dx lattice; dx is rescaled so that the
    // count lies in [nmin, nmax]
    const double area = (dam ? 0.5 * W : W - dx) * depth;
    const double est = area / (dx * dx);
    if (est > nmax) dx = std::sqrt(area / nmax);
    else if (est < nmin) dx = std::sqrt(area / nmin);
    if (int e = build_tank(W, Ht, depth, dam, dx, jseed, &cases[q])) {
      for (int u = 0; u < q; ++u) {
        tsph_case_free(cases[u]);
        cases[u] = nullptr;
      }
      return e;
    }
  }
  return 0;
}

extern "C" int tsph_case_free(tsph_case *c) {
  if (!c) return 0;
  tsph_particles &h = c->particles;
  void *ps[] = {h.type, h.index, h.p, h.rho, h.drho, h.r, h.v, h.a, h.r_o, h.v_o, h.rho_o, h.gnr, h.n};
  for (void *p : ps) std::free(p);
  std::free(c);
  return 0;
}
