// tinysph_dropin.hpp -- host-side drop-in for the reference's stage library.
//
// A case driver written against the reference (struct SPH_simulation with
// INIT/RUN calling free functions on a Particle_system, e.g.
// cases/damBreakLong_mDBC/SPH_simulationSymplectic.h:24-203) compiles against
// this header instead of src/*.h and runs its per-step pipeline on the GPU
// through libtinysph_hip.so:
//   * Particle_system / SPH_data / SPH_constants / SPH_special / Simulation_data
//     keep the reference's field names (src/SPH_data.h:5-125,
//     src/SPH_particle_system.h:7-213) as host mirrors;
//   * the stage functions keep the reference signatures and call the C ABI;
//   * the geometry routines (src/SPH_create.h, src/SPH_create_bt.h) build the
//     host mirror exactly as the reference does.
// The pair-loop variant is chosen at compile time like the reference chooses it
// by #include: define TSPH_DROPIN_FORCES to TSPH_FORCES_WCSPH (default,
// src/SPH_interaction.h), TSPH_FORCES_DBC (src/SPH_interactionDBC.h) or
// TSPH_FORCES_BT (src/SPH_interactionBT.h) before including this header.  Define
// TSPH_DROPIN_RALE instead for the ALE drivers (cases/ALESPH_dambreakLong/):
// Compute_Forces is then src/RALESPH_interaction.h's, the leap-frog halves of
// both ALE schemes and the ALE-SPH MUSCL stages (Compute_Density,
// Kernel_gradient_approx, Integrate_density_compute_pressure,
// Compute_Acceleration_BT) are provided, and the geometry routines create ALE
// particles (src/ALESPH_create.h).  Buffer zones: DRAW_CreateInletY/X,
// DRAW_CreateOutletY, BZ_Integrate_inlet/outlet (src/SPH_buffer_zones.h).
//
// Data stays on the device between stages.  The host vectors are refreshed by
// tsph_sync_host(particles); host edits after the first stage call must be
// followed by tsph_mark_host_dirty(particles).  Errors throw tsph_error.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "tinysph.h"

#ifndef TSPH_DROPIN_FORCES
#define TSPH_DROPIN_FORCES TSPH_FORCES_WCSPH
#endif

// ---- src/SPH_defs.h:31-93 ----
#define SQ(x) ((x) * (x))
#define NORM(x, y) sqrt(SQ(x) + SQ(y))
#define LEN(x0, y0, x1, y1) sqrt(SQ(x1 - x0) + SQ(y1 - y0))
#define POS(x, y, X, Y) (y + (x) * (Y))

enum {
  fluid = 0, wall = 1, virt = 2, mov_a = 3, mov_b = 4, mov_c = 5, inlet = 6, inlet1 = 7,
  inlet2 = 8, inlet3 = 9, inlet5 = 10, outlet = 11, outlet1 = 12, outlet2 = 13, outlet3 = 14,
  outlet4 = 15, outlet5 = 16, outletf = 17
};

typedef struct { double x, y; } tdouble2;
inline tdouble2 TDouble2(double x, double y) { tdouble2 p = {x, y}; return p; }
inline tdouble2 operator+(const tdouble2 &a, const tdouble2 &b) { return TDouble2(a.x + b.x, a.y + b.y); }
inline tdouble2 operator-(const tdouble2 &a, const tdouble2 &b) { return TDouble2(a.x - b.x, a.y - b.y); }
inline tdouble2 operator*(const tdouble2 &a, const double &b) { return TDouble2(a.x * b, a.y * b); }
inline tdouble2 operator/(const tdouble2 &a, const double &b) { return TDouble2(a.x / b, a.y / b); }
using realvec = tdouble2;
using real = double;
using idx = int;

struct tsph_error : std::runtime_error {
  explicit tsph_error(const std::string &m) : std::runtime_error(m) {}
};

inline void tsph_check(int rc, const char *what) {
  if (rc != TSPH_OK) throw tsph_error(std::string(what) + ": " + tsph_last_error());
}

// ---- src/SPH_data.h:5-125 ----
struct SPH_data {
  std::vector<idx> index;
  std::vector<unsigned short int> part_type;
  std::vector<real> p, rho, drho;
  std::vector<realvec> r, v, a;
  std::vector<realvec> r_o, v_o, v_oo;
  std::vector<real> rho_o, rho_oo;
};

struct SPH_constants {
  real h, kap, dp, cs, rho0;
  realvec graviy;
  real avisc, delta, m;
};

struct SPH_special {
  std::vector<realvec> gnr;  // mDBC ghost nodes (walls only, in creation order)
  std::vector<realvec> n;    // BT normals (walls only, in creation order)
  // ALE-SPH volume variables (src/SPH_data.h:53-66), one entry per particle
  std::vector<real> omega, omegarho;
  std::vector<realvec> omegav;
  std::vector<real> omega_o, omegarho_o;
  std::vector<realvec> omegav_o;
  std::vector<real> domega, domegarho;
  std::vector<realvec> domegav;
  std::vector<real> rhoTemp;  // density-reinitialisation buffer, one entry per particle
};

struct Interaction_data {
  unsigned int ncx, ncy;
};

struct Simulation_data {
  real x_0, x_m, y_0, y_m;
  real hh, kh;
  int nvl;
  Simulation_data(real _x_0, real _x_m, real _y_0, real _y_m)
      : x_0(_x_0), x_m(_x_m), y_0(_y_0), y_m(_y_m), hh(0), kh(0), nvl(0) {}
};

// ---- src/SPH_particle_system.h:7-213 ----
// BufferZone, src/SPH_inlet_outlet_data.h:18-46 (the fields the buffer-zone
// stages read)
struct BufferZone {
  realvec orientation;
  real x0, y0, xm, ym;
  real bl, bh;
  unsigned int btype;
  realvec v_b;
  BufferZone(real _x0, real _xm, real _y0, real _ym, unsigned int, realvec _v0, realvec _orientation,
             unsigned int _btype)
      : orientation(_orientation), x0(_x0), y0(_y0), xm(_xm), ym(_ym), btype(_btype), v_b(_v0) {
    bl = std::fabs(xm - x0);
    bh = std::fabs(ym - y0);
    if (orientation.y != 0 && orientation.x == 0) { bl = std::fabs(ym - y0); bh = std::fabs(xm - x0); }
  }
};

struct Particle_system {
  unsigned int np;
  std::vector<BufferZone> zones;  // src/SPH_particle_system.h:22
  SPH_data data;
  SPH_constants data_const;
  SPH_special special;
  Interaction_data pairs;
  unsigned int np_out;

  // device side
  tsph_ctx *dev = nullptr;
  int precision = TSPH_PREC_FP64;
  // particle slots on the device; 0: the particle count at Divide_To_Cells,
  // plus max(4096, np/4) slots for inlet appends when buffer zones exist
  int device_capacity = 0;
  uint64_t removed_seen = 0;   // device removal count already added to np_out
  uint64_t rejected_seen = 0;  // device count of refused inlet appends already reported
  bool host_dirty = true;   // host vectors newer than the device
  bool dev_newer = false;   // device newer than the host vectors
  size_t zones_on_device = 0;
  bool history_on_device = false;  // v_oo, rho_oo, rhoTemp held on the device

  void Add_particle(short unsigned int _type, real _p, real _rho, realvec _r, realvec _v, realvec _a) {
    np++;
    data.part_type.push_back(_type);
    data.index.push_back(np - 1);
    data.p.push_back(_p);
    data.rho.push_back(_rho);
    data.drho.push_back(0.);
    data.r.push_back(_r);
    data.v.push_back(_v);
    data.a.push_back(_a);
    data.r_o.push_back(_r);
    data.rho_o.push_back(_rho);
    data.rho_oo.push_back(_rho);
    data.v_o.push_back(_v);
    data.v_oo.push_back(_v);
    special.rhoTemp.push_back(_rho);
    host_dirty = true;
  }

  // src/SPH_particle_system.h:51-108 (the volume variables the Riemann-ALE
  // path reads; the MUSCL gradients and *_oo copies are not used by it)
  void ALESPH_Add_particle(short unsigned int _type, real _p, real _rho, realvec _r, realvec _v,
                           realvec _a) {
    Add_particle(_type, _p, _rho, _r, _v, _a);
    const real w = data_const.dp * data_const.dp;
    special.omega.push_back(w);
    special.omegav.push_back(_v * w);
    special.omegarho.push_back(_rho * w);
    special.omega_o.push_back(w);
    special.omegav_o.push_back(_v * w);
    special.omegarho_o.push_back(_rho * w);
    special.domega.push_back(0.);
    special.domegarho.push_back(0.);
    special.domegav.push_back(TDouble2(0., 0.));
  }

  Particle_system(double _h, double _kap, double _cs, double _rho0, double _visco, double _delta,
                  double _dp) {
    np = 0;
    np_out = 0;
    pairs.ncx = pairs.ncy = 0;
    data_const.h = _h;
    data_const.kap = _kap;
    data_const.cs = _cs;
    data_const.rho0 = _rho0;
    data_const.dp = _dp;
    data_const.graviy = {0., 9.81};
    data_const.m = _rho0 / (1 / _dp * 1 / _dp);
    data_const.avisc = _visco;
    data_const.delta = _delta;
  }
  Particle_system(const Particle_system &) = delete;
  Particle_system &operator=(const Particle_system &) = delete;
  ~Particle_system() {
    if (dev) tsph_destroy(dev);
  }
};

// ---------------------------------------------------------------------------
// host <-> device
// ---------------------------------------------------------------------------
inline void tsph_mark_host_dirty(Particle_system &ps) { ps.host_dirty = true; }

// per-particle copy of a wall-only extra (gnr / n): walls come first in every
// reference case, so entry k of the extra belongs to particle k
inline std::vector<double> tsph_expand_extra(const Particle_system &ps, const std::vector<realvec> &e) {
  std::vector<double> out(2 * (size_t)ps.np, 0.0);
  for (size_t k = 0; k < e.size() && k < ps.np; ++k) {
    out[2 * k] = e[k].x;
    out[2 * k + 1] = e[k].y;
  }
  return out;
}

inline void tsph_push_zones(Particle_system &ps) {
  for (; ps.zones_on_device < ps.zones.size(); ++ps.zones_on_device) {
    const BufferZone &b = ps.zones[ps.zones_on_device];
    tsph_zone z{};
    z.x0 = b.x0; z.xm = b.xm; z.y0 = b.y0; z.ym = b.ym; z.vx = b.v_b.x; z.vy = b.v_b.y;
    z.ox = (int)b.orientation.x; z.oy = (int)b.orientation.y; z.btype = (int)b.btype;
    tsph_check(tsph_bz_add_zone(ps.dev, &z, nullptr), "tsph_bz_add_zone");
  }
}

inline void tsph_push(Particle_system &ps) {
  if (!ps.dev) throw tsph_error("Divide_To_Cells must be called before the first stage");
  tsph_push_zones(ps);
  if (!ps.host_dirty) return;
  const int n = (int)ps.np;
  std::vector<int32_t> type(n), index(n);
  for (int i = 0; i < n; ++i) {
    type[i] = ps.data.part_type[i];
    index[i] = ps.data.index[i];
  }
  auto flat = [n](const std::vector<realvec> &v) {
    std::vector<double> o(2 * (size_t)n);
    for (int i = 0; i < n; ++i) { o[2 * i] = v[i].x; o[2 * i + 1] = v[i].y; }
    return o;
  };
  std::vector<double> r = flat(ps.data.r), v = flat(ps.data.v), a = flat(ps.data.a),
                      ro = flat(ps.data.r_o), vo = flat(ps.data.v_o);
  std::vector<double> gnr, nn;
  tsph_particles h{};
  h.np = n;
  h.type = type.data();
  h.index = index.data();
  h.p = ps.data.p.data();
  h.rho = ps.data.rho.data();
  h.drho = ps.data.drho.data();
  h.r = r.data(); h.v = v.data(); h.a = a.data(); h.r_o = ro.data(); h.v_o = vo.data();
  h.rho_o = ps.data.rho_o.data();
  if (!ps.special.gnr.empty()) { gnr = tsph_expand_extra(ps, ps.special.gnr); h.gnr = gnr.data(); }
  if (!ps.special.n.empty()) { nn = tsph_expand_extra(ps, ps.special.n); h.n = nn.data(); }
  tsph_check(tsph_upload(ps.dev, &h), "tsph_upload");
  if (!ps.special.omega.empty()) {
    const SPH_special &sp = ps.special;
    std::vector<double> ov = flat(sp.omegav), ovo = flat(sp.omegav_o), dov = flat(sp.domegav);
    tsph_ale_fields f{};
    f.np = n;
    f.omega = const_cast<double *>(sp.omega.data());
    f.omegarho = const_cast<double *>(sp.omegarho.data());
    f.omega_o = const_cast<double *>(sp.omega_o.data());
    f.omegarho_o = const_cast<double *>(sp.omegarho_o.data());
    f.domega = const_cast<double *>(sp.domega.data());
    f.domegarho = const_cast<double *>(sp.domegarho.data());
    f.omegav = ov.data(); f.omegav_o = ovo.data(); f.domegav = dov.data();
    tsph_check(tsph_ale_upload(ps.dev, &f), "tsph_ale_upload");
  }
  if (ps.history_on_device) {
    std::vector<double> voo = flat(ps.data.v_oo);
    tsph_history hh{};
    hh.np = n;
    hh.v_oo = voo.data();
    hh.rho_oo = ps.data.rho_oo.data();
    hh.rho_temp = ps.special.rhoTemp.data();
    tsph_check(tsph_history_upload(ps.dev, &hh), "tsph_history_upload");
  }
  ps.host_dirty = false;
}

// The Verlet time layer and the reinitialisation buffer move to the device on
// the first stage that needs them.  The device sets them from its current state
// (v_oo = v, rho_oo = rho, rhoTemp = rho); the host's values, which are still
// Add_particle's, replace them while the particle set is the host's (no
// particle removed or added on the device yet).
inline void tsph_need_history(Particle_system &ps) {
  tsph_push(ps);
  if (ps.history_on_device) return;
  tsph_check(tsph_history_enable(ps.dev), "tsph_history_enable");
  ps.history_on_device = true;
  const int n = (int)ps.np;
  std::vector<double> voo(2 * (size_t)n);
  for (int i = 0; i < n; ++i) { voo[2 * i] = ps.data.v_oo[i].x; voo[2 * i + 1] = ps.data.v_oo[i].y; }
  tsph_history hh{};
  hh.np = n;
  hh.v_oo = voo.data();
  hh.rho_oo = ps.data.rho_oo.data();
  hh.rho_temp = ps.special.rhoTemp.data();
  const int rc = tsph_history_upload(ps.dev, &hh);
  if (rc != TSPH_OK && rc != TSPH_EINVAL) tsph_check(rc, "tsph_history_upload");
}

inline void tsph_sync_host(Particle_system &ps) {
  if (!ps.dev || !ps.dev_newer) return;
  int n = 0;
  tsph_check(tsph_get_np(ps.dev, &n), "tsph_get_np");
  std::vector<int32_t> type(n), index(n);
  std::vector<double> p(n), rho(n), drho(n), rho_o(n), r(2 * (size_t)n), v(2 * (size_t)n),
      a(2 * (size_t)n), ro(2 * (size_t)n), vo(2 * (size_t)n);
  tsph_particles h{};
  h.type = type.data(); h.index = index.data(); h.p = p.data(); h.rho = rho.data();
  h.drho = drho.data(); h.rho_o = rho_o.data(); h.r = r.data(); h.v = v.data(); h.a = a.data();
  h.r_o = ro.data(); h.v_o = vo.data();
  tsph_check(tsph_download(ps.dev, &h), "tsph_download");
  const int m = h.np;
  auto vec = [m](const std::vector<double> &f) {
    std::vector<realvec> o(m);
    for (int i = 0; i < m; ++i) o[i] = {f[2 * i], f[2 * i + 1]};
    return o;
  };
  // np_out counts removals (src/SPH_particle_system.h:115), never appends:
  // take the device's removal count since the last synchronisation
  tsph_counters cnt{};
  tsph_check(tsph_get_counters(ps.dev, &cnt), "tsph_get_counters");
  if (cnt.removed < ps.removed_seen) ps.removed_seen = 0;  // counters were reset
  ps.np_out += (unsigned)(cnt.removed - ps.removed_seen);
  ps.removed_seen = cnt.removed;
  ps.np = (unsigned)m;
  ps.data.part_type.assign(type.begin(), type.begin() + m);
  ps.data.index.assign(index.begin(), index.begin() + m);
  ps.data.p.assign(p.begin(), p.begin() + m);
  ps.data.rho.assign(rho.begin(), rho.begin() + m);
  ps.data.drho.assign(drho.begin(), drho.begin() + m);
  ps.data.rho_o.assign(rho_o.begin(), rho_o.begin() + m);
  ps.data.r = vec(r); ps.data.v = vec(v); ps.data.a = vec(a); ps.data.r_o = vec(ro); ps.data.v_o = vec(vo);
  ps.data.v_oo.resize(m);
  ps.data.rho_oo.resize(m);
  ps.special.rhoTemp.resize(m);
  if (ps.history_on_device) {
    std::vector<double> voo(2 * (size_t)n);
    tsph_history hh{};
    hh.v_oo = voo.data();
    hh.rho_oo = ps.data.rho_oo.data();
    hh.rho_temp = ps.special.rhoTemp.data();
    tsph_check(tsph_history_download(ps.dev, &hh), "tsph_history_download");
    ps.data.v_oo = vec(voo);
  }
  if (!ps.special.omega.empty()) {
    SPH_special &sp = ps.special;
    std::vector<double> om(n), orho(n), omo(n), orhoo(n), dom(n), dorho(n), ov(2 * (size_t)n),
        ovo(2 * (size_t)n), dov(2 * (size_t)n);
    tsph_ale_fields f{};
    f.omega = om.data(); f.omegarho = orho.data(); f.omega_o = omo.data();
    f.omegarho_o = orhoo.data(); f.domega = dom.data(); f.domegarho = dorho.data();
    f.omegav = ov.data(); f.omegav_o = ovo.data(); f.domegav = dov.data();
    tsph_check(tsph_ale_download(ps.dev, &f), "tsph_ale_download");
    sp.omega.assign(om.begin(), om.begin() + m);
    sp.omegarho.assign(orho.begin(), orho.begin() + m);
    sp.omega_o.assign(omo.begin(), omo.begin() + m);
    sp.omegarho_o.assign(orhoo.begin(), orhoo.begin() + m);
    sp.domega.assign(dom.begin(), dom.begin() + m);
    sp.domegarho.assign(dorho.begin(), dorho.begin() + m);
    sp.omegav = vec(ov); sp.omegav_o = vec(ovo); sp.domegav = vec(dov);
  }
  ps.dev_newer = false;
}

inline void tsph_ran(Particle_system &ps) { ps.dev_newer = true; }

// ---------------------------------------------------------------------------
// stage functions (reference signatures)
// ---------------------------------------------------------------------------

// src/SPH_find_pairs.h:5-42: sizes the grid and creates the device context
inline void Divide_To_Cells(Particle_system &ps, Simulation_data sd) {
  tsph_params prm{};
  prm.h = ps.data_const.h; prm.kap = ps.data_const.kap; prm.cs = ps.data_const.cs;
  prm.rho0 = ps.data_const.rho0; prm.avisc = ps.data_const.avisc; prm.delta = ps.data_const.delta;
  prm.dp = ps.data_const.dp;
  prm.x_0 = sd.x_0; prm.x_m = sd.x_m; prm.y_0 = sd.y_0; prm.y_m = sd.y_m;
  prm.precision = ps.precision;
  const int n0 = (int)(ps.np > 0 ? ps.np : 1);
  // an inlet appends up to one buffer layer per pass, and appended particles
  // stay until an outlet or the domain bound removes them
  const int room = ps.zones.empty() ? 0 : std::max(4096, n0 / 4);
  prm.capacity = ps.device_capacity > 0 ? ps.device_capacity : n0 + room;
  if (ps.dev) tsph_destroy(ps.dev);
  ps.dev = nullptr;
  ps.zones_on_device = 0;
  ps.removed_seen = ps.rejected_seen = 0;
  tsph_check(tsph_create(&prm, &ps.dev), "tsph_create");
  int ncx = 0, ncy = 0;
  tsph_check(tsph_grid(ps.dev, &ncx, &ncy), "tsph_grid");
  ps.pairs.ncx = (unsigned)ncx;
  ps.pairs.ncy = (unsigned)ncy;
  ps.host_dirty = true;
}

// src/SPH_find_pairs.h:74-92
inline void Clear_cells(Particle_system &ps) {
  if (ps.dev) tsph_check(tsph_clear_cells(ps.dev), "Clear_cells");
}

// src/SPH_find_pairs.h:45-72
inline void Particle_To_Cell(Particle_system &ps, Simulation_data) {
  tsph_push(ps);
  tsph_check(tsph_particle_to_cell(ps.dev), "Particle_To_Cell");
  tsph_ran(ps);
}

// src/SPH_pressure.h:65-89
inline void Density_to_pressure(Particle_system &ps) {
  tsph_push(ps);
  tsph_check(tsph_density_to_pressure(ps.dev), "Density_to_pressure");
  tsph_ran(ps);
}

#ifndef TSPH_DROPIN_RALE
// src/SPH_interaction.h:8 / src/SPH_interactionDBC.h:8 / src/SPH_interactionBT.h:8
inline void Compute_Forces(Particle_system &ps) {
  tsph_push(ps);
  tsph_check(tsph_compute_forces(ps.dev, TSPH_DROPIN_FORCES), "Compute_Forces");
  tsph_ran(ps);
}
#else
// src/RALESPH_interaction.h:12-287 (first-order Riemann-ALE fluxes)
inline void Compute_Forces(Particle_system &ps) {
  tsph_push(ps);
  tsph_check(tsph_rale_forces(ps.dev), "Compute_Forces");
  tsph_ran(ps);
}
// src/RALESPH_integration.h:10-36
inline void Integrate_LeapFrog_partOne(Particle_system &ps, double dt) {
  tsph_push(ps);
  tsph_check(tsph_rale_leapfrog_one(ps.dev, dt), "Integrate_LeapFrog_partOne");
  tsph_ran(ps);
}
// src/RALESPH_integration.h:38-107
inline void Integrate_LeapFrog_partTwo(Particle_system &ps, double dt, int step) {
  tsph_push(ps);
  tsph_check(tsph_rale_leapfrog_two(ps.dev, dt, step), "Integrate_LeapFrog_partTwo");
  tsph_ran(ps);
}
// ALE-SPH with MUSCL reconstruction (cases/ALESPH_dambreakLong/SPH_simulation.h).
// TSPH_DROPIN_MINMOD selects the minmod-weighted reconstruction.
#ifndef TSPH_DROPIN_MINMOD
#define TSPH_DROPIN_MINMOD 0
#endif
// src/ALESPH_density_and_omega_bt.h:9-281
inline void Compute_Density(Particle_system &ps) {
  tsph_push(ps);
  tsph_check(tsph_ale_compute_density(ps.dev, TSPH_DROPIN_MINMOD), "Compute_Density");
  tsph_ran(ps);
}
// src/ALESPH_kernel_gradient.h:10-190
inline void Kernel_gradient_approx(Particle_system &ps) {
  tsph_push(ps);
  tsph_check(tsph_ale_kernel_gradient(ps.dev), "Kernel_gradient_approx");
  tsph_ran(ps);
}
// src/ALESPH_integration.h:86-135
inline void Integrate_density_compute_pressure(Particle_system &ps, double dt) {
  tsph_push(ps);
  tsph_check(tsph_ale_density_pressure(ps.dev, dt), "Integrate_density_compute_pressure");
  tsph_ran(ps);
}
// src/ALESPH_acceleration_bt.h:12-295
inline void Compute_Acceleration_BT(Particle_system &ps) {
  tsph_push(ps);
  tsph_check(tsph_ale_compute_acceleration(ps.dev, TSPH_DROPIN_MINMOD), "Compute_Acceleration_BT");
  tsph_ran(ps);
}
// src/ALESPH_integration.h:9-31
inline void Integrate_LeapFrog_partOne_ALESPH(Particle_system &ps, double dt) {
  tsph_push(ps);
  tsph_check(tsph_ale_leapfrog_one(ps.dev, dt), "Integrate_LeapFrog_partOne_ALESPH");
  tsph_ran(ps);
}
// src/ALESPH_integration.h:33-83
inline void Integrate_LeapFrog_partTwo_ALESPH(Particle_system &ps, double dt, int step) {
  tsph_push(ps);
  tsph_check(tsph_ale_leapfrog_two(ps.dev, dt, step), "Integrate_LeapFrog_partTwo_ALESPH");
  tsph_ran(ps);
}
#endif

// src/SPH_mDBC.h:9-50
inline void mDBC_compute_density_mdbcGeo(Particle_system &ps, Simulation_data) {
  tsph_push(ps);
  tsph_check(tsph_mdbc_density(ps.dev), "mDBC_compute_density_mdbcGeo");
  tsph_ran(ps);
}

// src/SPH_BT.h:9-50
inline void BT_compute_density_BTGeo(Particle_system &ps, Simulation_data) {
  tsph_push(ps);
  tsph_check(tsph_bt_density(ps.dev), "BT_compute_density_BTGeo");
  tsph_ran(ps);
}

// src/SPH_integration.h:266-309
inline void Integrate_SymplecticPredictor(Particle_system &ps, double dt) {
  tsph_push(ps);
  tsph_check(tsph_symplectic_predictor(ps.dev, dt), "Integrate_SymplecticPredictor");
  tsph_ran(ps);
}
inline void Integrate_SymplecticCorrector(Particle_system &ps, double dt) {
  tsph_push(ps);
  tsph_check(tsph_symplectic_corrector(ps.dev, dt), "Integrate_SymplecticCorrector");
  tsph_ran(ps);
}

#ifndef TSPH_DROPIN_RALE
// src/SPH_integration.h:103-128
inline void Integrate_Verlet(Particle_system &ps, double dt) {
  tsph_need_history(ps);
  tsph_check(tsph_verlet(ps.dev, dt), "Integrate_Verlet");
  tsph_ran(ps);
}
// src/SPH_integration.h:130-158
inline void Integrate_Euler(Particle_system &ps, double dt) {
  tsph_need_history(ps);
  tsph_check(tsph_euler(ps.dev, dt), "Integrate_Euler");
  tsph_ran(ps);
}
// src/SPH_integration.h:161-183
inline void Integrate_LeapFrog_partOne_withDensity(Particle_system &ps, double dt) {
  tsph_push(ps);
  tsph_check(tsph_leapfrog_one(ps.dev, dt), "Integrate_LeapFrog_partOne_withDensity");
  tsph_ran(ps);
}
// src/SPH_integration.h:185-230
inline void Integrate_LeapFrog_partTwo_withDensity(Particle_system &ps, double dt, int step) {
  tsph_push(ps);
  tsph_check(tsph_leapfrog_two(ps.dev, dt, step), "Integrate_LeapFrog_partTwo_withDensity");
  tsph_ran(ps);
}
// src/SPH_density_reinitialization.h:8-121
inline void Density_reinitialization_sumation(Particle_system &ps) {
  tsph_need_history(ps);
  tsph_check(tsph_density_reinit_sum(ps.dev), "Density_reinitialization_sumation");
  tsph_ran(ps);
}
// src/SPH_density_reinitialization.h:123-139 (the warnings are counted in
// tsph_counters.reinit_out_of_range)
inline void Density_reinitialization_overwrite(Particle_system &ps) {
  tsph_need_history(ps);
  tsph_check(tsph_density_reinit_overwrite(ps.dev), "Density_reinitialization_overwrite");
  tsph_ran(ps);
}
#endif

// src/SPH_btdebug.h:4-31
inline void RemoveParticlesOutOfDomain(Particle_system &ps, Simulation_data) {
  tsph_push(ps);
  tsph_check(tsph_remove_out_of_domain(ps.dev), "RemoveParticlesOutOfDomain");
  tsph_ran(ps);
}

// src/SPH_water_level.h:87-150
inline real Water_Elevation(Particle_system &ps, Simulation_data, real x, real y0) {
  tsph_push(ps);
  double wl = 0;
  tsph_check(tsph_water_elevation(ps.dev, x, y0, 0, &wl), "Water_Elevation");
  return wl;
}
inline real Water_Elevation_inverse(Particle_system &ps, Simulation_data, real x, real y0) {
  tsph_push(ps);
  double wl = 0;
  tsph_check(tsph_water_elevation(ps.dev, x, y0, 1, &wl), "Water_Elevation_inverse");
  return wl;
}
// src/SPH_measure_pressure.h:6-15
inline real Pressure_in_point(Particle_system &ps, Simulation_data, realvec ar) {
  tsph_push(ps);
  double out[4];
  tsph_check(tsph_interpolate_point(ps.dev, ar.x, ar.y, out), "Pressure_in_point");
  return out[1];
}

// src/SPH_data_to_vtk.h:4-51.  With the state on the device the copies are
// queued at the call and the file is written in the background
// (tsph_io_wait(particles.dev) joins); before the first stage it is written
// from the host vectors.
inline void write_to_ASCII_VTK(Particle_system &ps, std::string output_file_name) {
  if (ps.dev) {
    tsph_push(ps);
    tsph_check(tsph_write_vtk(ps.dev, output_file_name.c_str()), "write_to_ASCII_VTK");
    return;
  }
  const int n = (int)ps.np;
  std::vector<int32_t> type(n);
  std::vector<double> r(2 * (size_t)n), v(2 * (size_t)n);
  for (int i = 0; i < n; ++i) {
    type[i] = ps.data.part_type[i];
    r[2 * i] = ps.data.r[i].x; r[2 * i + 1] = ps.data.r[i].y;
    v[2 * i] = ps.data.v[i].x; v[2 * i + 1] = ps.data.v[i].y;
  }
  tsph_particles h{};
  h.np = n;
  h.type = type.data(); h.r = r.data(); h.v = v.data();
  h.p = ps.data.p.data(); h.rho = ps.data.rho.data();
  tsph_check(tsph_format_vtk(&h, output_file_name.c_str()), "write_to_ASCII_VTK");
}

// src/SPH_dataInterpolation.h:13-91 (grid on the device, file as
// write_INTERPOLATED_mesh_to_ASCII_VTK, src/SPH_grid_to_vtk.h:139-243)
inline void GenerateInterpol(Particle_system &ps, Simulation_data, std::string filename, real x0,
                             real y0, real xm, real ym) {
  tsph_push(ps);
  tsph_check(tsph_write_interpolated_vtk(ps.dev, x0, y0, xm, ym, filename.c_str()),
             "GenerateInterpol");
}

// write_mesh_to_ASCII_VTK, src/SPH_grid_to_vtk.h:6-67 (the reference takes the
// system by value; the cell counts are the device's last Particle_To_Cell)
inline void write_mesh_to_ASCII_VTK(Particle_system &ps, Simulation_data sd, std::string filename) {
  tsph_push(ps);
  tsph_check(tsph_write_grid_vtk(ps.dev, sd.x_0, sd.y_0, filename.c_str()), "write_mesh_to_ASCII_VTK");
}

// Output_hydrostatic_pressure, src/SPH_measure_pressure.h:17-34
inline void Output_hydrostatic_pressure(Particle_system &ps, std::string file_name) {
  tsph_push(ps);
  tsph_check(tsph_write_hydrostatic_pressure(ps.dev, file_name.c_str()), "Output_hydrostatic_pressure");
}

// ---- the drivers' text output, src/SPH_output_info.h (host only) ----
// Output_file, src/SPH_output_info.h:8-43: the constants and the padded domain
inline void Output_file(SPH_constants constants, Simulation_data simulation_data,
                        std::string info_file_name) {
  std::ofstream f(info_file_name);
  auto line = [&f](const char *label, real value) { f << label << value << '\n'; };
  f << "--- SIMULATION INFO --- " << '\n' << "\n" << "-> SPH constants: " << '\n';
  line("Alpha: ", constants.avisc);
  line("Cs: ", constants.cs);
  line("Dp: ", constants.dp);
  f << "External forces: [" << constants.graviy.x << "," << constants.graviy.y << "]" << '\n';
  line("h: ", constants.h);
  line("Kap: ", constants.kap);
  line("m: ", constants.m);
  line("rho0: ", constants.rho0);
  f << "\n" << "-> Simulation info: " << '\n';
  line("h = h(dp): ", simulation_data.hh);
  line("kh: ", simulation_data.kh);
  f << "Number of virual layers (for walls) nvl: " << simulation_data.nvl << '\n';
  f << "\n" << "-> Domain size (with respect to cells): " << '\n';
  f << "[x_min, y_min]: [" << simulation_data.x_0 << " , " << simulation_data.y_0 << "] " << '\n';
  f << "[x_max, y_max]: [" << simulation_data.x_m << " , " << simulation_data.y_m << "] " << '\n';
  f.close();
  std::cout << "Info file generated." << std::endl;
}
// src/SPH_output_info.h:46-56 (empty in the reference)
inline void Add_to_output_file(std::string) {}
inline void Write_basic_simulation_info() {}
// src/SPH_output_info.h:58-85: one "time value" line appended per call
inline void tsph_append_series(const std::string &path, real t, real value) {
  std::ofstream f(path, std::ios_base::app);
  f << t << " " << value << '\n';
}
inline void Write_water_level(real step, real wl, std::string output_file_name_wl) {
  tsph_append_series(output_file_name_wl, step, wl);
}
inline void Write_pressure(real step, real p, std::string output_file_name_p) {
  tsph_append_series(output_file_name_p, step, p);
}
inline void Write_kinetic_energy(real step, real p, std::string output_file_name_p) {
  tsph_append_series(output_file_name_p, step, p);
}

// src/SPH_buffer_zones.h:10-107.  The reference appends without bound; the
// device has a fixed number of slots (device_capacity), so a pass whose appends
// do not fit is refused whole and reported here as an error instead of letting
// the inlet silently stop feeding.
inline void BZ_Integrate_inlet(Particle_system &ps, unsigned int zoneIndex, double dt) {
  tsph_push(ps);
  tsph_check(tsph_bz_integrate_inlet(ps.dev, (int)zoneIndex, dt), "BZ_Integrate_inlet");
  tsph_ran(ps);
  tsph_counters cnt{};
  tsph_check(tsph_get_counters(ps.dev, &cnt), "tsph_get_counters");
  if (cnt.zone_rejected < ps.rejected_seen) ps.rejected_seen = 0;  // counters were reset
  if (cnt.zone_rejected > ps.rejected_seen) {
    const uint64_t k = cnt.zone_rejected - ps.rejected_seen;
    ps.rejected_seen = cnt.zone_rejected;
    throw tsph_error("BZ_Integrate_inlet: " + std::to_string(k) +
                     " inlet appends refused (device capacity; raise Particle_system::device_capacity)");
  }
}
// src/SPH_buffer_zones.h:111-206
inline void BZ_Integrate_outlet(Particle_system &ps, unsigned int zoneIndex, double dt) {
  tsph_push(ps);
  tsph_check(tsph_bz_integrate_outlet(ps.dev, (int)zoneIndex, dt), "BZ_Integrate_outlet");
  tsph_ran(ps);
}

// ---------------------------------------------------------------------------
// geometry, src/SPH_create.h and src/SPH_create_bt.h (host side, same arithmetic)
// With TSPH_DROPIN_RALE the routines add ALE particles, as their namesakes in
// src/ALESPH_create.h do (same positions, ALESPH_Add_particle).
// ---------------------------------------------------------------------------
inline void tsph_add_particle(Particle_system &ps, short unsigned int type, real p, real rho,
                              realvec r, realvec v, realvec a) {
#ifdef TSPH_DROPIN_RALE
  ps.ALESPH_Add_particle(type, p, rho, r, v, a);
#else
  ps.Add_particle(type, p, rho, r, v, a);
#endif
}

// src/SPH_pressure.h:49-63
inline real Pressure_to_density(real p, real rho0, real c0) {
  const real b_const = c0 * c0 * rho0 / 7.;
  return (pow(p / b_const + 1, 1. / 7)) * rho0;
}

// src/SPH_create.h:546-581
inline void Create_Wall_x_mdbc(Particle_system &ps, double x_0, double x_1, double y, int orientation,
                               double y_w) {
  const real dp = ps.data_const.dp, kh = ps.data_const.h * ps.data_const.kap;
  const real len = LEN(x_0, y, x_1, y);
  const int npart = (int)(len / dp);
  const int nvl = std::ceil(kh / dp);
  realvec r, v = {0, 0}, a = {0, 0}, gn;
  for (int l = 0; l < nvl; l++) {
    r.y = y + dp * l * orientation;
    for (int i = 0; i < npart + 1; i++) {
      r.x = x_0 + i * dp;
      tsph_add_particle(ps, wall, 0, ps.data_const.rho0, r, v, a);
      gn.x = r.x;
      gn.y = y_w + fabs(r.y - y_w);
      ps.special.gnr.push_back(gn);
    }
  }
}

// src/SPH_create.h:583-619
inline void Create_Wall_y_mdbc(Particle_system &ps, double x, double y_0, double y_1, int orientation,
                               double x_w) {
  const real dp = ps.data_const.dp, kh = ps.data_const.h * ps.data_const.kap;
  const real len = LEN(x, y_0, x, y_1);
  const int npart = (int)(len / dp);
  const int nvl = std::ceil(kh / dp);
  realvec r, v = {0, 0}, a = {0, 0}, gn;
  for (int l = 0; l < nvl; l++) {
    r.x = x + dp * l * orientation;
    for (int i = 0; i < npart + 1; i++) {
      r.y = y_0 + i * dp;
      tsph_add_particle(ps, wall, 0, ps.data_const.rho0, r, v, a);
      gn.x = x_w + fabs(r.x - x_w) * (-1) * (orientation);
      gn.y = r.y;
      ps.special.gnr.push_back(gn);
    }
  }
}

// src/SPH_create.h:621-652
inline void Create_corner_mdbc(Particle_system &ps, int orientationX, int orientationY, realvec xyv) {
  const real dp = ps.data_const.dp, kh = ps.data_const.h * ps.data_const.kap;
  const int nvl = std::ceil(kh / dp);
  realvec r, v = {0, 0}, a = {0, 0}, gn;
  for (int l = 0; l < nvl; l++) {
    r.y = xyv.y + dp * l * orientationY + dp / 2 * orientationY;
    for (int i = 0; i < nvl; i++) {
      r.x = xyv.x + dp * i * orientationX + dp / 2 * orientationX;
      tsph_add_particle(ps, wall, 0, ps.data_const.rho0, r, v, a);
      gn.x = xyv.x + fabs(r.x - xyv.x) * (-1) * orientationX;
      gn.y = xyv.y + fabs(r.y - xyv.y) * (-1) * orientationY;
      ps.special.gnr.push_back(gn);
    }
  }
}

// src/SPH_create.h:105-139
inline void Create_Block_of_Fluid(Particle_system &ps, double x_0, double y_0, double x_1, double y_1) {
  const real dp = ps.data_const.dp;
  const real len_x = LEN(x_0, y_0, x_1, y_0), len_y = LEN(x_0, y_0, x_0, y_1);
  const int npart_x = (int)(len_x / dp), npart_y = (int)(len_y / dp);
  realvec r, v = {0, 0}, a = {0, 0};
  for (int i = 0; i < npart_x + 1; i++) {
    r.x = x_0 + i * dp;
    for (int j = 0; j < npart_y + 1; j++) {
      r.y = y_0 + j * dp;
      tsph_add_particle(ps, fluid, 0, ps.data_const.rho0, r, v, a);
    }
  }
}

// src/SPH_create.h:689-727 (hydrostatic pressure and density)
inline void Create_Block_of_FluidWP(Particle_system &ps, double x_0, double y_0, double x_1, double y_1) {
  const real dp = ps.data_const.dp;
  const real len_x = LEN(x_0, y_0, x_1, y_0), len_y = LEN(x_0, y_0, x_0, y_1);
  const int npart_x = (int)(len_x / dp), npart_y = (int)(len_y / dp);
  realvec r, v = {0, 0}, a = {0, 0};
  for (int i = 0; i < npart_x + 1; i++) {
    r.x = x_0 + i * dp;
    for (int j = 0; j < npart_y + 1; j++) {
      r.y = y_0 + j * dp;
      const real rhoin = Pressure_to_density(9.81 * (y_1 - r.y) * 1000, 1000., ps.data_const.cs);
      tsph_add_particle(ps, fluid, 9.81 * (y_1 - r.y) * 1000, rhoin, r, v, a);
    }
  }
}

// src/SPH_create_bt.h:49-82
inline void Create_Wall_x_bt(Particle_system &ps, double x_0, double x_1, double y, int orientation, realvec n) {
  const real dp = ps.data_const.dp;
  const real len = LEN(x_0, y, x_1, y);
  const int npart = (int)(len / dp);
  realvec r, v = {0, 0}, a = {0, 0};
  for (int l = 0; l < 1; l++) {
    r.y = y + dp * l * orientation;
    for (int i = 0; i < npart + 1; i++) {
      r.x = x_0 + i * dp;
      tsph_add_particle(ps, wall, 0, ps.data_const.rho0, r, v, a);
      ps.special.n.push_back(n);
    }
  }
}

// src/SPH_create_bt.h:84-118
inline void Create_Wall_y_bt(Particle_system &ps, double x, double y_0, double y_1, int orientation, realvec n) {
  const real dp = ps.data_const.dp;
  const real len = LEN(x, y_0, x, y_1);
  const int npart = (int)(len / dp);
  realvec r, v = {0, 0}, a = {0, 0};
  for (int l = 0; l < 1; l++) {
    r.x = x + dp * l * orientation;
    for (int i = 0; i < npart + 1; i++) {
      r.y = y_0 + i * dp;
      tsph_add_particle(ps, wall, 0, ps.data_const.rho0, r, v, a);
      ps.special.n.push_back(n);
    }
  }
}

// src/SPH_create_bt.h:120-149
inline void Create_corner_bt(Particle_system &ps, int orientationX, int orientationY, realvec xyv, realvec n) {
  const real dp = ps.data_const.dp;
  realvec r, v = {0, 0}, a = {0, 0};
  for (int l = 0; l < 1; l++) {
    r.y = xyv.y + dp * l * orientationY;
    for (int i = 0; i < 1; i++) {
      r.x = xyv.x + dp * i * orientationX;
      tsph_add_particle(ps, wall, 0, ps.data_const.rho0, r, v, a);
      ps.special.n.push_back(n);
    }
  }
}

// src/SPH_create.h:439-470
inline void DRAW_CreateInletY(Particle_system &ps, double x, double y_0, double y_1, int orientation,
                              int nbl, realvec v0, int type) {
  const real dp = ps.data_const.dp;
  const real len = LEN(x, y_0, x, y_1);
  const unsigned int npart = std::ceil(len / dp) + 1;
  realvec r, v, a;
  for (int l = 0; l < nbl; l++) {
    r.x = x + dp * l * orientation;
    for (unsigned h = 0; h < npart; h++) {
      r.y = y_0 + h * dp;
      v.x = v0.x; v.y = v0.y;
      a.x = 0; a.y = 0;
      ps.Add_particle(type, 0, ps.data_const.rho0, r, v, a);
    }
  }
  ps.zones.push_back(BufferZone(x, x + nbl * dp * orientation, y_0, y_1, nbl, v0, {(real)orientation, 0}, type));
}

// src/SPH_create.h:472-503
inline void DRAW_CreateInletX(Particle_system &ps, double y, double x_0, double x_1, int orientation,
                              int nbl, realvec v0, int type) {
  const real dp = ps.data_const.dp;
  const real len = LEN(x_0, y, x_1, y);
  const unsigned int npart = std::ceil(len / dp) + 1;
  realvec r, v, a;
  for (int l = 0; l < nbl; l++) {
    r.y = y + dp * l * orientation;
    for (unsigned h = 0; h < npart; h++) {
      r.x = x_0 + h * dp;
      v.x = v0.x; v.y = v0.y;
      a.x = 0; a.y = 0;
      ps.Add_particle(type, 0, ps.data_const.rho0, r, v, a);
    }
  }
  ps.zones.push_back(BufferZone(x_0, x_1, y, y + nbl * dp * orientation, nbl, v0, {0, (real)orientation}, type));
}

// src/SPH_create.h:508-542 (a non-empty outlet is filled with particles of type inlet)
inline void DRAW_CreateOutletY(Particle_system &ps, double x, double y_0, double y_1, int orientation,
                               int nbl, realvec v0, int type, bool empty) {
  const real dp = ps.data_const.dp;
  if (empty == false) {
    const real len = LEN(x, y_0, x, y_1);
    const unsigned int npart = std::ceil(len / dp) + 1;
    realvec r, v, a;
    for (int l = 0; l < nbl; l++) {
      r.x = x + dp * l * orientation;
      for (unsigned h = 0; h < npart; h++) {
        r.y = y_0 + h * dp;
        v.x = v0.x; v.y = v0.y;
        a.x = 0; a.y = 0;
        ps.Add_particle(inlet, 0, ps.data_const.rho0, r, v, a);
      }
    }
  }
  ps.zones.push_back(BufferZone(x, x + nbl * dp * orientation, y_0, y_1, nbl, v0, {(real)orientation, 0}, type));
}
