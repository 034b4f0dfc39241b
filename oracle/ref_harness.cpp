// Test infrastructure only: a C-ABI harness around the reference's own stage
// functions, compiled from the reference headers where they lie (see
// oracle/Makefile).  It is used by tests/ to pin the oracle restatement
// (oracle/tsph_oracle.cpp) and may serve as the "reference" CPU baseline.
// Nothing in the product links or loads it.
//
// One shared object is built per pair-loop variant because every variant
// header defines a function named Compute_Forces:
//   REF_WCSPH  src/SPH_interaction.h     + src/SPH_mDBC.h  (mDBC cases)
//   REF_DBC    src/SPH_interactionDBC.h                   (DBC cases)
//   REF_BT     src/SPH_interactionBT.h   + src/SPH_BT.h    (BT cases)
// It is built without -fopenmp: the reference kernel returns a pointer to a
// static buffer (src/SPH_kernel.h:54), which is only well defined serially.

#include "SPH_defs.h"
#include "SPH_particle_system.h"
#include "SPH_find_pairs.h"
#include "SPH_pressure.h"
#if defined(REF_DBC)
#include "SPH_interactionDBC.h"
#elif defined(REF_BT)
#include "SPH_interactionBT.h"
#include "SPH_BT.h"
#elif defined(REF_RALE)
#include "RALESPH_interaction.h"
#else
#include "SPH_interaction.h"
#include "SPH_mDBC.h"
#endif
#if defined(REF_RALE)
#include "RALESPH_integration.h"
#include "ALESPH_density_and_omega_bt.h"
#include "ALESPH_acceleration_bt.h"
#include "ALESPH_kernel_gradient.h"
#include "ALESPH_integration.h"
#else
#include "SPH_integration.h"
#include "SPH_density_reinitialization.h"
#endif
#include "SPH_btdebug.h"
#include "SPH_water_level.h"
#include "SPH_measure_pressure.h"
#if defined(REF_BT)
#include "SPH_create_bt.h"
#elif defined(REF_RALE)
#include "ALESPH_create.h"
#else
#include "SPH_create.h"
#endif
#if defined(REF_WCSPH)
#include "SPH_buffer_zones.h"
#include "SPH_data_to_vtk.h"
#include "SPH_grid_to_vtk.h"
#include "SPH_output_info.h"
#endif

namespace {

struct RefCase {
  Particle_system ps;
  Simulation_data sd;
  RefCase(const double* q)
      : ps(q[0], q[1], q[2], q[3], q[4], q[5], q[6]), sd(q[7], q[8], q[9], q[10]) {
    sd.hh = q[0];
    sd.kh = q[0] * q[1];
    sd.nvl = 0;
  }
};

// Field ids: 0 r, 1 v, 2 a, 3 p, 4 rho, 5 drho, 6 r_o, 7 v_o, 8 rho_o, 9 gnr,
// 10 n, ALE: 11 omega, 12 omegarho, 13 omegav, 14 omega_o, 15 omegarho_o,
// 16 omegav_o, 17 domega, 18 domegarho, 19 domegav, MUSCL: 20 gradrho,
// 21 gradvx, 22 gradvy, history: 23 v_oo, 24 rho_oo, 25 rhoTemp.
std::vector<realvec>* vec2_field(RefCase* c, int f) {
  switch (f) {
    case 0: return &c->ps.data.r;
    case 1: return &c->ps.data.v;
    case 2: return &c->ps.data.a;
    case 6: return &c->ps.data.r_o;
    case 7: return &c->ps.data.v_o;
    case 9: return &c->ps.special.gnr;
    case 10: return &c->ps.special.n;
    case 13: return &c->ps.special.omegav;
    case 16: return &c->ps.special.omegav_o;
    case 19: return &c->ps.special.domegav;
    case 20: return &c->ps.special.gradrho;
    case 21: return &c->ps.special.gradvx;
    case 22: return &c->ps.special.gradvy;
    case 23: return &c->ps.data.v_oo;
    default: return nullptr;
  }
}

std::vector<real>* scal_field(RefCase* c, int f) {
  switch (f) {
    case 3: return &c->ps.data.p;
    case 4: return &c->ps.data.rho;
    case 5: return &c->ps.data.drho;
    case 8: return &c->ps.data.rho_o;
    case 11: return &c->ps.special.omega;
    case 12: return &c->ps.special.omegarho;
    case 14: return &c->ps.special.omega_o;
    case 15: return &c->ps.special.omegarho_o;
    case 17: return &c->ps.special.domega;
    case 18: return &c->ps.special.domegarho;
    case 24: return &c->ps.data.rho_oo;
    case 25: return &c->ps.special.rhoTemp;
    default: return nullptr;
  }
}

}  // namespace

extern "C" {

// q = {h, kap, cs, rho0, avisc, delta, dp, x_0, x_m, y_0, y_m} (padded domain)
void* ref_new(const double* q) { return new RefCase(q); }
void ref_free(void* c) { delete static_cast<RefCase*>(c); }

void ref_add(void* c, int type, double p, double rho, double rx, double ry,
             double vx, double vy) {
  static_cast<RefCase*>(c)->ps.Add_particle((unsigned short)type, p, rho, {rx, ry},
                                           {vx, vy}, {0., 0.});
}
void ref_add_ale(void* c, int type, double p, double rho, double rx, double ry, double vx,
                 double vy) {
  static_cast<RefCase*>(c)->ps.ALESPH_Add_particle((unsigned short)type, p, rho, {rx, ry},
                                                  {vx, vy}, {0., 0.});
}
void ref_push_gnr(void* c, double x, double y) {
  static_cast<RefCase*>(c)->ps.special.gnr.push_back({x, y});
}
void ref_push_n(void* c, double x, double y) {
  static_cast<RefCase*>(c)->ps.special.n.push_back({x, y});
}

int ref_np(void* c) { return (int)static_cast<RefCase*>(c)->ps.np; }
int ref_np_out(void* c) { return (int)static_cast<RefCase*>(c)->ps.np_out; }
double ref_mass(void* c) { return static_cast<RefCase*>(c)->ps.data_const.m; }

// Field ids: 0 r, 1 v, 2 a, 3 p, 4 rho, 5 drho, 6 r_o, 7 v_o, 8 rho_o, 9 gnr, 10 n.
// Returns the number of doubles written/read (0 for an unknown field).
int ref_field_len(void* vc, int f) {
  RefCase* c = static_cast<RefCase*>(vc);
  if (auto* v = vec2_field(c, f)) return 2 * (int)v->size();
  if (auto* s = scal_field(c, f)) return (int)s->size();
  return 0;
}
int ref_get(void* vc, int f, double* out) {
  RefCase* c = static_cast<RefCase*>(vc);
  if (auto* v = vec2_field(c, f)) {
    for (size_t i = 0; i < v->size(); ++i) { out[2 * i] = (*v)[i].x; out[2 * i + 1] = (*v)[i].y; }
    return 2 * (int)v->size();
  }
  if (auto* s = scal_field(c, f)) {
    for (size_t i = 0; i < s->size(); ++i) out[i] = (*s)[i];
    return (int)s->size();
  }
  return 0;
}
int ref_set(void* vc, int f, const double* in) {
  RefCase* c = static_cast<RefCase*>(vc);
  if (auto* v = vec2_field(c, f)) {
    for (size_t i = 0; i < v->size(); ++i) { (*v)[i].x = in[2 * i]; (*v)[i].y = in[2 * i + 1]; }
    return 2 * (int)v->size();
  }
  if (auto* s = scal_field(c, f)) {
    for (size_t i = 0; i < s->size(); ++i) (*s)[i] = in[i];
    return (int)s->size();
  }
  return 0;
}
void ref_get_type(void* vc, int* out) {
  RefCase* c = static_cast<RefCase*>(vc);
  for (size_t i = 0; i < c->ps.data.part_type.size(); ++i) out[i] = c->ps.data.part_type[i];
}
void ref_get_index(void* vc, int* out) {
  RefCase* c = static_cast<RefCase*>(vc);
  for (size_t i = 0; i < c->ps.data.index.size(); ++i) out[i] = c->ps.data.index[i];
}

// ---- cell list ----
void ref_divide_to_cells(void* vc) {
  RefCase* c = static_cast<RefCase*>(vc);
  Divide_To_Cells(c->ps, c->sd);
}
void ref_clear_cells(void* vc) { Clear_cells(static_cast<RefCase*>(vc)->ps); }
void ref_particle_to_cell(void* vc) {
  RefCase* c = static_cast<RefCase*>(vc);
  Particle_To_Cell(c->ps, c->sd);
}
void ref_grid(void* vc, int* ncx, int* ncy) {
  RefCase* c = static_cast<RefCase*>(vc);
  *ncx = (int)c->ps.pairs.ncx;
  *ncy = (int)c->ps.pairs.ncy;
}
// CSR export of the cell lists: start[ncells+1], items[np]
void ref_cell_csr(void* vc, int* start, int* items) {
  RefCase* c = static_cast<RefCase*>(vc);
  int k = 0;
  for (size_t cell = 0; cell < c->ps.cells.size(); ++cell) {
    start[cell] = k;
    for (int p : c->ps.cells[cell].cp) items[k++] = p;
  }
  start[c->ps.cells.size()] = k;
}

// ---- stages ----
void ref_density_to_pressure(void* vc) { Density_to_pressure(static_cast<RefCase*>(vc)->ps); }
void ref_compute_forces(void* vc) { Compute_Forces(static_cast<RefCase*>(vc)->ps); }
void ref_wall_density(void* vc) {
  RefCase* c = static_cast<RefCase*>(vc);
#if defined(REF_BT)
  BT_compute_density_BTGeo(c->ps, c->sd);
#elif defined(REF_WCSPH)
  mDBC_compute_density_mdbcGeo(c->ps, c->sd);
#else
  (void)c;
#endif
}
#if defined(REF_RALE)
void ref_rale_leapfrog_one(void* vc, double dt) {
  Integrate_LeapFrog_partOne(static_cast<RefCase*>(vc)->ps, dt);
}
void ref_rale_leapfrog_two(void* vc, double dt, int step) {
  Integrate_LeapFrog_partTwo(static_cast<RefCase*>(vc)->ps, dt, step);
}
// ALE-SPH with MUSCL reconstruction (cases/ALESPH_dambreakLong/SPH_simulation.h:78-125)
void ref_ale_compute_density(void* vc) { Compute_Density(static_cast<RefCase*>(vc)->ps); }
void ref_ale_compute_acceleration(void* vc) {
  Compute_Acceleration_BT(static_cast<RefCase*>(vc)->ps);
}
void ref_ale_kernel_gradient(void* vc) { Kernel_gradient_approx(static_cast<RefCase*>(vc)->ps); }
void ref_ale_density_pressure(void* vc, double dt) {
  Integrate_density_compute_pressure(static_cast<RefCase*>(vc)->ps, dt);
}
void ref_ale_leapfrog_one(void* vc, double dt) {
  Integrate_LeapFrog_partOne_ALESPH(static_cast<RefCase*>(vc)->ps, dt);
}
void ref_ale_leapfrog_two(void* vc, double dt, int step) {
  Integrate_LeapFrog_partTwo_ALESPH(static_cast<RefCase*>(vc)->ps, dt, step);
}
#else
void ref_predictor(void* vc, double dt) { Integrate_SymplecticPredictor(static_cast<RefCase*>(vc)->ps, dt); }
void ref_corrector(void* vc, double dt) { Integrate_SymplecticCorrector(static_cast<RefCase*>(vc)->ps, dt); }
void ref_verlet(void* vc, double dt) { Integrate_Verlet(static_cast<RefCase*>(vc)->ps, dt); }
void ref_euler(void* vc, double dt) { Integrate_Euler(static_cast<RefCase*>(vc)->ps, dt); }
void ref_leapfrog_one(void* vc, double dt) {
  Integrate_LeapFrog_partOne_withDensity(static_cast<RefCase*>(vc)->ps, dt);
}
void ref_leapfrog_two(void* vc, double dt, int step) {
  Integrate_LeapFrog_partTwo_withDensity(static_cast<RefCase*>(vc)->ps, dt, step);
}
void ref_reinit_sum(void* vc) { Density_reinitialization_sumation(static_cast<RefCase*>(vc)->ps); }
void ref_reinit_overwrite(void* vc) { Density_reinitialization_overwrite(static_cast<RefCase*>(vc)->ps); }
#endif
void ref_remove_out_of_domain(void* vc) {
  RefCase* c = static_cast<RefCase*>(vc);
  RemoveParticlesOutOfDomain(c->ps, c->sd);
}

// ---- measurements ----
double ref_water_elevation(void* vc, double x, double y0) {
  RefCase* c = static_cast<RefCase*>(vc);
  return Water_Elevation(c->ps, c->sd, x, y0);
}
double ref_water_elevation_inverse(void* vc, double x, double y0) {
  RefCase* c = static_cast<RefCase*>(vc);
  return Water_Elevation_inverse(c->ps, c->sd, x, y0);
}
double ref_sum_kernel_function(void* vc, double x, double y) {
  RefCase* c = static_cast<RefCase*>(vc);
  return Sum_kernel_function(c->ps, c->sd, {x, y});
}
void ref_interpolate_point(void* vc, double x, double y, double* out4) {
  RefCase* c = static_cast<RefCase*>(vc);
  std::array<real, 4> f = Kernel_velocity_approximation_TEST(c->ps, c->sd, {x, y});
  for (int i = 0; i < 4; ++i) out4[i] = f[i];
}

// ---- case geometry (the reference's own creation routines) ----
#if defined(REF_BT) || defined(REF_RALE)
void ref_create_wall_x_bt(void* vc, double x0, double x1, double y, int o, double nx, double ny) {
  Create_Wall_x_bt(static_cast<RefCase*>(vc)->ps, x0, x1, y, o, {nx, ny});
}
void ref_create_wall_y_bt(void* vc, double x, double y0, double y1, int o, double nx, double ny) {
  Create_Wall_y_bt(static_cast<RefCase*>(vc)->ps, x, y0, y1, o, {nx, ny});
}
void ref_create_corner_bt(void* vc, int ox, int oy, double cx, double cy, double nx, double ny) {
  Create_corner_bt(static_cast<RefCase*>(vc)->ps, ox, oy, {cx, cy}, {nx, ny});
}
#if defined(REF_RALE)
void ref_create_block_of_fluid(void* vc, double x0, double y0, double x1, double y1) {
  Create_Block_of_Fluid(static_cast<RefCase*>(vc)->ps, x0, y0, x1, y1);
}
#endif
#else
void ref_create_wall_x_mdbc(void* vc, double x0, double x1, double y, int o, double yw) {
  Create_Wall_x_mdbc(static_cast<RefCase*>(vc)->ps, x0, x1, y, o, yw);
}
void ref_create_wall_y_mdbc(void* vc, double x, double y0, double y1, int o, double xw) {
  Create_Wall_y_mdbc(static_cast<RefCase*>(vc)->ps, x, y0, y1, o, xw);
}
void ref_create_corner_mdbc(void* vc, int ox, int oy, double cx, double cy) {
  Create_corner_mdbc(static_cast<RefCase*>(vc)->ps, ox, oy, {cx, cy});
}
void ref_create_block_of_fluid(void* vc, double x0, double y0, double x1, double y1) {
  Create_Block_of_Fluid(static_cast<RefCase*>(vc)->ps, x0, y0, x1, y1);
}
#endif
#if defined(REF_WCSPH)
// ---- buffer zones (src/SPH_buffer_zones.h, src/SPH_create.h:439-542) ----
void ref_draw_create_inlet_y(void* vc, double x, double y0, double y1, int o, int nbl, double vx,
                             double vy, int type) {
  DRAW_CreateInletY(static_cast<RefCase*>(vc)->ps, x, y0, y1, o, nbl, {vx, vy}, type);
}
void ref_draw_create_inlet_x(void* vc, double y, double x0, double x1, int o, int nbl, double vx,
                             double vy, int type) {
  DRAW_CreateInletX(static_cast<RefCase*>(vc)->ps, y, x0, x1, o, nbl, {vx, vy}, type);
}
void ref_draw_create_outlet_y(void* vc, double x, double y0, double y1, int o, int nbl, double vx,
                              double vy, int type, int empty) {
  DRAW_CreateOutletY(static_cast<RefCase*>(vc)->ps, x, y0, y1, o, nbl, {vx, vy}, type, empty != 0);
}
int ref_zone_count(void* vc) { return (int)static_cast<RefCase*>(vc)->ps.zones.size(); }
// {x0, xm, y0, ym, nbl, bl, bh, btype, vbx, vby, ox, oy}
void ref_zone_get(void* vc, int i, double* out) {
  const BufferZone& z = static_cast<RefCase*>(vc)->ps.zones[i];
  const double v[12] = {z.x0, z.xm, z.y0, z.ym, (double)z.nbl, z.bl, z.bh, (double)z.btype,
                        z.v_b.x, z.v_b.y, z.orientation.x, z.orientation.y};
  for (int k = 0; k < 12; ++k) out[k] = v[k];
}
void ref_bz_integrate_inlet(void* vc, int zone, double dt) {
  BZ_Integrate_inlet(static_cast<RefCase*>(vc)->ps, (unsigned)zone, dt);
}
void ref_bz_integrate_outlet(void* vc, int zone, double dt) {
  BZ_Integrate_outlet(static_cast<RefCase*>(vc)->ps, (unsigned)zone, dt);
}
// write_to_ASCII_VTK, src/SPH_data_to_vtk.h:4-51
void ref_write_vtk(void* vc, const char* path) {
  write_to_ASCII_VTK(static_cast<RefCase*>(vc)->ps, std::string(path));
}
// write_INTERPOLATED_mesh_to_ASCII_VTK, src/SPH_grid_to_vtk.h:139-243
void ref_write_grid_vtk(int ncx, int ncy, double kh, double x0, double y0, const double* v,
                        const double* rho, const double* p, const char* path) {
  const int n = ncx * ncy;
  std::vector<realvec> vv(n);
  for (int t = 0; t < n; ++t) vv[t] = {v[2 * t], v[2 * t + 1]};
  std::vector<real> rr(rho, rho + n), pp(p, p + n);
  write_INTERPOLATED_mesh_to_ASCII_VTK(ncx, ncy, kh, x0, y0, vv, rr, pp, std::string(path));
}
// write_mesh_to_ASCII_VTK, src/SPH_grid_to_vtk.h:6-67 (cell counts of the last
// Particle_To_Cell)
void ref_write_mesh_vtk(void* vc, const char* path) {
  RefCase* c = static_cast<RefCase*>(vc);
  write_mesh_to_ASCII_VTK(c->ps, c->sd, std::string(path));
}
// Output_hydrostatic_pressure, src/SPH_measure_pressure.h:17-34
void ref_output_hydrostatic_pressure(void* vc, const char* path) {
  Output_hydrostatic_pressure(static_cast<RefCase*>(vc)->ps, std::string(path));
}
// Output_file, src/SPH_output_info.h:8-43, with the case's padded domain and
// the given nvl
void ref_output_file(void* vc, int nvl, const char* path) {
  RefCase* c = static_cast<RefCase*>(vc);
  Simulation_data sd = c->sd;
  sd.nvl = nvl;
  Output_file(c->ps.data_const, sd, std::string(path));
}
// Write_water_level / Write_pressure, src/SPH_output_info.h:58-76
void ref_write_water_level(double t, double value, const char* path) {
  Write_water_level(t, value, std::string(path));
}
void ref_write_pressure(double t, double value, const char* path) {
  Write_pressure(t, value, std::string(path));
}
#endif
void ref_create_block_of_fluid_wp(void* vc, double x0, double y0, double x1, double y1) {
  Create_Block_of_FluidWP(static_cast<RefCase*>(vc)->ps, x0, y0, x1, y1);
}

}  // extern "C"
