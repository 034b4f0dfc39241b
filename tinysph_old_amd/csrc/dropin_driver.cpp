// Example case driver written against the reference's stage API through
// include/tinysph_dropin.hpp, linked with libtinysph_hip.so.  Built four times:
//   tsph_dropin_mdbc   : the dam break with mDBC walls, stage order of
//                        cases/damBreakLong_mDBC/SPH_simulationSymplectic.h:33-203
//                        (with its water-level gauges and wall sensors) and
//                        geometry of cases/damBreakLong_mDBC/geometry.h
//   tsph_dropin_rale   : the Riemann-ALE dam break (-DTSPH_DROPIN_RALE), stage
//                        order of cases/ALESPH_dambreakLong/SPH_simulationINT.h:84-132
//                        and geometry of cases/ALESPH_dambreakLong/geometry.h
//   tsph_dropin_verlet : the mDBC dam break with the Verlet integrator and the
//                        density reinitialisation (-DTSPH_DROPIN_VERLET), stage order
//                        of cases/damBreakLong_mDBC/SPH_simulationVerlet_reinit.h
//   tsph_dropin_inflow : the inflow slab (-DTSPH_DROPIN_INFLOW), stage order of
//                        cases/inflowSlab_mDBC/SPH_simulationSymplectic.h:57-123,
//                        constants and geometry of cases/inflowSlab_mDBC/; a
//                        small block of fluid is added past the right domain
//                        bound, so that the run both appends (inlet) and removes
//                        (RemoveParticlesOutOfDomain)
// Apart from the state dumps, the bodies of INIT/RUN are the reference drivers'
// call sequences unchanged.
//
//   tsph_dropin_{mdbc,rale,verlet,inflow} DP STEPS INIT.bin FINAL.bin [OUTDIR]
//
// With OUTDIR, INIT writes OUTDIR/grid.vtk and OUTDIR/info.txt
// (write_mesh_to_ASCII_VTK, Output_file) and the mDBC dam break writes its
// gauge and sensor series (Write_water_level) there as its reference driver does.
//
// Dump format (little endian): int32 np, int32 n_extra, int32 np_out,
// int32 type[np], f64 r[2np], v[2np], rho[np], p[np], extra[2*n_extra] (gnr for
// mDBC, n for BT), then for ALE runs f64 omega[np], omegarho[np], omegav[2np].
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "tinysph_dropin.hpp"

namespace {

void dump(Particle_system &ps, const char *path) {
  tsph_sync_host(ps);
  FILE *f = std::fopen(path, "wb");
  if (!f) throw tsph_error(std::string("cannot write ") + path);
  const int np = (int)ps.np;
#ifdef TSPH_DROPIN_RALE
  const std::vector<realvec> &extra = ps.special.n;
#else
  const std::vector<realvec> &extra = ps.special.gnr;
#endif
  const int ne = (int)extra.size();
  const int np_out = (int)ps.np_out;
  std::fwrite(&np, 4, 1, f);
  std::fwrite(&ne, 4, 1, f);
  std::fwrite(&np_out, 4, 1, f);
  for (int i = 0; i < np; ++i) {
    const int32_t t = ps.data.part_type[i];
    std::fwrite(&t, 4, 1, f);
  }
  auto vec = [&](const std::vector<realvec> &v, int n) {
    for (int i = 0; i < n; ++i) {
      std::fwrite(&v[i].x, 8, 1, f);
      std::fwrite(&v[i].y, 8, 1, f);
    }
  };
  vec(ps.data.r, np);
  vec(ps.data.v, np);
  std::fwrite(ps.data.rho.data(), 8, np, f);
  std::fwrite(ps.data.p.data(), 8, np, f);
  vec(extra, ne);
#ifdef TSPH_DROPIN_RALE
  std::fwrite(ps.special.omega.data(), 8, np, f);
  std::fwrite(ps.special.omegarho.data(), 8, np, f);
  vec(ps.special.omegav, np);
#endif
  std::fclose(f);
}

// the case constants (constants.h of each reference case; dt scaled with dp)
struct Constants {
  double cs, visco, delta, dt_at_5mm;
};
#ifdef TSPH_DROPIN_INFLOW
constexpr Constants kCase = {42.48, 0.1, 0.1, 0.00005};  // cases/inflowSlab_mDBC/constants.h
#else
constexpr Constants kCase = {34.3, 0.02, 0.1, 0.0};  // cases/damBreakLong_mDBC/constants.h
#endif

struct SPH_simulation {
  double dp, hh, kap = 2.1, cs = kCase.cs, visco = kCase.visco, delta = kCase.delta, dt;
  int nvl;
  int step = 0, step_end;
  std::string out;  // output folder ("": no text output)
  Particle_system particles;
  Simulation_data simulation_data;

  SPH_simulation(double _dp, int steps, std::string outdir)
      : dp(_dp), hh(std::sqrt(2 * _dp * _dp)),
        dt(kCase.dt_at_5mm > 0 ? kCase.dt_at_5mm * _dp / 0.005 : 0.2 * std::sqrt(2 * _dp * _dp) / kCase.cs),
        nvl((int)std::ceil(2.1 * std::sqrt(2 * _dp * _dp) / _dp)), step_end(steps), out(outdir),
        particles(std::sqrt(2 * _dp * _dp), 2.1, kCase.cs, 1000., kCase.visco, kCase.delta, _dp),
        simulation_data(0., 2., 0., 1.) {
    // PREP_SIMULATION_DATA
    simulation_data.hh = hh;
    simulation_data.kh = hh * kap;
    simulation_data.nvl = nvl;
    simulation_data.x_0 -= nvl * dp;
    simulation_data.x_m += nvl * dp;
    simulation_data.y_0 -= nvl * dp;
    simulation_data.y_m += nvl * dp;
  }

  void GEOMETRY() {
#ifdef TSPH_DROPIN_RALE
    Create_Wall_x_bt(particles, 0. + dp, 1.61 - dp, 0., -1, {0., 1.});
    Create_Wall_y_bt(particles, 0., 0. + dp, 1., -1, {1., 0.});
    Create_Wall_y_bt(particles, 1.61, 0. + dp, 1., +1, {-1., 0.});
    Create_corner_bt(particles, -1, -1, {0., 0.}, {std::sqrt(2), std::sqrt(2)});
    Create_corner_bt(particles, +1, -1, {1.61, 0.}, {-std::sqrt(2), std::sqrt(2)});
    Create_Block_of_Fluid(particles, 0.0 + dp, 0 + dp, 0.6, 0.3);
#elif defined(TSPH_DROPIN_INFLOW)
    Create_Wall_x_mdbc(particles, 0.0, 0.4 - 3 * dp, 0.1, -1, 0.1 + dp / 2);
    Create_Wall_y_mdbc(particles, 0.4, 0 + dp, 0.1 - 3 * dp, -1, 0.4 + dp / 2);
    Create_Wall_x_mdbc(particles, 0.4 + dp, 1.2, 0., -1, 0 + dp / 2);
    Create_corner_mdbc(particles, -1, -1, {0.4 + dp / 2, 0. + dp / 2});
    Create_corner_mdbc(particles, -1, -1, {0.4 + dp / 2, 0.1 + dp / 2});
    DRAW_CreateInletY(particles, 0.1, 0.1 + dp, 0.2, -1, 4, {0.75, 0.}, inlet1);
    DRAW_CreateOutletY(particles, 1.0, 0.0 + dp, 0.5, +1, 4, {0.5, 0.}, outlet1, true);
    // one column of fluid past the removal bound x_m - 0.01 (src/SPH_btdebug.h:22)
    // and inside the cell grid
    Create_Block_of_Fluid(particles, 2.0 + 2.75 * dp, 0.5, 2.0 + 2.75 * dp, 0.5 + 4 * dp);
#else
    Create_Wall_x_mdbc(particles, 0. + dp, 1.61, 0., -1, dp / 2);
    Create_Wall_y_mdbc(particles, 0., 0 + dp, 1., -1, dp / 2);
    Create_Wall_y_mdbc(particles, 1.61, 0 + dp, 1., +1, 1.61 - dp / 2);
    Create_corner_mdbc(particles, -1, -1, {0. + dp / 2, 0. + dp / 2});
    Create_corner_mdbc(particles, +1, -1, {1.61 - dp / 2, 0. + dp / 2});
    Create_Block_of_Fluid(particles, 0 + dp, 0 + dp, 0.6, 0.3);
#endif
  }

  void INIT() {
    Divide_To_Cells(particles, simulation_data);
    Particle_To_Cell(particles, simulation_data);
    if (!out.empty()) {
      write_mesh_to_ASCII_VTK(particles, simulation_data, out + "/grid.vtk");
      Output_file(particles.data_const, simulation_data, out + "/info.txt");
    }
  }

#ifdef TSPH_DROPIN_RALE
  void RUN() {
    while (step < step_end) {
      step++;
      if (step > 1) Integrate_LeapFrog_partOne(particles, dt);
      Clear_cells(particles);
      Particle_To_Cell(particles, simulation_data);
      Density_to_pressure(particles);
      Compute_Forces(particles);
      Integrate_LeapFrog_partTwo(particles, dt, step);
      RemoveParticlesOutOfDomain(particles, simulation_data);
    }
  }
#elif defined(TSPH_DROPIN_VERLET)
  void RUN() {
    while (step < step_end) {
      step++;
      mDBC_compute_density_mdbcGeo(particles, simulation_data);
      Clear_cells(particles);
      Particle_To_Cell(particles, simulation_data);
      Density_to_pressure(particles);
      Compute_Forces(particles);
      if (step > 1 && step % 20 != 0) {
        Integrate_Verlet(particles, dt);
      } else {
        Integrate_Euler(particles, dt);
      }
      RemoveParticlesOutOfDomain(particles, simulation_data);
      if (step % 20 == 0) {
        Density_reinitialization_sumation(particles);
        Density_reinitialization_overwrite(particles);
      }
    }
  }
#elif defined(TSPH_DROPIN_INFLOW)
  void RUN() {
    while (step < step_end) {
      step++;
      mDBC_compute_density_mdbcGeo(particles, simulation_data);
      Density_to_pressure(particles);
      Integrate_SymplecticPredictor(particles, dt);
      BZ_Integrate_inlet(particles, 0, dt);
      BZ_Integrate_outlet(particles, 1, dt);
      Clear_cells(particles);
      Particle_To_Cell(particles, simulation_data);
      Density_to_pressure(particles);
      Compute_Forces(particles);
      Integrate_SymplecticCorrector(particles, dt);
      RemoveParticlesOutOfDomain(particles, simulation_data);
      Clear_cells(particles);
      Particle_To_Cell(particles, simulation_data);
      Density_to_pressure(particles);
      Compute_Forces(particles);
    }
  }
#else
  // cases/damBreakLong_mDBC/constants.h:58-94: gauges and sensors
  const unsigned save_WL_interval = 50, save_pressure_interval = 10;
  realvec wl_h1_pos = {0.3, 0.}, wl_h1_pos_inv = {0.3, 0.45};
  realvec wl_h2_pos = {0.865, 0.}, wl_h2_pos_inv = {0.865, 0.45};
  realvec wl_h3_pos = {1.114, 0.}, wl_h3_pos_inv = {1.114, 0.45};
  realvec wl_h4_pos = {1.3625, 0.}, wl_h4_pos_inv = {1.3625, 0.45};
  realvec p_h1_pos = {1.61, 0.003}, p_h2_pos = {1.61, 0.015}, p_h3_pos = {1.61, 0.03},
          p_h4_pos = {1.61, 0.08};

  void RUN() {
    while (step < step_end) {
      step++;
      mDBC_compute_density_mdbcGeo(particles, simulation_data);
      Density_to_pressure(particles);
      Integrate_SymplecticPredictor(particles, dt);
      Clear_cells(particles);
      Particle_To_Cell(particles, simulation_data);
      Density_to_pressure(particles);
      Compute_Forces(particles);
      Integrate_SymplecticCorrector(particles, dt);
      RemoveParticlesOutOfDomain(particles, simulation_data);
      Clear_cells(particles);
      Particle_To_Cell(particles, simulation_data);
      Density_to_pressure(particles);
      Compute_Forces(particles);
      if (out.empty()) continue;
      // cases/damBreakLong_mDBC/SPH_simulationSymplectic.h:156-201
      if (step % save_WL_interval == 0) {
        real timeCoef = pow(9.81 / 0.3, 0.5);
        real wl_h1 = Water_Elevation(particles, simulation_data, wl_h1_pos.x, wl_h1_pos.y);
        Write_water_level(step * dt * timeCoef, wl_h1 / 0.3, out + "/waterLevel_h1.dat");
        real wl_h1_inv = Water_Elevation_inverse(particles, simulation_data, wl_h1_pos_inv.x, wl_h1_pos_inv.y);
        Write_water_level(step * dt * timeCoef, wl_h1_inv / 0.3, out + "/waterLevel_h1_inv.dat");
        real wl_h2 = Water_Elevation(particles, simulation_data, wl_h2_pos.x, wl_h2_pos.y);
        Write_water_level(step * dt * timeCoef, wl_h2 / 0.3, out + "/waterLevel_h2.dat");
        real wl_h2_inv = Water_Elevation_inverse(particles, simulation_data, wl_h2_pos_inv.x, wl_h2_pos_inv.y);
        Write_water_level(step * dt * timeCoef, wl_h2_inv / 0.3, out + "/waterLevel_h2_inv.dat");
        real wl_h3 = Water_Elevation(particles, simulation_data, wl_h3_pos.x, wl_h3_pos.y);
        Write_water_level(step * dt * timeCoef, wl_h3 / 0.3, out + "/waterLevel_h3.dat");
        real wl_h3_inv = Water_Elevation_inverse(particles, simulation_data, wl_h3_pos_inv.x, wl_h3_pos_inv.y);
        Write_water_level(step * dt * timeCoef, wl_h3_inv / 0.3, out + "/waterLevel_h3_inv.dat");
        real wl_h4 = Water_Elevation(particles, simulation_data, wl_h4_pos.x, wl_h4_pos.y);
        Write_water_level(step * dt * timeCoef, wl_h4 / 0.3, out + "/waterLevel_h4.dat");
        real wl_h4_inv = Water_Elevation_inverse(particles, simulation_data, wl_h4_pos.x, wl_h4_pos_inv.y);
        Write_water_level(step * dt * timeCoef, wl_h4_inv / 0.3, out + "/waterLevel_h4_inv.dat");
      }
      if (step % save_pressure_interval == 0) {
        real timeCoef = pow(9.81 / 0.3, 0.5);
        real pCoef = 1. / (1000. * 9.81 * 0.3);
        real p_h1 = Pressure_in_point(particles, simulation_data, p_h1_pos);
        Write_water_level(step * dt * timeCoef, p_h1 * pCoef, out + "/pressure_p1.dat");
        real p_h2 = Pressure_in_point(particles, simulation_data, p_h2_pos);
        Write_water_level(step * dt * timeCoef, p_h2 * pCoef, out + "/pressure_p2.dat");
        real p_h3 = Pressure_in_point(particles, simulation_data, p_h3_pos);
        Write_water_level(step * dt * timeCoef, p_h3 * pCoef, out + "/pressure_p3.dat");
        real p_h4 = Pressure_in_point(particles, simulation_data, p_h4_pos);
        Write_water_level(step * dt * timeCoef, p_h4 * pCoef, out + "/pressure_p4.dat");
      }
    }
  }
#endif
};

}  // namespace

int main(int argc, char **argv) {
  if (argc != 5 && argc != 6) {
    std::fprintf(stderr, "usage: %s DP STEPS INIT.bin FINAL.bin [OUTDIR]\n", argv[0]);
    return 2;
  }
  try {
    SPH_simulation sim(std::atof(argv[1]), std::atoi(argv[2]), argc == 6 ? argv[5] : "");
    sim.GEOMETRY();
    dump(sim.particles, argv[3]);
    sim.INIT();
    sim.RUN();
    dump(sim.particles, argv[4]);
  } catch (const std::exception &e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  return 0;
}
