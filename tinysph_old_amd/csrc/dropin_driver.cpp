// Example case driver written against the reference's stage API through
// include/tinysph_dropin.hpp, linked with libtinysph_hip.so.  Built twice:
//   tsph_dropin_mdbc : the dam break with mDBC walls, stage order of
//                      cases/damBreakLong_mDBC/SPH_simulationSymplectic.h:33-116
//                      and geometry of cases/damBreakLong_mDBC/geometry.h
//   tsph_dropin_rale : the Riemann-ALE dam break (-DTSPH_DROPIN_RALE), stage
//                      order of cases/ALESPH_dambreakLong/SPH_simulationINT.h:84-132
//                      and geometry of cases/ALESPH_dambreakLong/geometry.h
//   tsph_dropin_verlet : the mDBC dam break with the Verlet integrator and the
//                      density reinitialisation (-DTSPH_DROPIN_VERLET), stage order
//                      of cases/damBreakLong_mDBC/SPH_simulationVerlet_reinit.h
// Apart from the state dumps, the bodies of INIT/RUN are the reference drivers'
// call sequences unchanged.
//
//   tsph_dropin_{mdbc,rale,verlet} DP STEPS INIT.bin FINAL.bin
//
// Dump format (little endian): int32 np, int32 n_extra, int32 type[np],
// f64 r[2np], v[2np], rho[np], p[np], extra[2*n_extra] (gnr for mDBC, n for
// BT), then for ALE runs f64 omega[np], omegarho[np], omegav[2np].
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
  std::fwrite(&np, 4, 1, f);
  std::fwrite(&ne, 4, 1, f);
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

struct SPH_simulation {
  double dp, hh, kap = 2.1, cs = 34.3, visco = 0.02, delta = 0.1, dt;
  int nvl;
  int step = 0, step_end;
  Particle_system particles;
  Simulation_data simulation_data;

  SPH_simulation(double _dp, int steps)
      : dp(_dp), hh(std::sqrt(2 * _dp * _dp)), dt(0.2 * std::sqrt(2 * _dp * _dp) / 34.3),
        nvl((int)std::ceil(2.1 * std::sqrt(2 * _dp * _dp) / _dp)), step_end(steps),
        particles(std::sqrt(2 * _dp * _dp), 2.1, 34.3, 1000., 0.02, 0.1, _dp),
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
#else
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
    }
  }
#endif
};

}  // namespace

int main(int argc, char **argv) {
  if (argc != 5) {
    std::fprintf(stderr, "usage: %s DP STEPS INIT.bin FINAL.bin\n", argv[0]);
    return 2;
  }
  try {
    SPH_simulation sim(std::atof(argv[1]), std::atoi(argv[2]));
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
