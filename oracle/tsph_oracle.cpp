// tsph_oracle.cpp -- the ORACLE: a CPU restatement of the reference tinySPH
// per-step particle pipeline, used only as the checker by tests/, by
// __graft_entry__.smoke() and by bench.py's CPU-baseline leg.  It is test
// infrastructure, never part of the shipped product.
//
// Every stage follows the reference operation by operation (same operand
// order, same float-rounding points, same neighbour-cell order, same
// ascending in-cell order), cited per function.  Two deliberate differences,
// neither of which changes a single bit of the results:
//   * kernel values are returned by value instead of through a static buffer
//     (src/SPH_kernel.h:54), so the pair loops may run under OpenMP;
//   * cell lists are stored as one CSR array instead of a std::vector per cell
//     (src/SPH_data.h:95-108); the in-cell order is still ascending particle
//     index (src/SPH_find_pairs.h:61-71).
// Pinning: tests/test_oracle_ref.py compares every stage here bit-for-bit with
// the reference's own functions compiled in place (oracle/_ref); the
// closed-form and invariant tests in tests/test_oracle_math.py cover the rest.
//
// Build: oracle/Makefile (-O3 -funroll-loops -ffp-contract=off; OpenMP variant
// liboracle_omp.so for the multi-core baseline).

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

struct V2 { double x, y; };
inline V2 operator+(V2 a, V2 b) { return {a.x + b.x, a.y + b.y}; }
inline V2 operator-(V2 a, V2 b) { return {a.x - b.x, a.y - b.y}; }
inline V2 operator*(V2 a, double b) { return {a.x * b, a.y * b}; }
inline V2 operator/(V2 a, double b) { return {a.x / b, a.y / b}; }

// Particle type ids, src/SPH_defs.h:44-65
enum { T_FLUID = 0, T_WALL = 1, T_INLET = 6, T_OUTLET = 11, T_OUTLETF = 17 };
const double EPS = 0.01;  // src/SPH_defs.h:69

struct Params {  // SPH_constants + Simulation_data, src/SPH_data.h:30-46,111-125
  double h, kap, cs, rho0, avisc, delta, dp;
  double x0, xm, y0, ym;
  double m, gx, gy;
};

// BufferZone, src/SPH_inlet_outlet_data.h:18-46 (nbl is never stored there)
struct Zone {
  V2 orientation;
  double x0, y0, xm, ym, bl, bh;
  int btype;
  V2 v_b;
};

struct Case {
  Params q;
  std::vector<Zone> zones;
  int np = 0, np_out = 0;
  std::vector<int> type, index;
  std::vector<double> p, rho, drho, rho_o;
  std::vector<V2> r, v, a, r_o, v_o, gnr, n;
  // ALE-SPH volume variables (SPH_special, src/SPH_data.h:54-67)
  std::vector<double> omega, omegarho, omega_o, omegarho_o, domega, domegarho;
  std::vector<V2> omegav, omegav_o, domegav;
  // ALE-SPH MUSCL gradients (src/SPH_data.h:68-70)
  std::vector<V2> gradrho, gradvx, gradvy;
  // second time layer of the Verlet integrator (src/SPH_data.h:23,25) and the
  // density-reinitialisation buffer (SPH_special::rhoTemp, src/SPH_data.h:82)
  std::vector<V2> v_oo;
  std::vector<double> rho_oo, rho_tmp;
  int ncx = 0, ncy = 0;
  std::vector<int> cstart, citems;  // CSR cell lists
  int nthreads = 1;
};

// src/SPH_kernel.h:52-77 (Wendland C2, 2-D) with the reference's casts:
// awen/bwen rounded to float, wqq1/wqq rounded to float, wqq2 a float product.
struct KW { double w, f; };
inline KW wendland(double r, double h) {
  const double awen = (double)(float)(0.557 / (h * h));
  const double bwen = (double)(float)(-2.7852 / (h * h * h));
  const double qq = r / h;
  const float wqq1 = (float)(1.0 - 0.5 * qq);
  const float wqq2 = wqq1 * wqq1;
  const float wqq = (float)(qq + qq + 1.0);
  KW k{0.0, 0.0};
  if (qq <= 2.0) {
    k.w = awen * (double)wqq * (double)wqq2 * (double)wqq2;
    k.f = bwen * qq * (double)wqq2 * (double)wqq1 / r;
  }
  if (r == 0.0) { k.w = awen; k.f = 0.0; }
  return k;
}

// src/SPH_viscouse_forces.h:7-25 (called with rho0 := 0.5*(rho_i+rho_j))
inline double artificial_viscosity(double h, double drs, double drdv, double rhobar,
                                   double c0, double alpha) {
  double av = 0.0;
  const double visco = drdv > alpha ? drdv : alpha;
  if (drdv < 0) {
    const double mu = h * drdv / (drs * drs + EPS * h * h);
    av = -visco * c0 * mu / rhobar;
  }
  return av;
}

// src/SPH_viscouse_forces.h:45-61
inline double diffusion_fourtakas(double h, double drs, double dr_z, double drdW,
                                  double arho, double nrho, double rho0, double c0,
                                  double delta, double m, double gravity) {
  const double cb = c0 * c0 * rho0 / 7.;
  const double dpH = 1. + dr_z * rho0 * gravity / cb;
  const double drhoH = rho0 * std::pow(dpH, 1. / 7) - rho0;
  const double psi = 2. * c0 * h * delta * ((nrho - arho) - drhoH) / (drs * drs + EPS * h * h);
  return psi * drdW * m / nrho;
}

// src/SPH_pressure.h:7-19
inline double tait_pressure(double rho, double rho0, double c0) {
  const double b = c0 * c0 * rho0 / 7.;
  const double rr = rho / rho0;
  return b * (std::pow(rr, 7.) - 1);
}

// src/SPH_pressure.h:49-63
inline double pressure_to_density(double p, double rho0, double c0) {
  const double b = c0 * c0 * rho0 / 7.;
  return std::pow(p / b + 1, 1. / 7) * rho0;
}

inline double kh_of(const Params& q) { return q.kap * q.h; }
inline double of_of(const Params& q) { return q.h * 0.273; }

// cell of a point, src/SPH_find_pairs.h:64-65 (truncation toward zero)
inline void cell_xy(const Params& q, V2 r, int& cx, int& cy) {
  const double kh = kh_of(q), of = of_of(q);
  cx = (int)((r.x - q.x0 + of) / kh);
  cy = (int)((r.y - q.y0 + of) / kh);
}

// 9-cell stencil in the reference order, src/SPH_interaction.h:26-36
inline void stencil(int c, int ncy, int ac[9]) {
  ac[0] = c; ac[1] = c + 1; ac[2] = c + 1 + ncy; ac[3] = c + ncy; ac[4] = c - 1 + ncy;
  ac[5] = c - 1; ac[6] = c - 1 - ncy; ac[7] = c - ncy; ac[8] = c + 1 - ncy;
}

inline int ncells(const Case& s) { return s.ncx * s.ncy; }

// ---------------------------------------------------------------------------
// Pair loops.  Accumulation order per active particle: the 9 cells in stencil
// order, then ascending particle index inside each cell; candidates with
// cl < 0 are skipped and no upper bound is checked (src/SPH_interaction.h:71-80).
// ---------------------------------------------------------------------------

// variant 0: fluid-only WCSPH, src/SPH_interaction.h:8-149
// variant 1: DBC, src/SPH_interactionDBC.h:8-155
// variant 2: BT boundary integral, src/SPH_interactionBT.h:8-192
void compute_forces(Case& s, int variant) {
  const Params& q = s.q;
  const int nc = ncells(s);
  const double h = q.h, m = q.m, c0 = q.cs, delta = q.delta, rho0 = q.rho0, dp = q.dp;
  const V2 g{q.gx, q.gy};
#pragma omp parallel for schedule(dynamic, 64) num_threads(s.nthreads)
  for (int c = 0; c < nc; ++c) {
    int ac[9];
    stencil(c, s.ncy, ac);
    for (int ii = s.cstart[c]; ii < s.cstart[c + 1]; ++ii) {
      const int ai = s.citems[ii];
      const int at = s.type[ai];
      if (variant == 1) {
        if (at == T_INLET || at == T_OUTLET) continue;
      } else if (at != T_FLUID) {
        continue;
      }
      const V2 ar = s.r[ai], av = s.v[ai];
      const double arho = s.rho[ai], ap = s.p[ai];
      V2 acc{0., 0.};
      double drho_sum = 0, diff_sum = 0, gamma = 0;
      for (int k = 0; k < 9; ++k) {
        const int cl = ac[k];
        if (cl < 0) continue;
        for (int jj = s.cstart[cl]; jj < s.cstart[cl + 1]; ++jj) {
          const int ni = s.citems[jj];
          if (ai == ni) continue;
          const int nt = s.type[ni];
          if (variant == 1 && at == T_WALL && nt == T_WALL) continue;
          const V2 nr = s.r[ni], nv = s.v[ni];
          const double nrho = s.rho[ni], np = s.p[ni];
          const V2 dr = ar - nr, dv = av - nv;
          const double drs = std::sqrt((dr.x * dr.x) + (dr.y * dr.y));
          const double drdv = dr.x * dv.x + dr.y * dv.y;
          const KW k2 = wendland(drs, h);
          const V2 dW = dr * k2.f;
          const double dvdW = dv.x * dW.x + dv.y * dW.y;
          const double drdW = dr.x * dW.x + dr.y * dW.y;
          const double p_temp = (ap + np) / (arho * nrho);
          const double visco = artificial_viscosity(h, drs, drdv, 0.5 * (arho + nrho), c0, q.avisc);
          const double diff = diffusion_fourtakas(h, drs, dr.y, drdW, arho, nrho, rho0, c0, delta, m, g.y);
          if (variant == 2) {
            // src/SPH_interactionBT.h:120-159
            if (nt == T_WALL) {
              const V2 nb = s.n[ni];
              acc.x += k2.w * (p_temp + visco) * dp * nb.x * nrho;
              acc.y += k2.w * (p_temp + visco) * dp * nb.y * nrho;
              const double dvnb = dv.x * nb.x + dv.y * nb.y;
              drho_sum -= dvnb * k2.w * nrho * dp;
            } else {
              acc.x -= dW.x * (p_temp + visco) * m;
              acc.y -= dW.y * (p_temp + visco) * m;
              drho_sum += dvdW * m;
              diff_sum -= diff;
            }
            if (nt == T_FLUID) gamma += k2.w * m / nrho;
          } else {
            acc.x -= dW.x * (p_temp + visco) * m;
            acc.y -= dW.y * (p_temp + visco) * m;
            drho_sum += dvdW * m;
            diff_sum -= diff;
          }
        }
      }
      if (variant == 2) {
        // src/SPH_interactionBT.h:166-178
        if (gamma > EPS) {
          s.a[ai] = acc / gamma - g;
          s.drho[ai] = drho_sum / gamma + diff_sum / gamma;
        } else {
          s.a[ai] = V2{0., 0.} - g;
          s.drho[ai] = 0;
        }
      } else if (variant == 1 && at == T_WALL) {
        s.a[ai] = V2{0., 0.};  // src/SPH_interactionDBC.h:137-139
        s.drho[ai] = drho_sum + diff_sum;
      } else {
        s.a[ai] = acc - g;
        s.drho[ai] = drho_sum + diff_sum;
      }
    }
  }
}

// determinant, src/SPH_kernel_approx.h:663-675
inline double det3(double a11, double a12, double a13, double a21, double a22, double a23,
                   double a31, double a32, double a33) {
  return (a11 * a22 * a33 + a12 * a23 * a31 + a13 * a21 * a32 - a31 * a22 * a13 -
          a32 * a23 * a11 - a33 * a21 * a12);
}

// src/SPH_kernel_approx.h:1141-1342 (fluid neighbours of the ghost node `ar`)
double mdbc_density_at(const Case& s, V2 ar, V2 gnd) {
  const Params& q = s.q;
  const double h = q.h, m = q.m;
  double a11 = 0, a12 = 0, a13 = 0, a21 = 0, a22 = 0, a23 = 0, a31 = 0, a32 = 0, a33 = 0;
  double drhodx = 0, drhody = 0, wsum = 0, rhowsum = 0;
  int cx, cy;
  cell_xy(q, ar, cx, cy);
  int ac[9];
  stencil(cy + cx * s.ncy, s.ncy, ac);
  for (int k = 0; k < 9; ++k) {
    const int cl = ac[k];
    if (cl < 0) continue;
    for (int jj = s.cstart[cl]; jj < s.cstart[cl + 1]; ++jj) {
      const int ni = s.citems[jj];
      if (s.type[ni] != T_FLUID) continue;
      const V2 nr = s.r[ni];
      const double nrho = s.rho[ni];
      const V2 dr = ar - nr;
      const double drs = std::sqrt(dr.x * dr.x + dr.y * dr.y);
      const KW kw = wendland(drs, h);
      wsum += kw.w;
      rhowsum += m * kw.w;
      drhodx += dr.x * kw.f * m;
      drhody += dr.y * kw.f * m;
      a11 += kw.w * (m / nrho);
      a12 += dr.x * kw.w * (m / nrho);
      a13 += dr.y * kw.w * (m / nrho);
      a21 += dr.x * kw.f * (m / nrho);
      a22 += dr.x * dr.x * kw.f * (m / nrho);
      a23 += dr.x * dr.y * kw.f * (m / nrho);
      a31 += dr.y * kw.f * (m / nrho);
      a32 += dr.y * dr.x * kw.f * (m / nrho);
      a33 += dr.y * dr.y * kw.f * (m / nrho);
    }
  }
  (void)wsum;
  const double det = det3(a11, a12, a13, a21, a22, a23, a31, a32, a33);
  double dens;
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
    dens = (dtemp + grx * gnd.x + gry * gnd.y);
  } else if (a11 > 0) {
    dens = rhowsum / a11;
  } else {
    dens = 1000.;
  }
  return dens;
}

// src/SPH_mDBC.h:9-50
void mdbc_density(Case& s) {
#pragma omp parallel for schedule(static) num_threads(s.nthreads)
  for (int p = 0; p < s.np; ++p) {
    if (s.type[p] != T_WALL) continue;
    const V2 gn = s.gnr[p];
    const V2 gnd = s.r[p] - gn;
    double rho = mdbc_density_at(s, gn, gnd);
    if (rho < 1000.) rho = 1000.;
    s.rho[p] = rho;
  }
}

// src/SPH_kernel_approx.h:1838-1953
double bt_density_at(const Case& s, V2 ar) {
  const Params& q = s.q;
  const double h = q.h, m = q.m;
  double wsum = 0, rhowsum = 0;
  int cx, cy;
  cell_xy(q, ar, cx, cy);
  int ac[9];
  stencil(cy + cx * s.ncy, s.ncy, ac);
  for (int k = 0; k < 9; ++k) {
    const int cl = ac[k];
    if (cl < 0) continue;
    for (int jj = s.cstart[cl]; jj < s.cstart[cl + 1]; ++jj) {
      const int ni = s.citems[jj];
      if (s.type[ni] != T_FLUID) continue;
      const V2 dr = ar - s.r[ni];
      const double drs = std::sqrt(dr.x * dr.x + dr.y * dr.y);
      const KW kw = wendland(drs, h);
      wsum += kw.w * m / s.rho[ni];
      rhowsum += m * kw.w;
    }
  }
  return wsum > 0. ? rhowsum / wsum : 1000.;
}

// src/SPH_BT.h:9-50
void bt_density(Case& s) {
#pragma omp parallel for schedule(static) num_threads(s.nthreads)
  for (int p = 0; p < s.np; ++p) {
    if (s.type[p] != T_WALL) continue;
    double rho = bt_density_at(s, s.r[p]);
    if (rho < 1000.) rho = 1000.;
    s.rho[p] = rho;
  }
}

// Shepard sum of m W / rho over all types, src/SPH_water_level.h:6-82
double sum_kernel_function(const Case& s, V2 ar) {
  const Params& q = s.q;
  const double kh = kh_of(q), of = of_of(q);
  // the reference stores the cell coordinates in unsigned ints (:17,23-24)
  const unsigned npx = (unsigned)(int)((ar.x - q.x0 + of) / kh);
  const unsigned npy = (unsigned)(int)((ar.y - q.y0 + of) / kh);
  const int c = (int)(npy + npx * (unsigned)s.ncy);
  int ac[9];
  stencil(c, s.ncy, ac);
  double msum = 0;
  for (int k = 0; k < 9; ++k) {
    const int cl = ac[k];
    if (cl < 0) continue;
    for (int jj = s.cstart[cl]; jj < s.cstart[cl + 1]; ++jj) {
      const int ni = s.citems[jj];
      const V2 dr = ar - s.r[ni];
      const double drs = std::sqrt(dr.x * dr.x + dr.y * dr.y);  // pow(x,2) is exact
      const KW kw = wendland(drs, q.h);
      msum += q.m * kw.w / s.rho[ni];
    }
  }
  return msum;
}

// Shepard interpolation of rho, p, v over all types, src/testing.h:6-148
void interpolate_point(const Case& s, V2 ar, double out[4]) {
  const Params& q = s.q;
  const double kh = kh_of(q), of = of_of(q);
  const unsigned npx = (unsigned)(int)((ar.x - q.x0 + of) / kh);
  const unsigned npy = (unsigned)(int)((ar.y - q.y0 + of) / kh);
  const int c = (int)(npy + npx * (unsigned)s.ncy);
  int ac[9];
  stencil(c, s.ncy, ac);
  V2 vw{0., 0.};
  double wsum = 0, rhow = 0, pw = 0, mw = 0;
  for (int k = 0; k < 9; ++k) {
    const int cl = ac[k];
    if (cl < 0) continue;
    for (int jj = s.cstart[cl]; jj < s.cstart[cl + 1]; ++jj) {
      const int ni = s.citems[jj];
      const V2 dr = ar - s.r[ni];
      const double drs = std::sqrt(dr.x * dr.x + dr.y * dr.y);
      const KW kw = wendland(drs, q.h);
      wsum += kw.w;
      vw = vw + s.v[ni] * kw.w;
      rhow = rhow + s.rho[ni] * kw.w;
      pw = pw + s.p[ni] * kw.w;
      mw += q.m * kw.w / s.rho[ni];
    }
  }
  V2 vel{0., 0.};
  double rho = 0, p = 0;
  if (mw > 0.5) {
    vel = V2{0., 0.} + vw / wsum;
    rho = 0 + rhow / wsum;
    p = 0 + pw / wsum;
  }
  if (wsum == 0) vel = V2{0., 0.};
  out[0] = rho; out[1] = p; out[2] = vel.x; out[3] = vel.y;
}

// ---------------------------------------------------------------------------
// Riemann-ALE-SPH, first order (the ALESPH_dambreakLong driver run.sh selects:
// cases/ALESPH_dambreakLong/SPH_simulationINT.h)
// ---------------------------------------------------------------------------

// velRiemannLinearized, src/ALESPH_riemann_solver.h:86-96
inline double vel_riemann(double rhoL, double rhoR, double vL, double vR, double rhoAvg, double c) {
  return 0.5 * (vR + vL) - 0.5 * (rhoR - rhoL) * c / rhoAvg;
}
// densRiemannLinearizedWithLimiter, src/ALESPH_riemann_solver.h:20-34
inline double dens_riemann_lim(double rhoL, double rhoR, double vL, double vR, double rhoAvg, double c) {
  const double d = vL - vR;
  const double mx = d > 0 ? d : 0;
  const double e = 3. * mx;
  const double beta = e < c ? e : c;
  return 0.5 * (rhoL + rhoR) + 0.5 * beta * (vL - vR) * rhoAvg / (c * c);
}

// RALESPH Compute_Forces, src/RALESPH_interaction.h:12-287.  Quirk kept: the
// normalisation sum gamma_sum is declared per cell and never reset between the
// cell's particles (:48, :257, :277-280), so a fluid particle without any
// neighbour in its support takes the "gamma == 0" branch only if no earlier
// fluid particle of its cell had one.
void rale_forces(Case& s) {
  const Params& q = s.q;
  const int nc = ncells(s);
  const double h = q.h, m = q.m, c0 = q.cs, rho0 = q.rho0, dp = q.dp;
  const V2 g{q.gx, q.gy};
#pragma omp parallel for schedule(dynamic, 64) num_threads(s.nthreads)
  for (int c = 0; c < nc; ++c) {
    int ac[9];
    stencil(c, s.ncy, ac);
    double gamma_sum = 0;
    for (int ii = s.cstart[c]; ii < s.cstart[c + 1]; ++ii) {
      const int ai = s.citems[ii];
      if (s.type[ai] != T_FLUID) continue;
      const V2 ar = s.r[ai], av = s.v[ai];
      const double arho = s.rho[ai];
      const double ao = s.omega[ai];
      double domega_sum = 0, domegarho_sum = 0;
      V2 domegav_sum{0., 0.};
      for (int k = 0; k < 9; ++k) {
        const int cl = ac[k];
        if (cl < 0) continue;
        for (int jj = s.cstart[cl]; jj < s.cstart[cl + 1]; ++jj) {
          const int ni = s.citems[jj];
          if (ai == ni) continue;
          const V2 nr = s.r[ni], nv = s.v[ni];
          const double nrho = s.rho[ni];
          const double no = s.omega[ni];
          const V2 dr = ar - nr, dv = av - nv;
          const double drs = std::sqrt((dr.x * dr.x) + (dr.y * dr.y));
          const KW kw = wendland(drs, h);
          const double W = kw.w;
          const V2 dW = dr * kw.f;
          const double dvdW = dv.x * dW.x + dv.y * dW.y;
          const V2 drn = dr * (1 / drs) * (-1);
          const double vL = av.x * drn.x + av.y * drn.y;
          const double vR = nv.x * drn.x + nv.y * drn.y;
          const double avgc = c0;
          const double avgrho = 0.5 * (arho + nrho);
          const double vss = vel_riemann(arho, nrho, vL, vR, avgrho, avgc);
          double rhos = dens_riemann_lim(arho, nrho, vL, vR, 0.5 * (arho + nrho), avgc);
          double ps = tait_pressure(rhos, rho0, c0);
          const V2 vs = drn * vss + ((av + nv) * 0.5 - drn * (vL + vR) * 0.5);
          const V2 vz = av;
          const double vsdW = (vs.x - vz.x) * dW.x + (vs.y - vz.y) * dW.y;
          if (s.type[ni] == T_WALL) {
            const V2 nb = s.n[ni] * (-1);
            // wall state from the one-sided problem, linear EOS (:224-225)
            rhos = arho - ((0 - av.x) * nb.x + (0 - av.y) * nb.y) * arho / c0;
            ps = c0 * c0 * (rhos - rho0) * 1;
            domegav_sum.x -= 2 * W * dp * (ps * nb.x);
            domegav_sum.y -= 2 * W * dp * (ps * nb.y);
            const double dvn = dv.x * nb.x + dv.y * nb.y;
            domega_sum -= dvn * W * dp;
          } else {
            domegav_sum.x -= 2 * no * ((rhos * vs.x * (vs.x - vz.x) + ps) * dW.x +
                                      (rhos * vs.x * (vs.y - vz.y)) * dW.y);
            domegav_sum.y -= 2 * no * ((rhos * vs.y * (vs.x - vz.x)) * dW.x +
                                      (rhos * vs.y * (vs.y - vz.y) + ps) * dW.y);
            domega_sum -= dvdW * no;
            domegarho_sum -= 2 * rhos * no * vsdW;
          }
          gamma_sum += W * m / nrho;
        }
      }
      if (gamma_sum == 0) {
        s.domega[ai] = 0.;
        s.domegarho[ai] = 0.;
        s.domegav[ai] = V2{0., 0.};
      } else {
        s.domega[ai] = domega_sum * ao;
        s.domegarho[ai] = domegarho_sum * ao;
        s.domegav[ai] = domegav_sum * ao - g * ao * arho;
      }
    }
  }
}

// Integrate_LeapFrog_partOne, src/RALESPH_integration.h:10-36 (fluid only)
void rale_leapfrog_one(Case& s, double dt) {
#pragma omp parallel for schedule(static) num_threads(s.nthreads)
  for (int p = 0; p < s.np; ++p) {
    if (s.type[p] != T_FLUID) continue;
    s.omega_o[p] = s.omega[p];
    s.omegarho_o[p] = s.omegarho[p];
    s.omegav_o[p] = s.omegav[p];
    s.omega[p] = s.omega[p] + s.domega[p] * dt * 0.5;
    s.omegarho[p] = s.omegarho[p] + s.domegarho[p] * dt * 0.5;
    s.omegav[p] = s.omegav[p] + s.domegav[p] * dt * 0.5;
    s.rho[p] = s.omegarho[p] / (s.omega[p]);
    s.v[p] = s.omegav[p] / (s.rho[p] * s.omega[p]);
  }
}

// Integrate_LeapFrog_partTwo, src/RALESPH_integration.h:38-107 (fluid only)
void rale_leapfrog_two(Case& s, double dt, int step) {
#pragma omp parallel for schedule(static) num_threads(s.nthreads)
  for (int p = 0; p < s.np; ++p) {
    if (s.type[p] != T_FLUID) continue;
    if (step == 1) {
      s.omega_o[p] = s.omega[p];
      s.omegarho_o[p] = s.omegarho[p];
      s.omegav_o[p] = s.omegav[p];
      s.omega[p] = s.omega[p] + s.domega[p] * dt * 0.5;
      s.omegarho[p] = s.omegarho[p] + s.domegarho[p] * dt * 0.5;
      s.omegav[p] = s.omegav[p] + s.domegav[p] * dt * 0.5;
      s.rho[p] = s.omegarho[p] / (s.omega[p]);
      s.v[p] = s.omegav[p] / (s.rho[p] * s.omega[p]);
      s.r[p] = s.r[p] + s.v[p] * dt * 0.5;
    } else {
      s.r[p] = s.r[p] + s.v[p] * dt;
      s.omega[p] = s.omega_o[p] + s.domega[p] * dt;
      s.omegarho[p] = s.omegarho_o[p] + s.domegarho[p] * dt;
      s.omegav[p] = s.omegav_o[p] + s.domegav[p] * dt;
      s.rho[p] = s.omegarho[p] / (s.omega[p]);
      s.v[p] = s.omegav[p] / (s.rho[p] * s.omega[p]);
    }
  }
}

// ---------------------------------------------------------------------------
// ALE-SPH with MUSCL reconstruction, the pipeline of
// cases/ALESPH_dambreakLong/SPH_simulation.h:78-125:
//   Integrate_LeapFrog_partOne_ALESPH  src/ALESPH_integration.h:9-31
//   Compute_Density                    src/ALESPH_density_and_omega_bt.h:9-281
//   Kernel_gradient_approx             src/ALESPH_kernel_gradient.h:10-190
//   Integrate_density_compute_pressure src/ALESPH_integration.h:86-135
//   Compute_Acceleration_BT            src/ALESPH_acceleration_bt.h:12-295
//   Integrate_LeapFrog_partTwo_ALESPH  src/ALESPH_integration.h:33-83
// ---------------------------------------------------------------------------

// MUSCLleft / MUSCLright with weight alpha, src/ALESPH_riemann_solver.h:195-235
inline double muscl_left(double A, V2 g, V2 dr, double alpha) {
  return A + alpha * 0.5 * (g.x * dr.x + g.y * dr.y);
}
inline double muscl_right(double A, V2 g, V2 dr, double alpha) {
  return A - alpha * 0.5 * (g.x * dr.x + g.y * dr.y);
}
// minmod_r, src/ALESPH_riemann_solver.h:180-191, with r_CD the vector from
// the reconstructing particle C to the other particle D.  The reference never
// calls it; for equal values (aD == aC) the ratio is undefined there and the
// restatement takes phi = 0 (no reconstruction).
inline double minmod_phi(double aC, double aD, V2 gC, V2 rCD) {
  if (aD - aC == 0) return 0.;
  const double r = 2 * (gC.x * rCD.x + gC.y * rCD.y) / (aD - aC) - 1.;
  const double mn = 1 < r ? 1 : r;  // MIN(1, r)
  return 0 > mn ? 0 : mn;            // MAX(0, MIN(1, r))
}
// cs, src/ALESPH_riemann_solver.h:148-159 (gamma = 2)
inline double ale_cs(double rho, double rho0, double c0) {
  const double gamma = 2.;
  return c0 * std::pow((rho / rho0), 0.5 * (gamma - 1));
}

// The reconstructed left/right states of one pair.  limited = 0 follows the
// reference (alpha = 1); limited = 1 weights each reconstruction by the
// minmod limiter of its own variable.
struct MusclLR { double rhoL, rhoR; V2 vL, vR; };
inline MusclLR muscl_states_of(double arho, double nrho, V2 av, V2 nv, V2 ga, V2 gn, V2 gax,
                               V2 gay, V2 gnx, V2 gny, V2 dr, int limited) {
  const V2 rij = V2{0., 0.} - dr;  // r_j - r_i
  double a1 = 1, a2 = 1, a3 = 1, a4 = 1, a5 = 1, a6 = 1;
  if (limited) {
    a1 = minmod_phi(arho, nrho, ga, rij);
    a2 = minmod_phi(nrho, arho, gn, dr);
    a3 = minmod_phi(av.x, nv.x, gax, rij);
    a4 = minmod_phi(av.y, nv.y, gay, rij);
    a5 = minmod_phi(nv.x, av.x, gnx, dr);
    a6 = minmod_phi(nv.y, av.y, gny, dr);
  }
  MusclLR o;
  o.rhoL = muscl_left(arho, ga, dr, a1);
  o.rhoR = muscl_right(nrho, gn, dr, a2);
  o.vL.x = muscl_left(av.x, gax, dr, a3);
  o.vL.y = muscl_left(av.y, gay, dr, a4);
  o.vR.x = muscl_right(nv.x, gnx, dr, a5);
  o.vR.y = muscl_right(nv.y, gny, dr, a6);
  return o;
}
inline MusclLR muscl_states(const Case& s, int ai, int ni, V2 dr, int limited) {
  return muscl_states_of(s.rho[ai], s.rho[ni], s.v[ai], s.v[ni], s.gradrho[ai], s.gradrho[ni],
                         s.gradvx[ai], s.gradvy[ai], s.gradvx[ni], s.gradvy[ni], dr, limited);
}

// Compute_Density, src/ALESPH_density_and_omega_bt.h:9-281.  Active: every
// particle except outletf / outlet / inlet / wall (:58-61).  Quirk kept: the
// fluid-neighbour volume flux reads `dvdW`, which the reference declares but
// never assigns (:135, :237); its own build (g++ -O3, src/Makefile flags)
// evaluates that term as zero, and so does this restatement -- only wall
// neighbours change omega (:219-232).
void ale_compute_density(Case& s, int limited) {
  const Params& q = s.q;
  const int nc = ncells(s);
  const double h = q.h, m = q.m, c0 = q.cs, dp = q.dp;
#pragma omp parallel for schedule(dynamic, 64) num_threads(s.nthreads)
  for (int c = 0; c < nc; ++c) {
    int ac[9];
    stencil(c, s.ncy, ac);
    for (int ii = s.cstart[c]; ii < s.cstart[c + 1]; ++ii) {
      const int ai = s.citems[ii];
      const int at = s.type[ai];
      if (at == T_OUTLETF || at == T_OUTLET || at == T_INLET || at == T_WALL) continue;
      const V2 ar = s.r[ai], av = s.v[ai];
      const double aomega = s.omega[ai];
      double domega_temp = 0, domegarho_temp = 0, gamma = 0;
      for (int k = 0; k < 9; ++k) {
        const int cl = ac[k];
        if (cl < 0) continue;
        for (int jj = s.cstart[cl]; jj < s.cstart[cl + 1]; ++jj) {
          const int ni = s.citems[jj];
          if (ai == ni) continue;
          const V2 nr = s.r[ni], nv = s.v[ni];
          const double nrho = s.rho[ni], nomega = s.omega[ni];
          const V2 dr = ar - nr, dv = av - nv;
          const double drs = std::sqrt((dr.x * dr.x) + (dr.y * dr.y));
          const V2 drn = dr * (1 / drs) * (-1);
          const KW kw = wendland(drs, h);
          const double W = kw.w;
          const V2 dW = dr * kw.f;
          const MusclLR lr = muscl_states(s, ai, ni, dr, limited);
          const double vl = lr.vL.x * drn.x + lr.vL.y * drn.y;
          const double vr = lr.vR.x * drn.x + lr.vR.y * drn.y;
          const double avgc = c0;
          const double rhos = dens_riemann_lim(lr.rhoL, lr.rhoR, vl, vr, 0.5 * (lr.rhoL + lr.rhoR), avgc);
          const double vss = vel_riemann(lr.rhoL, lr.rhoR, vl, vr, 0.5 * (lr.rhoL + lr.rhoR), avgc);
          const V2 vs = drn * vss + ((av + nv) * 0.5 - drn * (vl + vr) * 0.5);
          const V2 vz = av;
          const double vsdW = (vs.x - vz.x) * dW.x + (vs.y - vz.y) * dW.y;
          if (s.type[ni] == T_WALL) {
            const V2 nb = s.n[ni] * (-1);
            const double dvn = dv.x * nb.x + dv.y * nb.y;
            domega_temp -= dvn * W * dp;
          } else {
            domega_temp -= 0. * nomega;  // the unassigned dvdW, as the reference build evaluates it
            domegarho_temp -= 2 * rhos * nomega * vsdW;
          }
          gamma += W * m / nrho;
        }
      }
      if (gamma == 0) {
        s.domega[ai] = 0.;
        s.domegarho[ai] = 0.;
      } else {
        s.domega[ai] = domega_temp * aomega;
        s.domegarho[ai] = domegarho_temp * aomega;
      }
    }
  }
}

// Kernel_gradient_approx, src/ALESPH_kernel_gradient.h:10-190 (fluid only)
void ale_kernel_gradient(Case& s) {
  const Params& q = s.q;
  const int nc = ncells(s);
  const double h = q.h, m = q.m, dp = q.dp;
#pragma omp parallel for schedule(dynamic, 64) num_threads(s.nthreads)
  for (int c = 0; c < nc; ++c) {
    int ac[9];
    stencil(c, s.ncy, ac);
    for (int ii = s.cstart[c]; ii < s.cstart[c + 1]; ++ii) {
      const int ai = s.citems[ii];
      if (s.type[ai] != T_FLUID) continue;
      const V2 ar = s.r[ai], av = s.v[ai];
      const double arho = s.rho[ai];
      V2 grt{0., 0.}, gxt{0., 0.}, gyt{0., 0.};
      double gamma = 0;
      for (int k = 0; k < 9; ++k) {
        const int cl = ac[k];
        if (cl < 0) continue;
        for (int jj = s.cstart[cl]; jj < s.cstart[cl + 1]; ++jj) {
          const int ni = s.citems[jj];
          if (ai == ni) continue;
          const V2 nr = s.r[ni], nv = s.v[ni];
          const double nrho = s.rho[ni], nomega = s.omega[ni];
          const V2 dr = ar - nr;
          const double drs = std::sqrt((dr.x * dr.x) + (dr.y * dr.y));
          const KW kw = wendland(drs, h);
          const double W = kw.w;
          const V2 dW = dr * kw.f;
          if (s.type[ni] == T_WALL) {
            const V2 nb = s.n[ni] * (-1);
            grt.x += (nrho - arho) * nb.x * W * dp;
            grt.y += (nrho - arho) * nb.y * W * dp;
            gxt.x += (nv.x - av.x) * nb.x * W * dp;
            gxt.y += (nv.x - av.x) * nb.y * W * dp;
            gyt.x += (nv.y - av.y) * nb.x * W * dp;
            gyt.y += (nv.y - av.y) * nb.y * W * dp;
          } else {
            grt.x += (nrho - arho) * nomega * dW.x;
            grt.y += (nrho - arho) * nomega * dW.y;
            gxt.x += (nv.x - av.x) * nomega * dW.x;
            gxt.y += (nv.x - av.x) * nomega * dW.y;
            gyt.x += (nv.y - av.y) * nomega * dW.x;
            gyt.y += (nv.y - av.y) * nomega * dW.y;
          }
          gamma += W * m / nrho;
        }
      }
      if (gamma == 0) {
        s.gradrho[ai] = s.gradvx[ai] = s.gradvy[ai] = V2{0., 0.};
      } else {
        s.gradrho[ai] = grt * 2.;
        s.gradvx[ai] = gxt * 2.;
        s.gradvy[ai] = gyt * 2.;
      }
    }
  }
}

// Integrate_density_compute_pressure, src/ALESPH_integration.h:86-135 (fluid
// only; linear EOS Compute_Pressure2, src/SPH_pressure.h:24-34, clamped at 0).
// Returns the count of densities outside [0, 1500] (the reference prints them).
int ale_density_pressure(Case& s, double dt) {
  int bad = 0;
  for (int p = 0; p < s.np; ++p) {
    if (s.type[p] != T_FLUID) continue;  // :95-101
    s.omega[p] = s.omega[p] + s.domega[p] * dt;
    s.omegarho[p] = s.omegarho[p] + s.domegarho[p] * dt;
    s.rho[p] = s.omegarho[p] / (s.omega[p]);
    if (s.rho[p] > 1500. || s.rho[p] < 0.) ++bad;
    s.p[p] = s.q.cs * s.q.cs * (s.rho[p] - s.q.rho0);
    if (s.p[p] < 0) s.p[p] = 0;
  }
  return bad;
}

// Compute_Acceleration_BT, src/ALESPH_acceleration_bt.h:12-295 (fluid only).
// Sound speed of the pair: the sum cs(rho_i) + cs(rho_j) (:178); v* from the
// unreconstructed densities with the reconstructed average (:185).
void ale_compute_acceleration(Case& s, int limited) {
  const Params& q = s.q;
  const int nc = ncells(s);
  const double h = q.h, m = q.m, dp = q.dp, c0 = q.cs, rho0 = q.rho0;
  const V2 g{q.gx, q.gy};
#pragma omp parallel for schedule(dynamic, 64) num_threads(s.nthreads)
  for (int c = 0; c < nc; ++c) {
    int ac[9];
    stencil(c, s.ncy, ac);
    for (int ii = s.cstart[c]; ii < s.cstart[c + 1]; ++ii) {
      const int ai = s.citems[ii];
      if (s.type[ai] != T_FLUID) continue;
      const V2 ar = s.r[ai], av = s.v[ai];
      const double arho = s.rho[ai], aomega = s.omega[ai];
      V2 omegaa_temp{0., 0.};
      double gamma = 0;
      for (int k = 0; k < 9; ++k) {
        const int cl = ac[k];
        if (cl < 0) continue;
        for (int jj = s.cstart[cl]; jj < s.cstart[cl + 1]; ++jj) {
          const int ni = s.citems[jj];
          if (ai == ni) continue;
          const V2 nr = s.r[ni], nv = s.v[ni];
          const double nrho = s.rho[ni], nomega = s.omega[ni];
          const V2 dr = ar - nr;
          const double drs = std::sqrt((dr.x * dr.x) + (dr.y * dr.y));
          const V2 drn = dr * (1 / drs) * (-1);
          const KW kw = wendland(drs, h);
          const double W = kw.w;
          const V2 dW = dr * kw.f;
          const MusclLR lr = muscl_states(s, ai, ni, dr, limited);
          const double vl = lr.vL.x * drn.x + lr.vL.y * drn.y;
          const double vr = lr.vR.x * drn.x + lr.vR.y * drn.y;
          const double avgc = ale_cs(arho, rho0, c0) + ale_cs(nrho, rho0, c0);
          double rhos = dens_riemann_lim(lr.rhoL, lr.rhoR, vl, vr, 0.5 * (lr.rhoL + lr.rhoR), avgc);
          const double vss = vel_riemann(arho, nrho, vl, vr, 0.5 * (lr.rhoL + lr.rhoR), avgc);
          const V2 vs = drn * vss + ((av + nv) * 0.5 - drn * (vl + vr) * 0.5);
          double ps = c0 * c0 * (rhos - rho0);
          const V2 vz = av;
          if (s.type[ni] == T_WALL) {
            const V2 nb = s.n[ni] * (-1);
            rhos = lr.rhoL - ((0 - lr.vL.x) * nb.x + (0 - lr.vL.y) * nb.y) * lr.rhoL / c0;
            ps = c0 * c0 * (rhos - 1000.) * 1;
            omegaa_temp.x -= 2 * W * dp * (ps * nb.x);
            omegaa_temp.y -= 2 * W * dp * (ps * nb.y);
          } else {
            omegaa_temp.x -= 2 * nomega * ((rhos * vs.x * (vs.x - vz.x) + ps) * dW.x +
                                           (rhos * vs.x * (vs.y - vz.y)) * dW.y);
            omegaa_temp.y -= 2 * nomega * ((rhos * vs.y * (vs.x - vz.x)) * dW.x +
                                           (rhos * vs.y * (vs.y - vz.y) + ps) * dW.y);
          }
          gamma += W * m / nrho;
        }
      }
      if (gamma == 0)
        s.domegav[ai] = V2{0., 0.} - g * aomega * arho;
      else
        s.domegav[ai] = omegaa_temp * aomega - g * aomega * arho;
    }
  }
}

// Integrate_LeapFrog_partOne_ALESPH, src/ALESPH_integration.h:9-31 (fluid only)
void ale_leapfrog_one(Case& s, double dt) {
  for (int p = 0; p < s.np; ++p) {
    if (s.type[p] != T_FLUID) continue;
    s.omegav[p] = s.omegav[p] + s.domegav[p] * dt * 0.5;
    s.v[p] = s.omegav[p] / (s.rho[p] * s.omega[p]);
  }
}

// Integrate_LeapFrog_partTwo_ALESPH, src/ALESPH_integration.h:33-83 (fluid only)
void ale_leapfrog_two(Case& s, double dt, int step) {
  for (int p = 0; p < s.np; ++p) {
    if (s.type[p] != T_FLUID) continue;
    if (step == 1) {
      s.omegav[p] = s.omegav[p] + s.domegav[p] * dt * 0.5;
      s.v[p] = s.omegav[p] / (s.rho[p] * s.omega[p]);
      s.r[p] = s.r[p] + s.v[p] * dt * 0.5;
    } else {
      s.omegav[p] = s.omegav[p] + s.domegav[p] * dt;
      s.v[p] = s.omegav[p] / (s.rho[p] * s.omega[p]);
      s.r[p] = s.r[p] + s.v[p] * dt;
    }
  }
}

// ---------------------------------------------------------------------------
// Buffer zones, src/SPH_buffer_zones.h:10-206
// ---------------------------------------------------------------------------
// Particle_system::Add_particle, src/SPH_particle_system.h:25-50 (wall extras
// are kept per particle here and get a zero entry; ALE variables, when
// present, the ALESPH_Add_particle values)
void add_particle(Case& s, int type, double p, double rho, V2 r, V2 v, V2 a) {
  s.np++;
  s.type.push_back(type);
  s.index.push_back(s.np - 1);
  s.p.push_back(p);
  s.rho.push_back(rho);
  s.drho.push_back(0.);
  s.r.push_back(r);
  s.v.push_back(v);
  s.a.push_back(a);
  s.r_o.push_back(r);
  s.rho_o.push_back(rho);
  s.rho_oo.push_back(rho);
  s.v_o.push_back(v);
  s.v_oo.push_back(v);
  s.rho_tmp.push_back(rho);
  s.gnr.push_back(V2{0., 0.});
  s.n.push_back(V2{0., 0.});
  if (!s.omega.empty()) {
    const double w = s.q.dp * s.q.dp;
    s.omega.push_back(w); s.omegarho.push_back(rho * w); s.omegav.push_back(v * w);
    s.omega_o.push_back(w); s.omegarho_o.push_back(rho * w); s.omegav_o.push_back(v * w);
    s.domega.push_back(0.); s.domegarho.push_back(0.); s.domegav.push_back(V2{0., 0.});
    s.gradrho.push_back(V2{0., 0.}); s.gradvx.push_back(V2{0., 0.}); s.gradvy.push_back(V2{0., 0.});
  }
}

// Remove_particle, src/SPH_particle_system.h:110-146 (erases every per-particle field)
void erase_particle(Case& s, int p) {
  auto er = [p](auto& vec) { if ((int)vec.size() > p) vec.erase(vec.begin() + p); };
  er(s.type); er(s.index); er(s.p); er(s.rho); er(s.drho); er(s.r); er(s.v); er(s.a);
  er(s.r_o); er(s.v_o); er(s.rho_o); er(s.gnr); er(s.n);
  er(s.v_oo); er(s.rho_oo); er(s.rho_tmp);
  if (!s.omega.empty()) {
    er(s.omega); er(s.omegarho); er(s.omegav); er(s.omega_o); er(s.omegarho_o); er(s.omegav_o);
    er(s.domega); er(s.domegarho); er(s.domegav); er(s.gradrho); er(s.gradvx); er(s.gradvy);
  }
  s.np--;
  s.np_out++;
}

// BZ_Integrate_inlet, src/SPH_buffer_zones.h:10-107.  The loop bound grows
// with every appended particle, so particles appended during the pass are
// visited in the same pass.  Returns the number of particles appended.
int bz_integrate_inlet(Case& s, int zi, double dt) {
  const Zone& z = s.zones[zi];
  const double b_len = z.bl;
  const V2 v_inl = z.v_b;
  const int btype = z.btype;
  int added = 0;
  if (z.orientation.x != 0) {
    const double x_b = z.x0;
    const int orientation = (int)z.orientation.x;
    if (orientation < 0) {
      for (int p = 0; p < s.np; p++) {
        if (s.type[p] != btype) continue;
        if (s.r[p].x > x_b) {
          s.type[p] = T_FLUID;
          add_particle(s, btype, 0, 1000, V2{s.r[p].x - b_len, s.r[p].y}, v_inl, V2{0., 0.});
          ++added;
        }
      }
    } else if (orientation > 0) {
      for (int p = 0; p < s.np; p++) {
        if (s.type[p] != btype) continue;
        s.r[p].x = s.r[p].x + s.v[p].x * dt;
        if (s.r[p].x < x_b) {
          s.type[p] = T_FLUID;
          add_particle(s, btype, 0, 1000, V2{s.r[p].x + b_len, s.r[p].y}, v_inl, V2{0., 0.});
          ++added;
        }
      }
    }
  } else if (z.orientation.y != 0) {
    const double x_b = z.y0;
    const int orientation = (int)z.orientation.y;
    if (orientation < 0) {
      for (int p = 0; p < s.np; p++) {
        if (s.type[p] != btype) continue;
        s.r[p].y = s.r[p].y + s.v[p].y * dt;
        if (s.r[p].y > x_b) {
          s.type[p] = T_FLUID;
          add_particle(s, btype, 0, 1000, V2{s.r[p].x, s.r[p].y - b_len}, v_inl, V2{0., 0.});
          ++added;
        }
      }
    } else if (orientation > 0) {
      for (int p = 0; p < s.np; p++) {
        if (s.type[p] != btype) continue;
        s.r[p].y = s.r[p].y + s.v[p].y * dt;
        if (s.r[p].y < x_b) {
          s.type[p] = T_FLUID;
          add_particle(s, btype, 0, 1000, V2{s.r[p].x, s.r[p].y + b_len}, v_inl, V2{0., 0.});
          ++added;
        }
      }
    }
  }
  return added;
}

// BZ_Integrate_outlet, src/SPH_buffer_zones.h:111-206: fluid beyond the zone
// start becomes the zone type (a = 0, drho = 0); zone particles past the zone
// or below dp/2 are erased inside the ascending loop, so the particle that
// slides into the slot is not examined.  y-oriented zones do nothing.
// Returns the number of particles removed.
int bz_integrate_outlet(Case& s, int zi, double dt) {
  (void)dt;
  const Zone& z = s.zones[zi];
  const double dp = s.q.dp;
  const double b_len = z.bl;
  const int btype = z.btype;
  int removed = 0;
  if (z.orientation.x != 0) {
    const double x_b = z.x0;
    const int orientation = (int)z.orientation.x;
    for (int p = 0; p < s.np; p++) {
      if (s.type[p] == T_FLUID) {
        if (orientation > 0 ? s.r[p].x > x_b : s.r[p].x < x_b) {
          s.type[p] = btype;
          s.a[p] = V2{0., 0};
          s.drho[p] = 0.;
        }
      } else if (s.type[p] == btype) {
        const bool out = orientation > 0 ? (s.r[p].x > x_b + b_len || s.r[p].y < dp / 2)
                                         : (s.r[p].x < x_b - b_len || s.r[p].y < dp / 2);
        if (out) {
          erase_particle(s, p);
          ++removed;
        }
      }
    }
  }
  return removed;
}

}  // namespace

extern "C" {

// q = {h, kap, cs, rho0, avisc, delta, dp, x_0, x_m, y_0, y_m}
void* orc_new(const double* q) {
  Case* s = new Case();
  Params& p = s->q;
  p.h = q[0]; p.kap = q[1]; p.cs = q[2]; p.rho0 = q[3]; p.avisc = q[4]; p.delta = q[5];
  p.dp = q[6]; p.x0 = q[7]; p.xm = q[8]; p.y0 = q[9]; p.ym = q[10];
  p.m = p.rho0 / (1 / p.dp * 1 / p.dp);  // src/SPH_particle_system.h:207
  p.gx = 0.;
  p.gy = 9.81;  // src/SPH_particle_system.h:205
  return s;
}
void orc_free(void* c) { delete static_cast<Case*>(c); }
void orc_set_threads(void* c, int n) { static_cast<Case*>(c)->nthreads = n > 0 ? n : 1; }
double orc_mass(void* c) { return static_cast<Case*>(c)->q.m; }
int orc_np(void* c) { return static_cast<Case*>(c)->np; }
int orc_np_out(void* c) { return static_cast<Case*>(c)->np_out; }

// Load a whole particle set in reference order.  r, v are interleaved xy;
// gnr / n may be NULL.  a, drho start at 0 and the "old" layers at the
// current values, like Add_particle (src/SPH_particle_system.h:25-50).
void orc_load(void* vc, int np, const int* type, const double* p, const double* rho,
              const double* r, const double* v, const double* gnr, const double* n) {
  Case& s = *static_cast<Case*>(vc);
  s.np = np;
  s.type.assign(type, type + np);
  s.index.resize(np);
  for (int i = 0; i < np; ++i) s.index[i] = i;
  s.p.assign(p, p + np);
  s.rho.assign(rho, rho + np);
  s.rho_o = s.rho;
  s.drho.assign(np, 0.);
  s.r.resize(np); s.v.resize(np); s.a.assign(np, V2{0., 0.});
  for (int i = 0; i < np; ++i) {
    s.r[i] = {r[2 * i], r[2 * i + 1]};
    s.v[i] = {v[2 * i], v[2 * i + 1]};
  }
  s.r_o = s.r; s.v_o = s.v;
  s.v_oo = s.v; s.rho_oo = s.rho; s.rho_tmp = s.rho;
  s.gnr.assign(np, V2{0., 0.});
  s.n.assign(np, V2{0., 0.});
  if (gnr) for (int i = 0; i < np; ++i) s.gnr[i] = {gnr[2 * i], gnr[2 * i + 1]};
  if (n) for (int i = 0; i < np; ++i) s.n[i] = {n[2 * i], n[2 * i + 1]};
}

// Field ids as the reference harness: 0 r, 1 v, 2 a, 3 p, 4 rho, 5 drho, 6 r_o,
// 7 v_o, 8 rho_o, 9 gnr, 10 n.
// ALE fields: 11 omega, 12 omegarho, 13 omegav, 14 omega_o, 15 omegarho_o,
// 16 omegav_o, 17 domega, 18 domegarho, 19 domegav, MUSCL: 20 gradrho,
// 21 gradvx, 22 gradvy; history: 23 v_oo, 24 rho_oo, 25 rhoTemp
static std::vector<V2>* v2f(Case& s, int f) {
  switch (f) { case 0: return &s.r; case 1: return &s.v; case 2: return &s.a; case 6: return &s.r_o;
    case 7: return &s.v_o; case 9: return &s.gnr; case 10: return &s.n; case 13: return &s.omegav;
    case 16: return &s.omegav_o; case 19: return &s.domegav; case 20: return &s.gradrho;
    case 21: return &s.gradvx; case 22: return &s.gradvy; case 23: return &s.v_oo;
    default: return nullptr; }
}
static std::vector<double>* sf(Case& s, int f) {
  switch (f) { case 3: return &s.p; case 4: return &s.rho; case 5: return &s.drho; case 8: return &s.rho_o;
    case 11: return &s.omega; case 12: return &s.omegarho; case 14: return &s.omega_o;
    case 15: return &s.omegarho_o; case 17: return &s.domega; case 18: return &s.domegarho;
    case 24: return &s.rho_oo; case 25: return &s.rho_tmp; default: return nullptr; }
}

// ALE volume variables as ALESPH_Add_particle sets them
// (src/SPH_particle_system.h:51-108): omega = dp^2, omega*rho, omega*v
void orc_init_ale(void* vc) {
  Case& s = *static_cast<Case*>(vc);
  const double w = s.q.dp * s.q.dp;
  s.omega.assign(s.np, w);
  s.omegarho.resize(s.np);
  s.omegav.resize(s.np);
  for (int i = 0; i < s.np; ++i) {
    s.omegarho[i] = s.rho[i] * (s.q.dp * s.q.dp);
    s.omegav[i] = s.v[i] * (s.q.dp * s.q.dp);
  }
  s.omega_o = s.omega;
  s.omegarho_o = s.omegarho;
  s.omegav_o = s.omegav;
  s.domega.assign(s.np, 0.);
  s.domegarho.assign(s.np, 0.);
  s.domegav.assign(s.np, V2{0., 0.});
  s.gradrho.assign(s.np, V2{0., 0.});
  s.gradvx.assign(s.np, V2{0., 0.});
  s.gradvy.assign(s.np, V2{0., 0.});
}
void orc_ale_compute_density(void* vc) { ale_compute_density(*static_cast<Case*>(vc), 0); }
void orc_ale_compute_density_limited(void* vc) { ale_compute_density(*static_cast<Case*>(vc), 1); }
void orc_ale_compute_acceleration(void* vc) { ale_compute_acceleration(*static_cast<Case*>(vc), 0); }
void orc_ale_compute_acceleration_limited(void* vc) {
  ale_compute_acceleration(*static_cast<Case*>(vc), 1);
}
void orc_ale_kernel_gradient(void* vc) { ale_kernel_gradient(*static_cast<Case*>(vc)); }
int orc_ale_density_pressure(void* vc, double dt) { return ale_density_pressure(*static_cast<Case*>(vc), dt); }
void orc_ale_leapfrog_one(void* vc, double dt) { ale_leapfrog_one(*static_cast<Case*>(vc), dt); }
void orc_ale_leapfrog_two(void* vc, double dt, int step) {
  ale_leapfrog_two(*static_cast<Case*>(vc), dt, step);
}
// the reconstructed states of one pair (test hook): in = arho, nrho, av.x, av.y,
// nv.x, nv.y, then the gradients ga, gn, gax, gay, gnx, gny and dr = r_i - r_j as
// x, y pairs (20 values); out = rhoL, rhoR, vL.x, vL.y, vR.x, vR.y
void orc_muscl_states(const double* in, int limited, double* out) {
  auto v2 = [&](int k) { return V2{in[k], in[k + 1]}; };
  const MusclLR o = muscl_states_of(in[0], in[1], v2(2), v2(4), v2(6), v2(8), v2(10), v2(12),
                                    v2(14), v2(16), v2(18), limited);
  out[0] = o.rhoL; out[1] = o.rhoR; out[2] = o.vL.x; out[3] = o.vL.y; out[4] = o.vR.x; out[5] = o.vR.y;
}

double orc_minmod_phi(double aC, double aD, double gx, double gy, double rx, double ry) {
  return minmod_phi(aC, aD, V2{gx, gy}, V2{rx, ry});
}
void orc_rale_forces(void* vc) { rale_forces(*static_cast<Case*>(vc)); }
// BufferZone(x0, xm, y0, ym, nbl, v0, orientation, btype), src/SPH_inlet_outlet_data.h:32-44
int orc_add_zone(void* vc, double x0, double xm, double y0, double ym, double vx, double vy,
                 double ox, double oy, int btype) {
  Case& s = *static_cast<Case*>(vc);
  Zone z;
  z.x0 = x0; z.xm = xm; z.y0 = y0; z.ym = ym; z.v_b = V2{vx, vy}; z.orientation = V2{ox, oy};
  z.btype = btype;
  z.bl = std::fabs(xm - x0);
  z.bh = std::fabs(ym - y0);
  if (z.orientation.x != 0) { z.bl = std::fabs(xm - x0); z.bh = std::fabs(ym - y0); }
  else if (z.orientation.y != 0) { z.bl = std::fabs(ym - y0); z.bh = std::fabs(xm - x0); }
  s.zones.push_back(z);
  return (int)s.zones.size() - 1;
}
int orc_bz_integrate_inlet(void* vc, int zone, double dt) {
  return bz_integrate_inlet(*static_cast<Case*>(vc), zone, dt);
}
int orc_bz_integrate_outlet(void* vc, int zone, double dt) {
  return bz_integrate_outlet(*static_cast<Case*>(vc), zone, dt);
}
void orc_rale_leapfrog_one(void* vc, double dt) { rale_leapfrog_one(*static_cast<Case*>(vc), dt); }
void orc_rale_leapfrog_two(void* vc, double dt, int step) {
  rale_leapfrog_two(*static_cast<Case*>(vc), dt, step);
}
int orc_get(void* vc, int f, double* out) {
  Case& s = *static_cast<Case*>(vc);
  if (auto* v = v2f(s, f)) { for (int i = 0; i < s.np; ++i) { out[2 * i] = (*v)[i].x; out[2 * i + 1] = (*v)[i].y; } return 2 * s.np; }
  if (auto* x = sf(s, f)) { std::memcpy(out, x->data(), sizeof(double) * s.np); return s.np; }
  return 0;
}
int orc_set(void* vc, int f, const double* in) {
  Case& s = *static_cast<Case*>(vc);
  if (auto* v = v2f(s, f)) { for (int i = 0; i < s.np; ++i) (*v)[i] = {in[2 * i], in[2 * i + 1]}; return 2 * s.np; }
  if (auto* x = sf(s, f)) { std::memcpy(x->data(), in, sizeof(double) * s.np); return s.np; }
  return 0;
}
void orc_get_type(void* vc, int* out) { Case& s = *static_cast<Case*>(vc); std::copy(s.type.begin(), s.type.end(), out); }
void orc_get_index(void* vc, int* out) { Case& s = *static_cast<Case*>(vc); std::copy(s.index.begin(), s.index.end(), out); }

// src/SPH_find_pairs.h:5-42
void orc_divide_to_cells(void* vc) {
  Case& s = *static_cast<Case*>(vc);
  const double kh = kh_of(s.q), of = of_of(s.q);
  s.ncx = (int)((s.q.xm - s.q.x0 + of) / kh) + 1;
  s.ncy = (int)((s.q.ym - s.q.y0 + of) / kh) + 1;
  s.cstart.assign(ncells(s) + 1, 0);
  s.citems.clear();
}
void orc_grid(void* vc, int* ncx, int* ncy) {
  Case& s = *static_cast<Case*>(vc);
  *ncx = s.ncx; *ncy = s.ncy;
}
// src/SPH_find_pairs.h:74-92
void orc_clear_cells(void* vc) {
  Case& s = *static_cast<Case*>(vc);
  std::fill(s.cstart.begin(), s.cstart.end(), 0);
  s.citems.clear();
}
// src/SPH_find_pairs.h:45-72: serial binning in ascending particle index.  A
// point outside the grid is undefined behaviour in the reference; here it
// returns -1 - p (the first offending particle) instead of writing anywhere.
int orc_particle_to_cell(void* vc) {
  Case& s = *static_cast<Case*>(vc);
  const int nc = ncells(s);
  std::vector<int> cell(s.np);
  std::vector<int> cnt(nc + 1, 0);
  for (int p = 0; p < s.np; ++p) {
    int cx, cy;
    cell_xy(s.q, s.r[p], cx, cy);
    const int c = cy + cx * s.ncy;
    if (c < 0 || c >= nc) return -1 - p;
    cell[p] = c;
    cnt[c + 1]++;
  }
  for (int c = 0; c < nc; ++c) cnt[c + 1] += cnt[c];
  s.cstart = cnt;
  s.citems.assign(s.np, 0);
  std::vector<int> fill(cnt.begin(), cnt.end() - 1);
  for (int p = 0; p < s.np; ++p) s.citems[fill[cell[p]]++] = p;
  return 0;
}
void orc_cell_csr(void* vc, int* start, int* items) {
  Case& s = *static_cast<Case*>(vc);
  std::copy(s.cstart.begin(), s.cstart.end(), start);
  std::copy(s.citems.begin(), s.citems.end(), items);
}

// src/SPH_pressure.h:65-89; returns the number of particles whose density is
// outside [0, 1500] (the reference prints a warning for each and continues).
int orc_density_to_pressure(void* vc) {
  Case& s = *static_cast<Case*>(vc);
  int bad = 0;
#pragma omp parallel for schedule(static) num_threads(s.nthreads) reduction(+ : bad)
  for (int p = 0; p < s.np; ++p) {
    if (s.type[p] == T_OUTLET || s.type[p] == T_INLET) continue;
    if (s.rho[p] > 1500. || s.rho[p] < 0.) ++bad;
    s.p[p] = tait_pressure(s.rho[p], s.q.rho0, s.q.cs);
  }
  return bad;
}

void orc_compute_forces(void* vc, int variant) { compute_forces(*static_cast<Case*>(vc), variant); }
void orc_mdbc_density(void* vc) { mdbc_density(*static_cast<Case*>(vc)); }
void orc_bt_density(void* vc) { bt_density(*static_cast<Case*>(vc)); }

// src/SPH_integration.h:266-287 (all particle types)
void orc_predictor(void* vc, double dt) {
  Case& s = *static_cast<Case*>(vc);
#pragma omp parallel for schedule(static) num_threads(s.nthreads)
  for (int p = 0; p < s.np; ++p) {
    s.rho_o[p] = s.rho[p];
    s.r_o[p] = s.r[p];
    s.v_o[p] = s.v[p];
    s.rho[p] = s.rho[p] + s.drho[p] * dt * 0.5;
    s.r[p] = s.r[p] + s.v[p] * dt * 0.5;
    s.v[p] = s.v[p] + s.a[p] * dt * 0.5;
  }
}

// src/SPH_integration.h:289-309 (all particle types)
void orc_corrector(void* vc, double dt) {
  Case& s = *static_cast<Case*>(vc);
#pragma omp parallel for schedule(static) num_threads(s.nthreads)
  for (int p = 0; p < s.np; ++p) {
    const double e = (-s.drho[p] / s.rho[p]) * dt;
    s.rho[p] = s.rho_o[p] * ((2. - e) / (2. + e));
    s.v[p] = s.v_o[p] + s.a[p] * dt;
    s.r[p] = s.r_o[p] + (s.v[p] + s.v_o[p]) * dt * 0.5;
  }
}

// Integrate_Verlet, src/SPH_integration.h:103-128 (all particle types)
void orc_verlet(void* vc, double dt) {
  Case& s = *static_cast<Case*>(vc);
#pragma omp parallel for schedule(static) num_threads(s.nthreads)
  for (int p = 0; p < s.np; ++p) {
    s.v[p] = s.v_oo[p] + s.a[p] * dt * 2;
    s.r[p] = s.r[p] + s.v_o[p] * dt + s.a[p] * dt * dt * 0.5;
    s.rho[p] = s.rho_oo[p] + s.drho[p] * dt * 2;
    s.v_oo[p] = s.v_o[p];
    s.v_o[p] = s.v[p];
    s.rho_oo[p] = s.rho_o[p];
    s.rho_o[p] = s.rho[p];
  }
}

// Integrate_Euler, src/SPH_integration.h:130-158 (all particle types)
void orc_euler(void* vc, double dt) {
  Case& s = *static_cast<Case*>(vc);
#pragma omp parallel for schedule(static) num_threads(s.nthreads)
  for (int p = 0; p < s.np; ++p) {
    s.v[p] = s.v[p] + s.a[p] * dt;
    s.r[p] = s.r[p] + s.v_o[p] * dt + s.a[p] * dt * dt * 0.5;
    s.rho[p] = s.rho[p] + s.drho[p] * dt;
    s.v_oo[p] = s.v_o[p];
    s.v_o[p] = s.v[p];
    s.rho_oo[p] = s.rho_o[p];
    s.rho_o[p] = s.rho[p];
  }
}

// Integrate_LeapFrog_partOne_withDensity, src/SPH_integration.h:161-183
void orc_leapfrog_one(void* vc, double dt) {
  Case& s = *static_cast<Case*>(vc);
#pragma omp parallel for schedule(static) num_threads(s.nthreads)
  for (int p = 0; p < s.np; ++p) {
    s.v_o[p] = s.v[p];
    s.rho_o[p] = s.rho[p];
    s.v[p] = s.v[p] + s.a[p] * dt * 0.5;
    s.rho[p] += s.drho[p] * dt * 0.5;
  }
}

// Integrate_LeapFrog_partTwo_withDensity, src/SPH_integration.h:185-230
void orc_leapfrog_two(void* vc, double dt, int step) {
  Case& s = *static_cast<Case*>(vc);
  if (step == 1) {
    for (int p = 0; p < s.np; ++p) {
      s.v[p] = s.v[p] + s.a[p] * dt * 0.5;
      s.r[p] = s.r[p] + s.v[p] * dt * 0.5;
    }
  } else {
#pragma omp parallel for schedule(static) num_threads(s.nthreads)
    for (int p = 0; p < s.np; ++p) {
      s.r[p] = s.r[p] + s.v[p] * dt;
      s.v[p] = s.v_o[p] + s.a[p] * dt;
      s.rho[p] = s.rho_o[p] + s.drho[p] * dt;
    }
  }
}

// Density_reinitialization_sumation, src/SPH_density_reinitialization.h:8-121:
// per fluid particle, over the 9 cells in stencil order (cl < 0 skipped; the
// reference reads past the last cell without a check, which no case reaches)
// and ascending index in a cell, every other particle adds m W to the density
// sum and m W / rho_j to the volume sum (W = 0 outside the support, W = awen
// for a coincident particle); rhoTemp = their ratio where the volume sum is
// non-zero, otherwise rhoTemp keeps its value.
void orc_reinit_sum(void* vc) {
  Case& s = *static_cast<Case*>(vc);
  const double h = s.q.h, m = s.q.m;
  const int nc = ncells(s);
#pragma omp parallel for schedule(dynamic, 64) num_threads(s.nthreads)
  for (int c = 0; c < nc; ++c) {
    int ac[9];
    stencil(c, s.ncy, ac);
    for (int k = s.cstart[c]; k < s.cstart[c + 1]; ++k) {
      const int ai = s.citems[k];
      if (s.type[ai] != T_FLUID) continue;
      const V2 ar = s.r[ai];
      double rho_sum = 0, wv_sum = 0;
      for (int q = 0; q < 9; ++q) {
        const int cl = ac[q];
        if (cl < 0 || cl >= nc) continue;
        for (int u = s.cstart[cl]; u < s.cstart[cl + 1]; ++u) {
          const int ni = s.citems[u];
          if (ai == ni) continue;
          const V2 dr = ar - s.r[ni];
          const double drs = std::sqrt((dr.x * dr.x) + (dr.y * dr.y));
          const double W = wendland(drs, h).w;
          rho_sum += m * W;
          wv_sum += m * W / s.rho[ni];
        }
      }
      if (wv_sum != 0) s.rho_tmp[ai] = rho_sum / wv_sum;
    }
  }
}

// Density_reinitialization_overwrite, src/SPH_density_reinitialization.h:123-139;
// returns the fluid particles whose density before the overwrite lay outside
// [990, 1010] (the reference prints a warning for each)
int orc_reinit_overwrite(void* vc) {
  Case& s = *static_cast<Case*>(vc);
  int bad = 0;
  for (int p = 0; p < s.np; ++p) {
    if (s.type[p] != T_FLUID) continue;
    if (s.rho[p] < 990 || s.rho[p] > 1010) ++bad;
    s.rho[p] = s.rho_tmp[p];
  }
  return bad;
}

// src/SPH_btdebug.h:4-31 with Remove_particle, src/SPH_particle_system.h:110-146.
// The erase happens inside the ascending loop without stepping back, so the
// particle that slides into slot p is never examined.  gnr / n are not erased
// by the reference; they are kept aligned with the fluid-last creation order
// by erasing them here only where they exist for that slot (walls precede
// fluid in every case, so the wall entries never move).
int orc_remove_out_of_domain(void* vc) {
  Case& s = *static_cast<Case*>(vc);
  const double x0 = s.q.x0, xm = s.q.xm, y0 = s.q.y0, ym = s.q.ym;
  int removed = 0;
  for (int p = 0; p < s.np; ++p) {
    if (s.type[p] != T_FLUID) continue;
    const V2 r = s.r[p];
    if ((r.y < (y0 - EPS)) || (r.y > ym + EPS) || (r.x < x0 - EPS) || (r.x > xm - EPS)) {
      auto er = [p](auto& vec) { vec.erase(vec.begin() + p); };
      er(s.type); er(s.index); er(s.p); er(s.rho); er(s.drho); er(s.r); er(s.v); er(s.a);
      er(s.r_o); er(s.v_o); er(s.rho_o); er(s.gnr); er(s.n);
      er(s.v_oo); er(s.rho_oo); er(s.rho_tmp);  // src/SPH_particle_system.h:139-144
      // ALE fields erased with the particle (ALESPH_Remove_particle semantics,
      // src/SPH_particle_system.h:148-184; the ALE driver's plain removal
      // leaves them unerased, which misaligns them in the reference)
      if (!s.omega.empty()) {
        er(s.omega); er(s.omegarho); er(s.omegav); er(s.omega_o); er(s.omegarho_o);
        er(s.omegav_o); er(s.domega); er(s.domegarho); er(s.domegav);
        er(s.gradrho); er(s.gradvx); er(s.gradvy);
      }
      s.np--; s.np_out++; removed++;
    }
  }
  return removed;
}

// src/SPH_water_level.h:87-117; returns NaN where the reference leaves WL
// uninitialised (no sample reaches the threshold).
double orc_water_elevation(void* vc, double x, double y0) {
  Case& s = *static_cast<Case*>(vc);
  const double dp_wl = s.q.dp * 0.00005;
  V2 ar{x, y0};
  const int nmax = int(s.q.ym / dp_wl);
  for (int i = 0; i < nmax; ++i) {
    ar.y += i * dp_wl;
    if (sum_kernel_function(s, ar) <= 0.5) return ar.y - i * dp_wl;
  }
  return NAN;
}
// src/SPH_water_level.h:120-150
double orc_water_elevation_inverse(void* vc, double x, double y0) {
  Case& s = *static_cast<Case*>(vc);
  const double dp_wl = s.q.dp * 0.00005;
  V2 ar{x, y0};
  const int nmax = int(s.q.ym / dp_wl);
  for (int i = 0; i < nmax; ++i) {
    ar.y -= i * dp_wl;
    if (sum_kernel_function(s, ar) >= 0.5) return ar.y + i * dp_wl;
  }
  return NAN;
}
double orc_sum_kernel_function(void* vc, double x, double y) {
  return sum_kernel_function(*static_cast<Case*>(vc), V2{x, y});
}
// src/SPH_measure_pressure.h:6-15 -> src/testing.h:6-148; out4 = {rho, p, vx, vy}
void orc_interpolate_point(void* vc, double x, double y, double* out4) {
  interpolate_point(*static_cast<Case*>(vc), V2{x, y}, out4);
}

// Scalar closures exported for the closed-form tests.
double orc_wendland_w(double r, double h) { return wendland(r, h).w; }
double orc_wendland_f(double r, double h) { return wendland(r, h).f; }
double orc_tait(double rho, double rho0, double c0) { return tait_pressure(rho, rho0, c0); }
double orc_pressure_to_density(double p, double rho0, double c0) { return pressure_to_density(p, rho0, c0); }
double orc_artificial_viscosity(double h, double drs, double drdv, double rhobar, double c0, double alpha) {
  return artificial_viscosity(h, drs, drdv, rhobar, c0, alpha);
}
double orc_diffusion_fourtakas(double h, double drs, double dr_z, double drdW, double arho, double nrho,
                               double rho0, double c0, double delta, double m, double g) {
  return diffusion_fourtakas(h, drs, dr_z, drdW, arho, nrho, rho0, c0, delta, m, g);
}

}  // extern "C"
