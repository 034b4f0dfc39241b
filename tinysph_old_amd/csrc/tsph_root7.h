// x^(1/7) for x near 1, for the hydrostatic correction of the density
// diffusion term (src/SPH_viscouse_forces.h:52: rho0 * pow(dpH, 1/7) - rho0,
// dpH = 1 + dz rho0 g / cb).  In every configured case |dpH - 1| stays below
// 1e-3, where the general-purpose pow spends most of the FP64 pair
// arithmetic on a logarithm and an exponential in extended precision.  Here
// (1 + u)^(1/7) - 1 is summed from its binomial series
//   t = u/7 - 3u^2/49 + 13u^3/343 - 65u^4/2401 + ...   (|u| < 1/64)
// with 10 terms (truncation below 1.2e-22), the rest by Horner with fused
// multiply-adds.  The reference's exponent is the double nearest 1/7, 1/7 - d
// with d = 7.93e-18, so its pow returns x^(1/7) x^(-d); the leading term is
// therefore u * (1/7 rounded), which is u/7 - d u.  For |u| < 1/64 the error
// of t stays below ~4e-19 (2e-3 of an ulp of 1), so 1 + t is the correctly
// rounded value except within that distance of a rounding boundary.  Other
// positive arguments take two Newton steps on y^7 = z from an FP32 estimate,
// after z = x 2^(-7q) is brought into [0.5, 64), and the factor
// x^(-d) = 1 - d ln x: within about one ulp, with far fewer registers than
// pow() needs (the pair loop's occupancy depends on it).  Plain C++ so that
// the host can check it against libm.
#pragma once

#include <cmath>
#include <limits>

#ifdef __HIPCC__
#define TSPH_HOST_DEVICE __host__ __device__
#else
#define TSPH_HOST_DEVICE
#endif

namespace tsph {

constexpr double kRoot7Range = 0.015625;    // |u| below which root7_minus1 applies
constexpr double kSeventhLoss = 7.93016446160826e-18;  // 1/7 - (double)(1/7)

// (1 + u)^(1/7 - d) - 1 for |u| < kRoot7Range (d = kSeventhLoss)
TSPH_HOST_DEVICE inline double root7_minus1(double u) {
  // binom(1/7, k) for k = 2..10
  constexpr double c2 = -3. / 49, c3 = 13. / 343, c4 = -65. / 2401, c5 = 351. / 16807,
                   c6 = -1989. / 117649, c7 = 81549. / 5764801, c8 = -489294. / 40353607,
                   c9 = 2990130. / 282475249, c10 = -18538806. / 1977326743;
  double q = std::fma(u, c10, c9);
  q = std::fma(u, q, c8);
  q = std::fma(u, q, c7);
  q = std::fma(u, q, c6);
  q = std::fma(u, q, c5);
  q = std::fma(u, q, c4);
  q = std::fma(u, q, c3);
  q = std::fma(u, q, c2);
  return std::fma(u * u, q, u * (1. / 7));
}

// x^(1/7) as pow(x, 1/7): 0 for +-0, NaN for negative x and NaN, inf for inf
TSPH_HOST_DEVICE inline double root7(double x) {
  const double u = x - 1.;  // exact for x in [0.5, 2] (Sterbenz)
  if (std::fabs(u) < kRoot7Range) return 1. + root7_minus1(u);
  if (!(x > 0.)) return x == 0. ? 0. : std::numeric_limits<double>::quiet_NaN();
  if (x == std::numeric_limits<double>::infinity()) return x;
  int e = 0;
  const double m = std::frexp(x, &e);  // x = m 2^e, m in [0.5, 1)
  int q = e / 7, r = e - 7 * q;
  if (r < 0) {
    r += 7;
    --q;
  }
  const double z = std::ldexp(m, r);  // x = z 2^(7q), z in [0.5, 64)
  double y = (double)std::exp2(std::log2((float)z) * (1.f / 7.f));
  for (int k = 0; k < 2; ++k) {
    const double y2 = y * y, y3 = y2 * y, y6 = y3 * y3;
    y = y - (y6 * y - z) / (7. * y6);
  }
  const float lnx = (float)e * 0.693147181f + std::log((float)m);
  return std::ldexp(std::fma(y, -kSeventhLoss * (double)lnx, y), q);
}

}  // namespace tsph
