// x^7 for the Tait EOS (src/SPH_pressure.h:7-19: B ((rho/rho0)^7 - 1), where the
// reference calls pow(rho/rho0, 7.)).  The product is formed in double-double
// arithmetic (Dekker's exact products by fused multiply-add: x^2, x^4 = (x^2)^2,
// x^6 = x^4 x^2, x^7 = x^6 x, each as a high double plus its low-order part),
// relative error ~1e-31, and rounded once, so the result is the correctly
// rounded x^7 except for exact values within that distance of a rounding
// boundary, where a correctly rounded pow (glibc's, within 0.52 ulp) also
// agrees in practice.  Outside |x| in [2^-140, 2^140] the low parts would
// underflow or the products overflow; there the plain product is returned.
// Plain C++ so that the host can check it against libm.
#pragma once

#include <cmath>

#ifdef __HIPCC__
#define TSPH_HOST_DEVICE_POW7 __host__ __device__
#else
#define TSPH_HOST_DEVICE_POW7
#endif

namespace tsph {

TSPH_HOST_DEVICE_POW7 inline double pow7(double x) {
  const double ax = std::fabs(x);
  if (!(ax >= 0x1p-140 && ax <= 0x1p140)) {  // zero, tiny, huge, inf, NaN
    const double x2 = x * x;
    return x2 * x2 * x2 * x;
  }
  // x^2 = h2 + l2 exactly
  const double h2 = x * x;
  const double l2 = std::fma(x, x, -h2);
  // x^4 = (h2 + l2)^2 ~ h2^2 + 2 h2 l2
  const double h4 = h2 * h2;
  const double l4 = std::fma(h2, h2, -h4) + 2. * h2 * l2;
  // x^6 = x^4 x^2 ~ h4 h2 + (h4 l2 + l4 h2)
  const double h6 = h4 * h2;
  const double l6 = std::fma(h4, h2, -h6) + (h4 * l2 + l4 * h2);
  // x^7 = x^6 x
  const double h7 = h6 * x;
  const double l7 = std::fma(h6, x, -h7) + l6 * x;
  return h7 + l7;
}

}  // namespace tsph
