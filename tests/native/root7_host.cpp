// Host build of the pair loops' x^(1/7) (tinysph_old_amd/csrc/tsph_root7.h)
// next to libm's pow, for tests/test_root7.py.
#include <cmath>

#include "tsph_root7.h"

extern "C" {
void root7_host(const double *x, double *y, long n) {
  for (long i = 0; i < n; ++i) y[i] = tsph::root7(x[i]);
}
void pow7_host(const double *x, double *y, long n) {
  for (long i = 0; i < n; ++i) y[i] = std::pow(x[i], 1. / 7);
}
}
