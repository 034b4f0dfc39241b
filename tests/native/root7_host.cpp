// Host build of the FP64 path's powers, x^(1/7) of the pair loops
// (tinysph_old_amd/csrc/tsph_root7.h) and x^7 of the EOS (tsph_pow7.h), next
// to libm's pow, for tests/test_root7.py.
#include <cmath>

#include "tsph_pow7.h"
#include "tsph_root7.h"

extern "C" {
void root7_host(const double *x, double *y, long n) {
  for (long i = 0; i < n; ++i) y[i] = tsph::root7(x[i]);
}
void pow7_host(const double *x, double *y, long n) {
  for (long i = 0; i < n; ++i) y[i] = std::pow(x[i], 1. / 7);
}
void tsph_pow7_host(const double *x, double *y, long n) {
  for (long i = 0; i < n; ++i) y[i] = tsph::pow7(x[i]);
}
void libm_pow7_host(const double *x, double *y, long n) {
  for (long i = 0; i < n; ++i) y[i] = std::pow(x[i], 7.);
}
}
