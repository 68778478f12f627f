// TEST INFRASTRUCTURE ONLY: a C entry point over the reference's own Faddeeva
// package (src/Faddeeva.cpp, the erf behind the generated soft-FMT weights,
// src/haskell/Expression.hs:316), compiled from the reference's sources by
// `make ref` into oracle/_ref/.  tests/test_sfmt.py uses it to check the
// oracle's scipy.special.erf on complex arguments.
#include <complex>

#include "Faddeeva.hh"

extern "C" void faddeeva_ref_erf(const double *re, const double *im, long n, double *out_re, double *out_im) {
  for (long i = 0; i < n; ++i) {
    const std::complex<double> z = Faddeeva::erf(std::complex<double>(re[i], im[i]));
    out_re[i] = z.real();
    out_im[i] = z.imag();
  }
}
