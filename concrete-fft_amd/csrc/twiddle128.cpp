// twiddle128.cpp -- double-double twiddles of the 128-bit negacyclic plan.
//
// Same table as init_negacyclic_twiddles (src/fft128/mod.rs:1804-1828):
// entry m + i (1 <= m < n, i < m) holds exp(i * pi * bitrev(2m + i, 2n) / 2n)
// for the CT stage with m blocks.  The reference evaluates it with its own
// double-double sincospi; here it is evaluated in IEEE binary128
// (libquadmath, 113-bit significand) and split into an unevaluated sum
// hi + lo of doubles, which is accurate to about 2^-106 relative.
// Built with the host compiler (g++) and linked with -lquadmath.
#include <quadmath.h>

#include <cstddef>

#include "host_tables.hpp"

namespace cfft_host {

static size_t bitrev_len(size_t i, size_t len) {
  unsigned logn = 0;
  while (((size_t)1 << logn) < len) ++logn;
  size_t r = 0;
  for (unsigned k = 0; k < logn; ++k) r |= ((i >> k) & 1u) << (logn - 1 - k);
  return r;
}

static void split(__float128 v, double *hi, double *lo) {
  const double h = (double)v;
  *hi = h;
  *lo = (double)(v - (__float128)h);
}

void negacyclic_twiddles128(size_t n, double *re0, double *re1, double *im0, double *im1) {
  for (size_t i = 0; i < n; ++i) re0[i] = re1[i] = im0[i] = im1[i] = 0.0;
  for (size_t m = 1; m < n; m *= 2) {
    for (size_t i = 0; i < m; ++i) {
      const size_t k = 2 * m + i, pos = m + i;
      const __float128 x = (__float128)bitrev_len(k, 2 * n) / (__float128)(2 * n);
      __float128 s, c;
      sincosq(M_PIq * x, &s, &c);
      split(c, &re0[pos], &re1[pos]);
      split(s, &im0[pos], &im1[pos]);
    }
  }
}

}  // namespace cfft_host
