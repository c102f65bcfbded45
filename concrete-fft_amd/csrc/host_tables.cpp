// host_tables.cpp -- plan-time tables of the f64 path, computed on the host.
//
// Twiddles are sin/cos of pi*x evaluated with the published sincospi method
// of SO answer 42792940 (the method the reference cites at
// src/fft_simd.rs:236): reduce x to t in [-1/4, 1/4] around the nearest
// half-integer, evaluate a degree-14 cosine and a degree-13 sine minimax
// polynomial in t^2 by fused Horner steps, and map by quadrant.  The
// polynomial coefficients are the published ones; they must match the
// reference's bit for bit so that twiddles -- and therefore transforms -- are
// bit-identical to it.
//
// Every twiddle of a plan of size n is an entry of one table
// T[j] = sincospi(-2 j / n), 0 <= j < n: the reference's per-pass tables
// (init_wt, src/fft_simd.rs:298-321, and init_twiddles,
// src/unordered.rs:349-389) evaluate sincospi at -2kp/len or +2kp/len with
// len | n; the argument is exact in binary floating point, sincospi is odd in
// sin and even in cos bit for bit, so those values are T[kp * n/len].
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "host_tables.hpp"

namespace cfft_host {

namespace {
// cos(pi t) and sin(pi t)/t minimax coefficients, highest degree first
// (published with the method; see the file comment).
constexpr double kCosPoly[8] = {-1.0369917389758117e-4, 1.9294935641298806e-3,
                                -2.5806887942825395e-2, 2.3533063028328211e-1,
                                -1.3352627688538006e+0, 4.0587121264167623e+0,
                                -4.9348022005446790e+0, 1.0};
constexpr double kSinPoly[6] = {4.6151442520157035e-4, -7.3700183130883555e-3,
                                8.2145868949323936e-2, -5.9926452893214921e-1,
                                2.5501640398732688e+0, -5.1677127800499516e+0};
constexpr double kPi = 3.1415926535897931e+0;
}  // namespace

void sincospi(double x, double *sin_out, double *cos_out) {
  const double zero_like = x * 0.0;  // keeps the sign; NaN for infinities
  if (!(std::fabs(x) < 9007199254740992.0)) x = zero_like;
  const double q = std::round(x + x);          // nearest multiple of 1/2, in halves
  const int64_t quadrant = (int64_t)q;
  const double t = std::fma(-0.5, q, x);       // exact reduction to [-1/4, 1/4]
  const double t2 = t * t;
  double c = kCosPoly[0];
  for (int i = 1; i < 8; ++i) c = std::fma(c, t2, kCosPoly[i]);
  double p = kSinPoly[0];
  for (int i = 1; i < 6; ++i) p = std::fma(p, t2, kSinPoly[i]);
  p = p * (t2 * t);
  double s = std::fma(t, kPi, p);
  if (quadrant & 2) {
    s = 0.0 - s;
    c = 0.0 - c;
  }
  if (quadrant & 1) {
    const double tmp = 0.0 - s;
    s = c;
    c = tmp;
  }
  if (x == std::floor(x)) s = zero_like;
  *sin_out = s;
  *cos_out = c;
}

std::vector<C64> twiddle_table(size_t n) {
  std::vector<C64> t(n);
  const double theta = -2.0 / (double)n;
  for (size_t j = 0; j < n; ++j) {
    double s, c;
    sincospi(theta * (double)j, &s, &c);
    t[j] = C64{c, s};
  }
  return t;
}

std::vector<C64> twist_table(size_t N) {
  const size_t n = N / 2;
  std::vector<C64> t(n);
  for (size_t j = 0; j < n; ++j) {
    double s, c;
    sincospi((double)j / (double)N, &s, &c);
    t[j] = C64{c, s};
  }
  return t;
}

std::vector<C64> untwist_table(size_t N) {
  std::vector<C64> t = twist_table(N);
  const double inv_n = 1.0 / (double)(N / 2);
  for (auto &z : t) z = C64{z.re * inv_n, (-z.im) * inv_n};
  return t;
}

static size_t reverse_bits(size_t v, unsigned nbits) {
  size_t r = 0;
  for (unsigned b = 0; b < nbits; ++b) r |= ((v >> b) & 1u) << (nbits - 1 - b);
  return r;
}

// bit_rev_twice / bit_rev_twice_inv, src/unordered.rs:1046-1059
size_t freq_position(unsigned nbits, unsigned base_nbits, size_t freq) {
  const size_t rev = reverse_bits(freq, nbits);
  const size_t low = ((size_t)1 << base_nbits) - 1;
  return (rev & ~low) | reverse_bits(rev & low, base_nbits);
}
size_t position_freq(unsigned nbits, unsigned base_nbits, size_t pos) {
  const size_t low = ((size_t)1 << base_nbits) - 1;
  return reverse_bits((pos & ~low) | reverse_bits(pos & low, base_nbits), nbits);
}

}  // namespace cfft_host
