// host_tables.hpp -- plan-time host tables (twiddles, twists, permutations).
#pragma once

#include <cstddef>
#include <vector>

namespace cfft_host {

struct C64 {
  double re, im;
};

void sincospi(double x, double *sin_out, double *cos_out);
// T[j] = (cos, sin)(-2 pi j / n)
std::vector<C64> twiddle_table(size_t n);
// zeta_j = (cos, sin)(pi j / N), j < N/2
std::vector<C64> twist_table(size_t N);
// conj(zeta_j) / (N/2)
std::vector<C64> untwist_table(size_t N);
// position of frequency `freq` in the permuted order (bit_rev_twice)
size_t freq_position(unsigned nbits, unsigned base_nbits, size_t freq);
// frequency held at position `pos` (bit_rev_twice_inv)
size_t position_freq(unsigned nbits, unsigned base_nbits, size_t pos);

// double-double twiddles of the 128-bit negacyclic plan
// (init_negacyclic_twiddles, src/fft128/mod.rs:1804-1828), computed in
// binary128 and split into hi/lo doubles.
void negacyclic_twiddles128(size_t n, double *re0, double *re1, double *im0, double *im1);

}  // namespace cfft_host
