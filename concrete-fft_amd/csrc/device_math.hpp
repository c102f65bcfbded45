// device_math.hpp -- gfx950 device arithmetic shared by the FFT kernels.
//
// The butterflies execute exactly the reference's f64 operation DAG
// (concrete-fft v0.5.1): the same additions, the same explicit fused
// multiply-adds and the same operand order, so a transform computed here is
// bit-identical to the reference's for the same plan.  All HIP code is built
// with -ffp-contract=off; every fused multiply-add is an explicit __fma_rn.
//
//   complex product        src/fft_simd.rs:220-233 (re = fma(a,x,-(b*y)),
//                          im = fma(a,y,b*x)); the twiddle is operand a
//   mul_j, mul_exp_*       src/fft_simd.rs:106-160
//   radix-2/4/8/16 DAGs    src/dif2.rs:106-113, src/dif4.rs:196-214,
//                          src/dif8.rs:326-364, src/dif16.rs:650-779 (the DIT
//                          cores and the unordered depth butterflies,
//                          src/unordered.rs:25-219, use the same DAGs)
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cfft_dev {

// 16-byte aligned so LDS and global accesses of one point are single
// 128-bit operations
struct alignas(16) c64 {
  double re, im;
};

__device__ __forceinline__ c64 mk(double re, double im) {
  c64 r;
  r.re = re;
  r.im = im;
  return r;
}
__device__ __forceinline__ c64 cadd(c64 a, c64 b) { return mk(a.re + b.re, a.im + b.im); }
__device__ __forceinline__ c64 csub(c64 a, c64 b) { return mk(a.re - b.re, a.im - b.im); }
// reference scalar complex product, `a` is the twiddle / left operand
__device__ __forceinline__ c64 cmul(c64 a, c64 x) {
  return mk(__fma_rn(a.re, x.re, -(a.im * x.im)), __fma_rn(a.re, x.im, a.im * x.re));
}
__device__ __forceinline__ c64 conj(c64 a) { return mk(a.re, -a.im); }

// multiply by +i (fwd) or -i (inv): swap_re_im(conj(z)) / conj(swap_re_im(z))
template <bool FWD>
__device__ __forceinline__ c64 mul_j(c64 z) {
  return FWD ? mk(-z.im, z.re) : mk(z.im, -z.re);
}

constexpr double kFrac1Sqrt2 = 0.7071067811865476;
constexpr double kH1X = 0.9238795325112867;    // cos(-pi/8)
constexpr double kH1Y = -0.38268343236508984;  // sin(-pi/8)

template <bool FWD>
__device__ __forceinline__ c64 mul_exp_pi_over_8(c64 z) {
  c64 t = cadd(z, mul_j<FWD>(z));
  return mk(kFrac1Sqrt2 * t.re, kFrac1Sqrt2 * t.im);
}
template <bool FWD>
__device__ __forceinline__ c64 mul_exp_neg_pi_over_8(c64 z) {
  return mul_exp_pi_over_8<!FWD>(z);
}
template <bool FWD>
__device__ __forceinline__ c64 mul_exp_pi_over_16(c64 z) {
  return cmul(mk(kH1X, FWD ? kH1Y : -kH1Y), z);
}
template <bool FWD>
__device__ __forceinline__ c64 mul_exp_17pi_over_16(c64 z) {
  return cmul(mk(-kH1Y, FWD ? -kH1X : kH1X), z);
}

// ---- radix-R butterflies on v[0..R), natural input and output order ----
template <bool FWD>
__device__ __forceinline__ void bf2(c64 *v) {
  c64 a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}

template <bool FWD>
__device__ __forceinline__ void bf4(c64 *v) {
  c64 apc = cadd(v[0], v[2]);
  c64 amc = csub(v[0], v[2]);
  c64 bpd = cadd(v[1], v[3]);
  c64 jbmd = mul_j<FWD>(csub(v[1], v[3]));
  v[0] = cadd(apc, bpd);
  v[1] = csub(amc, jbmd);
  v[2] = csub(apc, bpd);
  v[3] = cadd(amc, jbmd);
}

template <bool FWD>
__device__ __forceinline__ void bf8(c64 *x) {
  c64 a04 = cadd(x[0], x[4]);
  c64 s04 = csub(x[0], x[4]);
  c64 a26 = cadd(x[2], x[6]);
  c64 js26 = mul_j<FWD>(csub(x[2], x[6]));
  c64 a15 = cadd(x[1], x[5]);
  c64 s15 = csub(x[1], x[5]);
  c64 a37 = cadd(x[3], x[7]);
  c64 js37 = mul_j<FWD>(csub(x[3], x[7]));
  c64 t0 = cadd(a04, a26);
  c64 t4 = csub(s04, js26);
  c64 t2 = csub(a04, a26);
  c64 t6 = cadd(s04, js26);
  c64 t1 = cadd(a15, a37);
  c64 t5 = mul_exp_neg_pi_over_8<FWD>(csub(s15, js37));
  c64 t3 = mul_j<FWD>(csub(a15, a37));
  c64 t7 = mul_exp_pi_over_8<FWD>(cadd(s15, js37));
  x[0] = cadd(t0, t1);
  x[1] = cadd(t4, t5);
  x[2] = csub(t2, t3);
  x[3] = csub(t6, t7);
  x[4] = csub(t0, t1);
  x[5] = csub(t4, t5);
  x[6] = cadd(t2, t3);
  x[7] = cadd(t6, t7);
}

template <bool FWD>
__device__ __forceinline__ void bf16(c64 *x) {
  c64 a08 = cadd(x[0], x[8]);
  c64 s08 = csub(x[0], x[8]);
  c64 a4c = cadd(x[4], x[12]);
  c64 s4c = csub(x[4], x[12]);
  c64 a2a = cadd(x[2], x[10]);
  c64 s2a = csub(x[2], x[10]);
  c64 a6e = cadd(x[6], x[14]);
  c64 s6e = csub(x[6], x[14]);
  c64 a19 = cadd(x[1], x[9]);
  c64 s19 = csub(x[1], x[9]);
  c64 a5d = cadd(x[5], x[13]);
  c64 s5d = csub(x[5], x[13]);
  c64 a3b = cadd(x[3], x[11]);
  c64 s3b = csub(x[3], x[11]);
  c64 a7f = cadd(x[7], x[15]);
  c64 s7f = csub(x[7], x[15]);

  c64 js4c = mul_j<FWD>(s4c);
  c64 js6e = mul_j<FWD>(s6e);
  c64 js5d = mul_j<FWD>(s5d);
  c64 js7f = mul_j<FWD>(s7f);

  c64 a08p1a4c = cadd(a08, a4c);
  c64 s08mjs4c = csub(s08, js4c);
  c64 a08m1a4c = csub(a08, a4c);
  c64 s08pjs4c = cadd(s08, js4c);
  c64 a2ap1a6e = cadd(a2a, a6e);
  c64 s2amjs6e = csub(s2a, js6e);
  c64 a2am1a6e = csub(a2a, a6e);
  c64 s2apjs6e = cadd(s2a, js6e);
  c64 a19p1a5d = cadd(a19, a5d);
  c64 s19mjs5d = csub(s19, js5d);
  c64 a19m1a5d = csub(a19, a5d);
  c64 s19pjs5d = cadd(s19, js5d);
  c64 a3bp1a7f = cadd(a3b, a7f);
  c64 s3bmjs7f = csub(s3b, js7f);
  c64 a3bm1a7f = csub(a3b, a7f);
  c64 s3bpjs7f = cadd(s3b, js7f);

  c64 w8_s2amjs6e = mul_exp_neg_pi_over_8<FWD>(s2amjs6e);
  c64 j_a2am1a6e = mul_j<FWD>(a2am1a6e);
  c64 v8_s2apjs6e = mul_exp_pi_over_8<FWD>(s2apjs6e);

  c64 A0 = cadd(a08p1a4c, a2ap1a6e);
  c64 A1 = cadd(s08mjs4c, w8_s2amjs6e);
  c64 A2 = csub(a08m1a4c, j_a2am1a6e);
  c64 A3 = csub(s08pjs4c, v8_s2apjs6e);
  c64 A4 = csub(a08p1a4c, a2ap1a6e);
  c64 A5 = csub(s08mjs4c, w8_s2amjs6e);
  c64 A6 = cadd(a08m1a4c, j_a2am1a6e);
  c64 A7 = cadd(s08pjs4c, v8_s2apjs6e);

  c64 w8_s3bmjs7f = mul_exp_neg_pi_over_8<FWD>(s3bmjs7f);
  c64 j_a3bm1a7f = mul_j<FWD>(a3bm1a7f);
  c64 v8_s3bpjs7f = mul_exp_pi_over_8<FWD>(s3bpjs7f);

  c64 B0 = cadd(a19p1a5d, a3bp1a7f);
  c64 B1 = cadd(s19mjs5d, w8_s3bmjs7f);
  c64 B2 = csub(a19m1a5d, j_a3bm1a7f);
  c64 B3 = csub(s19pjs5d, v8_s3bpjs7f);
  c64 B4 = csub(a19p1a5d, a3bp1a7f);
  c64 B5 = csub(s19mjs5d, w8_s3bmjs7f);
  c64 B6 = cadd(a19m1a5d, j_a3bm1a7f);
  c64 B7 = cadd(s19pjs5d, v8_s3bpjs7f);

  c64 h1 = mul_exp_pi_over_16<FWD>(B1);
  c64 w8 = mul_exp_neg_pi_over_8<FWD>(B2);
  c64 h3 = mul_exp_17pi_over_16<FWD>(B3);
  c64 j4 = mul_j<FWD>(B4);
  c64 hd = mul_exp_17pi_over_16<!FWD>(B5);
  c64 v8 = mul_exp_pi_over_8<FWD>(B6);
  c64 hf = mul_exp_pi_over_16<!FWD>(B7);

  x[0] = cadd(A0, B0);
  x[1] = cadd(A1, h1);
  x[2] = cadd(A2, w8);
  x[3] = cadd(A3, h3);
  x[4] = csub(A4, j4);
  x[5] = csub(A5, hd);
  x[6] = csub(A6, v8);
  x[7] = csub(A7, hf);
  x[8] = csub(A0, B0);
  x[9] = csub(A1, h1);
  x[10] = csub(A2, w8);
  x[11] = csub(A3, h3);
  x[12] = cadd(A4, j4);
  x[13] = cadd(A5, hd);
  x[14] = cadd(A6, v8);
  x[15] = cadd(A7, hf);
}

template <int R, bool FWD>
__device__ __forceinline__ void butterfly(c64 *v) {
  if constexpr (R == 2) bf2<FWD>(v);
  else if constexpr (R == 4) bf4<FWD>(v);
  else if constexpr (R == 8) bf8<FWD>(v);
  else bf16<FWD>(v);
}

template <int R>
__device__ __forceinline__ constexpr int bitrev_digit(int k) {
  if constexpr (R == 2) return k;
  else if constexpr (R == 4) return ((k & 1) << 1) | ((k >> 1) & 1);
  else return ((k & 1) << 2) | (k & 2) | ((k >> 2) & 1);
}

// ---- streamed operands ----
// Polynomials are read and written once per product: nontemporal accesses
// keep them from evicting the twiddle and twist tables every product reuses
// (measured: +8 % on the N = 2048 batched product).
template <class T>
__device__ __forceinline__ T ld_stream(const T *p) {
  return __builtin_nontemporal_load(p);
}
template <class T>
__device__ __forceinline__ void st_stream(T v, T *p) {
  __builtin_nontemporal_store(v, p);
}
typedef double dvec2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ c64 ld_stream(const c64 *p) {
  const dvec2 v = __builtin_nontemporal_load(reinterpret_cast<const dvec2 *>(p));
  return mk(v.x, v.y);
}
__device__ __forceinline__ void st_stream(c64 v, c64 *p) {
  const dvec2 w = {v.re, v.im};
  __builtin_nontemporal_store(w, reinterpret_cast<dvec2 *>(p));
}

// ---- integer and torus conversions (caller-side convention) ----
// (double)v for any int64 v, correctly rounded: hi 2^32 + lo is formed
// exactly inside one fma and rounded once.
__device__ __forceinline__ double i64_to_f64(int64_t v) {
  const double hi = (double)(int32_t)(v >> 32);
  const double lo = (double)(uint32_t)v;
  return __fma_rn(hi, 4294967296.0, lo);
}
// (double)(int64)a * 2^-64: the same single rounding, scaled exactly
// (the result is a normal number for a != 0).
__device__ __forceinline__ double torus_to_f64(uint64_t a) {
  const double hi = (double)(int32_t)(a >> 32);
  const double lo = (double)(uint32_t)a * 5.421010862427522e-20;  // 2^-64, exact
  return __fma_rn(hi, 2.3283064365386963e-10, lo);                // 2^-32
}
// Nearest multiple of 2^-64 of frac(x), as a u64.  Reference formula (half
// away from zero in both roundings):
//   fract = x - round(x); v = round(fract * 2^64); (u64)(i64)v, v = 2^63 -> 2^63.
// Branch-free form with the same bits for every finite x (11 VALU operations):
//   t = x - rint(x) is exact and differs from fract only at half-integers,
//   where both are +-1/2, i.e. the same torus element 2^63;  T = t 2^64 is
//   exact, |T| <= 2^63;  trunc(T + copysign(0.5 - 2^-54, T)) rounds half away
//   from zero exactly (the sum is correctly rounded, and a value below a
//   half-integer can never round up to it);  the integer Tr, |Tr| <= 2^63,
//   splits exactly into hi = floor(Tr / 2^32) in [-2^31, 2^31] and
//   lo = Tr - hi 2^32 in [0, 2^32);  hi + 1.5 2^52 holds hi mod 2^32 (two's
//   complement, hi = 2^31 included) in its low mantissa word.
__device__ __forceinline__ uint64_t from_torus(double x) {
  const double t = x - __builtin_rint(x);
  const double T = t * 18446744073709551616.0;                                  // 2^64
  const double Tr = __builtin_trunc(T + __builtin_copysign(0.49999999999999994, T));
  const double hi = __builtin_floor(Tr * 2.3283064365386963e-10);               // 2^-32
  const double lo = __fma_rn(hi, -4294967296.0, Tr);
  const double hm = hi + 6755399441055744.0;                                    // 1.5 2^52
  return ((uint64_t)(uint32_t)__builtin_bit_cast(uint64_t, hm) << 32) | (uint64_t)(uint32_t)lo;
}

}  // namespace cfft_dev
