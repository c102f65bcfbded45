/*
 * fft128.c -- oracle restatement of concrete-fft's 128-bit negacyclic FFT.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Compile with -ffp-contract=off:
 * the error-free transforms below are exact only without contraction.
 *
 * Restates:
 *   src/fft128/f128_ops.rs:3-40     quick_two_sum, two_sum, two_diff, two_prod
 *   src/fft128/f128_ops.rs:276-409  add/sub (+_estimate), mul, sqr
 *   src/fft128/f128_ops.rs:411-449  div_f128_f128
 *   src/fft128/f128_ops.rs:496-575  sincospi_taylor, sincospi
 *   src/fft128/f128_ops.rs:578-619  the constant tables (PI, Taylor
 *                                   coefficients, cos/sin(k pi/16)); the
 *                                   values must be alike to reproduce the
 *                                   reference's twiddle bits
 *   src/fft128/mod.rs:285-349       cplx_add / cplx_sub / cplx_mul /
 *                                   cplx_mul_conj
 *   src/fft128/mod.rs:351-402       negacyclic_fwd_fft_scalar
 *   src/fft128/mod.rs:1104-1155     negacyclic_inv_fft_scalar
 *   src/fft128/mod.rs:1794-1828     bitreverse, init_negacyclic_twiddles
 * mul_variant 1 restates the x86 mul_f128x4/x8 (f128_ops.rs:836-841):
 * p2 = fma(a0, b1, fma(a1, b0, p2)); the x86 two_sum (magnitude-sorted
 * quick_two_sum, :631-644) returns the same exact error as two_sum, so the
 * add/sub estimates are shared.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>

typedef or_f128 f128;

static inline void quick_two_sum(double a, double b, double *s, double *e) {
  double t = a + b;
  *s = t;
  *e = b - (t - a);
}
static inline void two_sum(double a, double b, double *s, double *e) {
  double t = a + b;
  double bb = t - a;
  *s = t;
  *e = (a - (t - bb)) + (b - bb);
}
static inline void two_diff(double a, double b, double *s, double *e) {
  double t = a - b;
  double bb = t - a;
  *s = t;
  *e = (a - (t - bb)) - (b + bb);
}
static inline void two_prod(double a, double b, double *p, double *e) {
  double t = a * b;
  *p = t;
  *e = fma(a, b, -t);
}

static inline f128 mk(double a, double b) { f128 r = {a, b}; return r; }

/* add_estimate_f128_f128, f128_ops.rs:298-305 */
f128 or_f128_add_estimate(f128 a, f128 b) {
  double s, e;
  two_sum(a.hi, b.hi, &s, &e);
  e = e + (a.lo + b.lo);
  quick_two_sum(s, e, &s, &e);
  return mk(s, e);
}
/* sub_estimate_f128_f128, f128_ops.rs:346-354 */
f128 or_f128_sub_estimate(f128 a, f128 b) {
  double s, e;
  two_diff(a.hi, b.hi, &s, &e);
  e = e + a.lo;
  e = e - b.lo;
  quick_two_sum(s, e, &s, &e);
  return mk(s, e);
}
/* add_f128_f128, f128_ops.rs:308-318 */
f128 or_f128_add(f128 a, f128 b) {
  double s1, s2, t1, t2;
  two_sum(a.hi, b.hi, &s1, &s2);
  two_sum(a.lo, b.lo, &t1, &t2);
  s2 = s2 + t1;
  quick_two_sum(s1, s2, &s1, &s2);
  s2 = s2 + t2;
  quick_two_sum(s1, s2, &s1, &s2);
  return mk(s1, s2);
}
/* sub_f128_f128, f128_ops.rs:357-367 */
f128 or_f128_sub(f128 a, f128 b) {
  double s1, s2, t1, t2;
  two_diff(a.hi, b.hi, &s1, &s2);
  two_diff(a.lo, b.lo, &t1, &t2);
  s2 = s2 + t1;
  quick_two_sum(s1, s2, &s1, &s2);
  s2 = s2 + t2;
  quick_two_sum(s1, s2, &s1, &s2);
  return mk(s1, s2);
}
/* sub_f128_f64, f128_ops.rs:329-335 */
static f128 sub_f128_f64(f128 a, double b) {
  double s1, s2;
  two_diff(a.hi, b, &s1, &s2);
  s2 = s2 + a.lo;
  quick_two_sum(s1, s2, &s1, &s2);
  return mk(s1, s2);
}
/* add_f128_f64, f128_ops.rs:283-289 */
static f128 add_f128_f64(f128 a, double b) {
  double s1, s2;
  two_sum(a.hi, b, &s1, &s2);
  s2 = s2 + a.lo;
  quick_two_sum(s1, s2, &s1, &s2);
  return mk(s1, s2);
}
/* mul_f128_f128, f128_ops.rs:394-400 (scalar: no contraction) */
f128 or_f128_mul(f128 a, f128 b) {
  double p1, p2;
  two_prod(a.hi, b.hi, &p1, &p2);
  p2 = p2 + (a.hi * b.lo + a.lo * b.hi);
  quick_two_sum(p1, p2, &p1, &p2);
  return mk(p1, p2);
}
/* mul_f128x4 / mul_f128x8, f128_ops.rs:836-841 */
f128 or_f128_mul_fma(f128 a, f128 b) {
  double p1, p2;
  two_prod(a.hi, b.hi, &p1, &p2);
  p2 = fma(a.hi, b.lo, fma(a.lo, b.hi, p2));
  quick_two_sum(p1, p2, &p1, &p2);
  return mk(p1, p2);
}
/* mul_f128_f64, f128_ops.rs:379-385 */
static f128 mul_f128_f64(f128 a, double b) {
  double p1, p2;
  two_prod(a.hi, b, &p1, &p2);
  p2 = p2 + (a.lo * b);
  quick_two_sum(p1, p2, &p1, &p2);
  return mk(p1, p2);
}
/* sqr, f128_ops.rs:403-409 */
static f128 f128_sqr(f128 a) {
  double p1, p2;
  two_prod(a.hi, a.hi, &p1, &p2);
  p2 = p2 + 2.0 * (a.hi * a.lo);
  quick_two_sum(p1, p2, &p1, &p2);
  return mk(p1, p2);
}
/* div_f128_f128, f128_ops.rs:470-483 */
f128 or_f128_div(f128 a, f128 b) {
  double q1 = a.hi / b.hi;
  f128 r = or_f128_sub(a, mul_f128_f64(b, q1));
  double q2 = r.hi / b.hi;
  r = or_f128_sub(r, mul_f128_f64(b, q2));
  double q3 = r.hi / b.hi;
  double s, e;
  quick_two_sum(q1, q2, &s, &e);
  return add_f128_f64(mk(s, e), q3);
}

/* constants, f128_ops.rs:578-619 */
static const f128 F128_PI = {3.141592653589793, 1.2246467991473532e-16};
static const f128 SINPI_TAYLOR[9] = {
    {-5.16771278004997, 2.2665622825789447e-16},
    {2.5501640398773455, -7.931006345326556e-17},
    {-0.5992645293207921, 2.845026112698218e-17},
    {0.08214588661112823, -3.847292805297656e-18},
    {-0.0073704309457143504, -3.328281165603432e-19},
    {0.00046630280576761255, 1.0704561733683463e-20},
    {-2.1915353447830217e-5, 1.4648526682685598e-21},
    {7.952054001475513e-7, 1.736540361519021e-23},
    {-2.2948428997269873e-8, -7.376346207041088e-26},
};
static const f128 COSPI_TAYLOR[9] = {
    {-4.934802200544679, -3.1326477543698557e-16},
    {4.0587121264167685, -2.6602000824298645e-16},
    {-1.3352627688545895, 3.1815237892149862e-18},
    {0.2353306303588932, -1.2583065576724427e-18},
    {-0.02580689139001406, 1.170191067939226e-18},
    {0.0019295743094039231, -9.669517939986956e-20},
    {-0.0001046381049248457, -2.421206183964864e-21},
    {4.303069587032947e-6, -2.864010082936791e-22},
    {-1.3878952462213771e-7, -7.479362090417238e-24},
};
static const f128 SIN_K_PI_OVER_16[4] = {
    {0.19509032201612828, -7.991079068461731e-18},
    {0.3826834323650898, -1.0050772696461588e-17},
    {0.5555702330196022, 4.709410940561677e-17},
    {0.7071067811865476, -4.833646656726457e-17},
};
static const f128 COS_K_PI_OVER_16[4] = {
    {0.9807852804032304, 1.8546939997825006e-17},
    {0.9238795325112867, 1.7645047084336677e-17},
    {0.8314696123025452, 1.4073856984728024e-18},
    {0.7071067811865476, -4.833646656726457e-17},
};

/* sincospi_taylor, f128_ops.rs:496-512 */
static void sincospi_taylor(f128 x, f128 *s, f128 *c) {
  f128 sinc = F128_PI;
  f128 cos = mk(1.0, 0.0);
  f128 sqr = f128_sqr(x);
  f128 pow = mk(1.0, 0.0);
  for (int i = 0; i < 9; i++) {
    pow = or_f128_mul(pow, sqr);
    sinc = or_f128_add(sinc, or_f128_mul(SINPI_TAYLOR[i], pow));
    cos = or_f128_add(cos, or_f128_mul(COSPI_TAYLOR[i], pow));
  }
  *s = or_f128_mul(sinc, x);
  *c = cos;
}

static inline f128 neg(f128 a) { return mk(-a.hi, -a.lo); }

/* sincospi, f128_ops.rs:514-575 (inputs in [-1, 1]) */
void or_f128_sincospi(f128 x, f128 *so, f128 *co) {
  double p = round(x.hi * 2.0);
  f128 r = sub_f128_f64(x, p * 0.5);
  double q = round(r.hi * 16.0);
  r = sub_f128_f64(r, q * (1.0 / 16.0));
  long pi = (long)p, qi = (long)q;
  long qa = qi < 0 ? -qi : qi;
  f128 sin_r, cos_r, s, c;
  sincospi_taylor(r, &sin_r, &cos_r);
  if (qi == 0) {
    s = sin_r; c = cos_r;
  } else {
    f128 u = COS_K_PI_OVER_16[qa - 1], v = SIN_K_PI_OVER_16[qa - 1];
    if (qi > 0) {
      s = or_f128_add(or_f128_mul(u, sin_r), or_f128_mul(v, cos_r));
      c = or_f128_sub(or_f128_mul(u, cos_r), or_f128_mul(v, sin_r));
    } else {
      s = or_f128_sub(or_f128_mul(u, sin_r), or_f128_mul(v, cos_r));
      c = or_f128_add(or_f128_mul(u, cos_r), or_f128_mul(v, sin_r));
    }
  }
  if (pi == 0) { *so = s; *co = c; }
  else if (pi == 1) { *so = c; *co = neg(s); }
  else if (pi == -1) { *so = neg(c); *co = s; }
  else { *so = neg(s); *co = neg(c); }
}

/* bitreverse + init_negacyclic_twiddles, mod.rs:1794-1828 */
static size_t bitreverse(size_t i, size_t n) {
  int logn = 0;
  while (((size_t)1 << logn) < n) logn++;
  size_t r = 0;
  for (int k = 0; k < logn; k++) r |= ((i >> k) & 1) << (logn - k - 1);
  return r;
}
void or_fft128_init_twiddles(size_t n, double *re0, double *re1, double *im0, double *im1) {
  for (size_t i = 0; i < n; i++) re0[i] = re1[i] = im0[i] = im1[i] = 0.0;
  for (size_t m = 1; m < n; m *= 2) {
    for (size_t i = 0; i < m; i++) {
      size_t k = 2 * m + i, pos = m + i;
      f128 th = mk((double)bitreverse(k, 2 * n) / (double)(2 * n), 0.0);
      f128 s, c;
      or_f128_sincospi(th, &s, &c);
      re0[pos] = c.hi; re1[pos] = c.lo;
      im0[pos] = s.hi; im1[pos] = s.lo;
    }
  }
}

typedef f128 (*mulfn)(f128, f128);

/* cplx_mul (mod.rs:310-326): a * b; cplx_mul_conj (:329-345): a * conj(b) */
static inline void cplx_mul(mulfn mul, f128 ar, f128 ai, f128 br, f128 bi, f128 *or_, f128 *oi) {
  f128 rr = mul(ar, br), ri = mul(ar, bi), ir = mul(ai, br), ii = mul(ai, bi);
  *or_ = or_f128_sub_estimate(rr, ii);
  *oi = or_f128_add_estimate(ir, ri);
}
static inline void cplx_mul_conj(mulfn mul, f128 ar, f128 ai, f128 br, f128 bi, f128 *or_, f128 *oi) {
  f128 rr = mul(ar, br), ri = mul(ar, bi), ir = mul(ai, br), ii = mul(ai, bi);
  *or_ = or_f128_add_estimate(rr, ii);
  *oi = or_f128_sub_estimate(ir, ri);
}

/* negacyclic_fwd_fft_scalar, mod.rs:351-402 */
void or_fft128_fwd(size_t n, double *re0, double *re1, double *im0, double *im1,
                   const double *t_re0, const double *t_re1, const double *t_im0,
                   const double *t_im1, int mul_variant) {
  mulfn mul = mul_variant ? or_f128_mul_fma : or_f128_mul;
  size_t t = n;
  for (size_t m = 1; m < n; m *= 2) {
    t /= 2;
    for (size_t i = 0; i < m; i++) {
      f128 wr = mk(t_re0[m + i], t_re1[m + i]), wi = mk(t_im0[m + i], t_im1[m + i]);
      size_t start = 2 * i * t;
      for (size_t j = start; j < start + t; j++) {
        f128 z0r = mk(re0[j], re1[j]), z0i = mk(im0[j], im1[j]);
        f128 z1r = mk(re0[j + t], re1[j + t]), z1i = mk(im0[j + t], im1[j + t]);
        f128 z1wr, z1wi;
        cplx_mul(mul, z1r, z1i, wr, wi, &z1wr, &z1wi);
        f128 ar = or_f128_add_estimate(z0r, z1wr), ai = or_f128_add_estimate(z0i, z1wi);
        f128 br = or_f128_sub_estimate(z0r, z1wr), bi = or_f128_sub_estimate(z0i, z1wi);
        re0[j] = ar.hi; re1[j] = ar.lo; im0[j] = ai.hi; im1[j] = ai.lo;
        re0[j + t] = br.hi; re1[j + t] = br.lo; im0[j + t] = bi.hi; im1[j + t] = bi.lo;
      }
    }
  }
}

/* negacyclic_inv_fft_scalar, mod.rs:1104-1155 */
void or_fft128_inv(size_t n, double *re0, double *re1, double *im0, double *im1,
                   const double *t_re0, const double *t_re1, const double *t_im0,
                   const double *t_im1, int mul_variant) {
  mulfn mul = mul_variant ? or_f128_mul_fma : or_f128_mul;
  size_t t = 1;
  for (size_t m = n; m > 1;) {
    m /= 2;
    for (size_t i = 0; i < m; i++) {
      f128 wr = mk(t_re0[m + i], t_re1[m + i]), wi = mk(t_im0[m + i], t_im1[m + i]);
      size_t start = 2 * i * t;
      for (size_t j = start; j < start + t; j++) {
        f128 z0r = mk(re0[j], re1[j]), z0i = mk(im0[j], im1[j]);
        f128 z1r = mk(re0[j + t], re1[j + t]), z1i = mk(im0[j + t], im1[j + t]);
        f128 dr = or_f128_sub_estimate(z0r, z1r), di = or_f128_sub_estimate(z0i, z1i);
        f128 ar = or_f128_add_estimate(z0r, z1r), ai = or_f128_add_estimate(z0i, z1i);
        f128 br, bi;
        cplx_mul_conj(mul, dr, di, wr, wi, &br, &bi);
        re0[j] = ar.hi; re1[j] = ar.lo; im0[j] = ai.hi; im1[j] = ai.lo;
        re0[j + t] = br.hi; re1[j + t] = br.lo; im0[j + t] = bi.hi; im1[j + t] = bi.lo;
      }
    }
    t *= 2;
  }
}
