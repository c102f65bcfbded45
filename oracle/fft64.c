/*
 * fft64.c -- oracle restatement of concrete-fft's f64 path.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Compile with -ffp-contract=off.
 *
 * Restates, as scalar C with the reference's operation order:
 *   src/fft_simd.rs:106-160  mul_j / mul_exp_* primitives
 *   src/fft_simd.rs:220-233  scalar complex mul (fma pattern)
 *   src/fft_simd.rs:236-296  sincospi64 (published method, SO 42792940)
 *   src/fft_simd.rs:298-321  init_wt
 *   src/dif{2,4,8,16}.rs     Stockham DIF cores, last butterflies, recursion
 *   src/dit{2,4,8,16}.rs     Stockham DIT cores, recursion
 *   src/ordered.rs:206-373   get_fn_ptr / Plan::new / fwd / inv
 *   src/unordered.rs:25-489  depth butterflies, process_x{2,4,8},
 *                            init_twiddles, fwd_depth, inv_depth
 *   src/unordered.rs:659-900 Plan::new, fwd_monomial
 *   src/unordered.rs:1040-1059 bit_rev, bit_rev_twice, bit_rev_twice_inv
 * The scalar backend is the reference's `Scalar` FftSimd (lane count 1), so
 * every Stockham pass uses the `stockham_core_generic` form.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef or_c64 c64;

/* cos(-pi/8), sin(-pi/8): src/fft_simd.rs:49-52 */
static const double H1X = 0.9238795325112867;
static const double H1Y = -0.38268343236508984;
static const double FRAC_1_SQRT_2 = 0.7071067811865476;

/* ---- primitives (src/fft_simd.rs:106-233) ---- */
static inline c64 cadd(c64 a, c64 b) { c64 r = {a.re + b.re, a.im + b.im}; return r; }
static inline c64 csub(c64 a, c64 b) { c64 r = {a.re - b.re, a.im - b.im}; return r; }
/* Scalar::mul, src/fft_simd.rs:220-233: re = fma(a, x, -b*y), im = fma(a, y, b*x) */
static inline c64 cmul(c64 ab, c64 xy) {
  c64 r;
  r.re = fma(ab.re, xy.re, -(ab.im * xy.im));
  r.im = fma(ab.re, xy.im, ab.im * xy.re);
  return r;
}
/* FftSimdExt::mul_j, src/fft_simd.rs:113-120 (conj = sign flip of im) */
static inline c64 mul_j(int fwd, c64 z) {
  c64 r;
  if (fwd) { r.re = -z.im; r.im = z.re; }   /* swap_re_im(conj(z)) */
  else     { r.re = z.im;  r.im = -z.re; }  /* conj(swap_re_im(z)) */
  return r;
}
/* src/fft_simd.rs:122-130 */
static inline c64 mul_exp_pi_over_8(int fwd, c64 z) {
  c64 t = cadd(z, mul_j(fwd, z));
  c64 r = {FRAC_1_SQRT_2 * t.re, FRAC_1_SQRT_2 * t.im};
  return r;
}
static inline c64 mul_exp_neg_pi_over_8(int fwd, c64 z) { return mul_exp_pi_over_8(!fwd, z); }
/* src/fft_simd.rs:132-159 */
static inline c64 mul_exp_pi_over_16(int fwd, c64 z) {
  c64 w = {H1X, fwd ? H1Y : -H1Y};
  return cmul(w, z);
}
static inline c64 mul_exp_17pi_over_16(int fwd, c64 z) {
  c64 w = {-H1Y, fwd ? -H1X : H1X};
  return cmul(w, z);
}
static inline c64 mul_exp_neg_pi_over_16(int fwd, c64 z) { return mul_exp_pi_over_16(!fwd, z); }
static inline c64 mul_exp_neg_17pi_over_16(int fwd, c64 z) { return mul_exp_17pi_over_16(!fwd, z); }

/* ---- sincospi64: src/fft_simd.rs:236-296 (SO 42792940) ---- */
void or_sincospi64(double a, double *sp, double *cp) {
  double az = a * 0.0;
  a = (fabs(a) < 9007199254740992.0) ? a : az;
  double r = round(a + a);
  int64_t i = (int64_t)r;
  double t = fma(-0.5, r, a);
  double s = t * t;
  r = -1.0369917389758117e-4;
  r = fma(r, s, 1.9294935641298806e-3);
  r = fma(r, s, -2.5806887942825395e-2);
  r = fma(r, s, 2.3533063028328211e-1);
  r = fma(r, s, -1.3352627688538006e+0);
  r = fma(r, s, 4.0587121264167623e+0);
  r = fma(r, s, -4.9348022005446790e+0);
  double c = fma(r, s, 1.0000000000000000e+0);
  r = 4.6151442520157035e-4;
  r = fma(r, s, -7.3700183130883555e-3);
  r = fma(r, s, 8.2145868949323936e-2);
  r = fma(r, s, -5.9926452893214921e-1);
  r = fma(r, s, 2.5501640398732688e+0);
  r = fma(r, s, -5.1677127800499516e+0);
  double st = s * t;
  r = r * st;
  double sn = fma(t, 3.1415926535897931e+0, r);
  if ((i & 2) != 0) {
    sn = 0.0 - sn;
    c = 0.0 - c;
  }
  if ((i & 1) != 0) {
    double tt = 0.0 - sn;
    sn = c;
    c = tt;
  }
  if (a == floor(a)) sn = az;
  *sp = sn;
  *cp = c;
}

/* ---- init_wt: src/fft_simd.rs:298-321 ---- */
void or_init_wt(int r, size_t n, c64 *w, c64 *w_inv) {
  if (n < (size_t)r) return;
  size_t nr = n / (size_t)r;
  double theta = -2.0 / (double)n;
  for (size_t i = 0; i < 2 * n; i++) {
    w[i].re = NAN; w[i].im = NAN;
  }
  for (size_t p = 0; p < nr; p++) {
    for (size_t k = 1; k < (size_t)r; k++) {
      double s, c;
      or_sincospi64(theta * (double)(k * p), &s, &c);
      c64 z = {c, s}, zc = {c, -s};
      w[p + k * nr] = z;
      w[n + (size_t)r * p + k] = z;
      w_inv[p + k * nr] = zc;
      w_inv[n + (size_t)r * p + k] = zc;
    }
  }
}

int or_algo_radix(int algo) {
  switch (algo) {
    case OR_DIF2: case OR_DIT2: return 2;
    case OR_DIF4: case OR_DIT4: return 4;
    case OR_DIF8: case OR_DIT8: return 8;
    case OR_DIF16: case OR_DIT16: return 16;
    default: return 0;
  }
}
static int is_dit(int algo) { return algo == OR_DIT2 || algo == OR_DIT4 || algo == OR_DIT8 || algo == OR_DIT16; }

/* ---- radix-r butterflies without twiddles (the `last_butterfly` of each
 *      module; the cores compute the same DAG) ---- */
/* src/dif2.rs:106-113 */
static void bf2(int fwd, c64 *v) {
  (void)fwd;
  c64 a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}
/* src/dif4.rs:196-214 */
static void bf4(int fwd, c64 *v) {
  c64 apc = cadd(v[0], v[2]);
  c64 amc = csub(v[0], v[2]);
  c64 bpd = cadd(v[1], v[3]);
  c64 jbmd = mul_j(fwd, csub(v[1], v[3]));
  v[0] = cadd(apc, bpd);
  v[1] = csub(amc, jbmd);
  v[2] = csub(apc, bpd);
  v[3] = cadd(amc, jbmd);
}
/* src/dif8.rs:326-364 */
static void bf8(int fwd, c64 *x) {
  c64 a04 = cadd(x[0], x[4]);
  c64 s04 = csub(x[0], x[4]);
  c64 a26 = cadd(x[2], x[6]);
  c64 js26 = mul_j(fwd, csub(x[2], x[6]));
  c64 a15 = cadd(x[1], x[5]);
  c64 s15 = csub(x[1], x[5]);
  c64 a37 = cadd(x[3], x[7]);
  c64 js37 = mul_j(fwd, csub(x[3], x[7]));
  c64 a04_p1_a26 = cadd(a04, a26);
  c64 s04_mj_s26 = csub(s04, js26);
  c64 a04_m1_a26 = csub(a04, a26);
  c64 s04_pj_s26 = cadd(s04, js26);
  c64 a15_p1_a37 = cadd(a15, a37);
  c64 w8_s15_mj_s37 = mul_exp_neg_pi_over_8(fwd, csub(s15, js37));
  c64 j_a15_m1_a37 = mul_j(fwd, csub(a15, a37));
  c64 v8_s15_pj_s37 = mul_exp_pi_over_8(fwd, cadd(s15, js37));
  x[0] = cadd(a04_p1_a26, a15_p1_a37);
  x[1] = cadd(s04_mj_s26, w8_s15_mj_s37);
  x[2] = csub(a04_m1_a26, j_a15_m1_a37);
  x[3] = csub(s04_pj_s26, v8_s15_pj_s37);
  x[4] = csub(a04_p1_a26, a15_p1_a37);
  x[5] = csub(s04_mj_s26, w8_s15_mj_s37);
  x[6] = cadd(a04_m1_a26, j_a15_m1_a37);
  x[7] = cadd(s04_pj_s26, v8_s15_pj_s37);
}
/* src/dif16.rs:650-779 */
static void bf16(int fwd, c64 *x) {
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

  c64 js4c = mul_j(fwd, s4c);
  c64 js6e = mul_j(fwd, s6e);
  c64 js5d = mul_j(fwd, s5d);
  c64 js7f = mul_j(fwd, s7f);

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

  c64 w8_s2amjs6e = mul_exp_neg_pi_over_8(fwd, s2amjs6e);
  c64 j_a2am1a6e = mul_j(fwd, a2am1a6e);
  c64 v8_s2apjs6e = mul_exp_pi_over_8(fwd, s2apjs6e);

  c64 A0 = cadd(a08p1a4c, a2ap1a6e);
  c64 A1 = cadd(s08mjs4c, w8_s2amjs6e);
  c64 A2 = csub(a08m1a4c, j_a2am1a6e);
  c64 A3 = csub(s08pjs4c, v8_s2apjs6e);
  c64 A4 = csub(a08p1a4c, a2ap1a6e);
  c64 A5 = csub(s08mjs4c, w8_s2amjs6e);
  c64 A6 = cadd(a08m1a4c, j_a2am1a6e);
  c64 A7 = cadd(s08pjs4c, v8_s2apjs6e);

  c64 w8_s3bmjs7f = mul_exp_neg_pi_over_8(fwd, s3bmjs7f);
  c64 j_a3bm1a7f = mul_j(fwd, a3bm1a7f);
  c64 v8_s3bpjs7f = mul_exp_pi_over_8(fwd, s3bpjs7f);

  c64 B0 = cadd(a19p1a5d, a3bp1a7f);
  c64 B1 = cadd(s19mjs5d, w8_s3bmjs7f);
  c64 B2 = csub(a19m1a5d, j_a3bm1a7f);
  c64 B3 = csub(s19pjs5d, v8_s3bpjs7f);
  c64 B4 = csub(a19p1a5d, a3bp1a7f);
  c64 B5 = csub(s19mjs5d, w8_s3bmjs7f);
  c64 B6 = cadd(a19m1a5d, j_a3bm1a7f);
  c64 B7 = cadd(s19pjs5d, v8_s3bpjs7f);

  c64 h1 = mul_exp_pi_over_16(fwd, B1);
  c64 w8 = mul_exp_neg_pi_over_8(fwd, B2);
  c64 h3 = mul_exp_17pi_over_16(fwd, B3);
  c64 j4 = mul_j(fwd, B4);
  c64 hd = mul_exp_neg_17pi_over_16(fwd, B5);
  c64 v8 = mul_exp_pi_over_8(fwd, B6);
  c64 hf = mul_exp_neg_pi_over_16(fwd, B7);

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

static void bf(int r, int fwd, c64 *v) {
  switch (r) {
    case 2: bf2(fwd, v); break;
    case 4: bf4(fwd, v); break;
    case 8: bf8(fwd, v); break;
    case 16: bf16(fwd, v); break;
    default: abort();
  }
}

static int ilog2(size_t n) {
  int l = 0;
  while (((size_t)1 << l) < n) l++;
  return l;
}

/* ---- Stockham DIF core (stockham_core_generic, e.g. src/dif4.rs:118-166):
 *      y[q + s*(r*p + k)] = w_k * butterfly_k(x[k*n/r + p*s + q]),
 *      twiddle w_k = w[r*p*s + k] (w = second half of the 2n table) ---- */
static void dif_core(int r, int fwd, size_t n, size_t s, const c64 *x, c64 *y, const c64 *w) {
  size_t nr = n / (size_t)r;
  c64 v[16];
  for (size_t p = 0; p < nr / s; p++) {
    for (size_t q = 0; q < s; q++) {
      for (int k = 0; k < r; k++) v[k] = x[(size_t)k * nr + p * s + q];
      bf(r, fwd, v);
      y[q + s * ((size_t)r * p)] = v[0];
      for (int k = 1; k < r; k++) y[q + s * ((size_t)r * p + k)] = cmul(w[(size_t)r * p * s + k], v[k]);
    }
  }
}
/* DIT core (e.g. src/dit4.rs:100-151): twiddle the inputs, then butterfly.
 * reads y[r*s*p + j*s + q], writes x[k*n/r + p*s + q] */
static void dit_core(int r, int fwd, size_t n, size_t s, c64 *x, const c64 *y, const c64 *w) {
  size_t nr = n / (size_t)r;
  c64 v[16];
  for (size_t p = 0; p < nr / s; p++) {
    for (size_t q = 0; q < s; q++) {
      v[0] = y[(size_t)r * s * p + q];
      for (int j = 1; j < r; j++) v[j] = cmul(w[(size_t)r * p * s + j], y[(size_t)r * s * p + (size_t)j * s + q]);
      bf(r, fwd, v);
      for (int k = 0; k < r; k++) x[(size_t)k * nr + p * s + q] = v[k];
    }
  }
}
/* end passes: stockham_dif{2,4,8,16}_end (src/dif4.rs:216-244) read x and
 * write y (or x in place); stockham_dit*_end (src/dit4.rs:191-218) read y
 * (or x in place) and write x.  With s*r == n the index map is identical. */
static void end_pass(int r, int fwd, size_t n, const c64 *src, c64 *dst) {
  size_t s = n / (size_t)r;
  c64 v[16];
  for (size_t q = 0; q < s; q++) {
    for (int k = 0; k < r; k++) v[k] = src[q + (size_t)k * s];
    bf(r, fwd, v);
    for (int k = 0; k < r; k++) dst[q + (size_t)k * s] = v[k];
  }
}

/* DIF recursion, e.g. src/dif4.rs:276-303: `L` = log2 of the remaining
 * sub-transform length n/s; a core pass runs while L > log2(r). */
static void dif_rec(int r, int fwd, int write_to_x, size_t n, size_t s, c64 *x, c64 *y,
                    const c64 *w, int L) {
  int lr = ilog2((size_t)r);
  if (L <= lr) {
    int re = 1 << L;
    if (write_to_x) end_pass(re, fwd, n, x, x);
    else end_pass(re, fwd, n, x, y);
  } else {
    dif_core(r, fwd, n, s, x, y, w);
    dif_rec(r, fwd, !write_to_x, n, s * (size_t)r, y, x, w, L - lr);
  }
}
/* DIT recursion, e.g. src/dit4.rs:299-318: recurse first, then the core. */
static void dit_rec(int r, int fwd, int read_from_x, size_t n, size_t s, c64 *x, c64 *y,
                    const c64 *w, int L) {
  int lr = ilog2((size_t)r);
  if (L <= lr) {
    int re = 1 << L;
    if (read_from_x) end_pass(re, fwd, n, x, x);
    else end_pass(re, fwd, n, y, x);
  } else {
    dit_rec(r, fwd, !read_from_x, n, s * (size_t)r, y, x, w, L - lr);
    dit_core(r, fwd, n, s, x, y, w);
  }
}

/* fn_ptr body, src/lib.rs:160-226: recursion entry with s = 1, buf as x,
 * scratch as y; `w` is the second half of the 2n table (src/ordered.rs:342-347) */
static void ordered_run(int algo, int fwd, size_t n, c64 *buf, c64 *scratch, const c64 *tw) {
  if (n <= 1) return; /* do_nothing, src/ordered.rs:204 */
  int r = or_algo_radix(algo);
  const c64 *w = tw + n;
  int L = ilog2(n);
  if (is_dit(algo)) dit_rec(r, fwd, 1, n, 1, buf, scratch, w, L);
  else dif_rec(r, fwd, 1, n, 1, buf, scratch, w, L);
}

/* Dif2/Dit2 use the same fn pointer for fwd and inv (src/dif2.rs:187-205);
 * their butterfly has no mul_j, so the fwd flag is irrelevant. */
void or_ordered_fwd(int algo, size_t n, c64 *buf, c64 *scratch, const c64 *tw) {
  ordered_run(algo, 1, n, buf, scratch, tw);
}
void or_ordered_inv(int algo, size_t n, c64 *buf, c64 *scratch, const c64 *tw_inv) {
  ordered_run(algo, 0, n, buf, scratch, tw_inv);
}

/* ---- unordered depth butterflies, src/unordered.rs:25-219 ---- */
static void fwd_butterfly(int r, c64 *z, const c64 *w /* r-1 */) {
  /* fwd_butterfly_x{2,4,8}: the DIF core DAG with fwd = true */
  bf(r, 1, z);
  for (int k = 1; k < r; k++) z[k] = cmul(w[k - 1], z[k]);
}
static void inv_butterfly(int r, c64 *z, const c64 *w) {
  /* inv_butterfly_x{2,4,8}: twiddle then butterfly with fwd = false */
  for (int k = 1; k < r; k++) z[k] = cmul(w[k - 1], z[k]);
  bf(r, 0, z);
}
static int bitrev_small(int r, int k) {
  int lr = ilog2((size_t)r), o = 0;
  for (int b = 0; b < lr; b++) o |= ((k >> b) & 1) << (lr - 1 - b);
  return o;
}
/* fwd_process_x{2,4,8}, src/unordered.rs:222-277: in place, output k of the
 * butterfly is stored in chunk bitrev(k) */
static void fwd_process(int r, c64 *z, size_t n, const c64 *w) {
  size_t m = n / (size_t)r;
  c64 v[8];
  for (size_t j = 0; j < m; j++) {
    for (int k = 0; k < r; k++) v[k] = z[(size_t)k * m + j];
    fwd_butterfly(r, v, w + (size_t)(r - 1) * j);
    for (int k = 0; k < r; k++) z[(size_t)bitrev_small(r, k) * m + j] = v[k];
  }
}
/* inv_process_x{2,4,8}, src/unordered.rs:233-293: inputs read from chunk
 * bitrev(k), outputs stored in natural order */
static void inv_process(int r, c64 *z, size_t n, const c64 *w) {
  size_t m = n / (size_t)r;
  c64 v[8];
  for (size_t j = 0; j < m; j++) {
    for (int k = 0; k < r; k++) v[k] = z[(size_t)bitrev_small(r, k) * m + j];
    inv_butterfly(r, v, w + (size_t)(r - 1) * j);
    for (int k = 0; k < r; k++) z[(size_t)k * m + j] = v[k];
  }
}

static int depth_radix(size_t n, size_t base_n) {
  /* src/unordered.rs:361-367, :407-413 */
  if (n == 2 * base_n) return 2;
  if (n == 4 * base_n) return 4;
  return 8;
}

/* init_twiddles, src/unordered.rs:349-389, complex_per_reg = 1 (Scalar) */
static void init_twiddles(size_t n, size_t base_n, int base_r, c64 *w, size_t wlen, c64 *w_inv,
                          size_t wilen) {
  double theta = 2.0 / (double)n;
  if (n <= base_n) {
    or_init_wt(base_r, n, w, w_inv);
  } else {
    int r = depth_radix(n, base_n);
    size_t m = n / (size_t)r;
    size_t head = (size_t)(r - 1) * m;
    c64 *w_next = w + head;
    c64 *w_inv_here = w_inv + (wilen - head);
    for (size_t p = 0; p < m; p++) {
      for (int k = 1; k < r; k++) {
        double sk, ck;
        or_sincospi64(theta * (double)((size_t)k * p), &sk, &ck);
        size_t idx = (size_t)(r - 1) * p + (size_t)(k - 1);
        w[idx].re = ck; w[idx].im = -sk;
        w_inv_here[idx].re = ck; w_inv_here[idx].im = sk;
      }
    }
    init_twiddles(n / (size_t)r, base_n, base_r, w_next, wlen - head, w_inv, wilen - head);
  }
}

/* fwd_depth, src/unordered.rs:391-439 */
static void fwd_depth(c64 *z, size_t n, const c64 *w, int base_algo, size_t base_n, c64 *scratch) {
  if (n == base_n) {
    ordered_run(base_algo, 1, n, z, scratch, w);
  } else {
    int r = depth_radix(n, base_n);
    size_t m = n / (size_t)r;
    fwd_process(r, z, n, w);
    for (int c = 0; c < r; c++) fwd_depth(z + (size_t)c * m, m, w + (size_t)(r - 1) * m, base_algo, base_n, scratch);
  }
}
/* inv_depth, src/unordered.rs:442-489 */
static void inv_depth(c64 *z, size_t n, const c64 *w, size_t wlen, int base_algo, size_t base_n,
                      c64 *scratch) {
  if (n == base_n) {
    ordered_run(base_algo, 0, n, z, scratch, w);
  } else {
    int r = depth_radix(n, base_n);
    size_t m = n / (size_t)r;
    size_t tail = (size_t)(r - 1) * m;
    for (int c = 0; c < r; c++) inv_depth(z + (size_t)c * m, m, w, wlen - tail, base_algo, base_n, scratch);
    inv_process(r, z, n, w + (wlen - tail));
  }
}

/* ---- permutation, src/unordered.rs:1040-1059 ---- */
static size_t bit_rev(unsigned nbits, size_t i) {
  size_t o = 0;
  for (unsigned b = 0; b < nbits; b++) o |= ((i >> b) & 1) << (nbits - 1 - b);
  return o;
}
size_t or_bit_rev_twice(unsigned nbits, unsigned base_nbits, size_t i) {
  size_t i_rev = bit_rev(nbits, i);
  size_t bottom_mask = ((size_t)1 << base_nbits) - 1;
  size_t bottom_bits = bit_rev(base_nbits, i_rev);
  return (i_rev & ~bottom_mask) | bottom_bits;
}
size_t or_bit_rev_twice_inv(unsigned nbits, unsigned base_nbits, size_t i) {
  size_t bottom_mask = ((size_t)1 << base_nbits) - 1;
  size_t bottom_bits = bit_rev(base_nbits, i);
  size_t i_rev = (i & ~bottom_mask) | bottom_bits;
  return bit_rev(nbits, i_rev);
}

/* ---- unordered::Plan::new, src/unordered.rs:659-747 (UserProvided) ---- */
int or_uplan_new(or_uplan *p, size_t n, int base_algo, size_t base_n) {
  memset(p, 0, sizeof(*p));
  if (n == 0 || (n & (n - 1)) != 0) return -1;
  if (base_n == 0 || (base_n & (base_n - 1)) != 0) return -2;
  if (base_n > n) return -3;
  if (base_n != n && base_n < 32) return -4;
  if (ilog2(base_n) > 10) return -5;
  if (or_algo_radix(base_algo) == 0) return -6;
  p->n = n; p->base_n = base_n; p->base_algo = base_algo;
  size_t tl = n + base_n;
  p->tw = (c64 *)malloc(tl * sizeof(c64));
  p->tw_inv = (c64 *)malloc(tl * sizeof(c64));
  p->mono = (c64 *)malloc(n * sizeof(c64));
  p->indices = (size_t *)malloc(n * sizeof(size_t));
  for (size_t i = 0; i < tl; i++) { p->tw[i].re = p->tw[i].im = NAN; p->tw_inv[i] = p->tw[i]; }
  init_twiddles(n, base_n, or_algo_radix(base_algo), p->tw, tl, p->tw_inv, tl);
  double theta = -2.0 / (double)n;
  for (size_t i = 0; i < n; i++) {
    double s, c;
    or_sincospi64(theta * (double)i, &s, &c);
    p->mono[i].re = c; p->mono[i].im = s;
  }
  unsigned nbits = (unsigned)ilog2(n), bbits = (unsigned)ilog2(base_n);
  for (size_t i = 0; i < n; i++) p->indices[i] = or_bit_rev_twice_inv(nbits, bbits, i);
  return 0;
}
void or_uplan_free(or_uplan *p) {
  free(p->tw); free(p->tw_inv); free(p->mono); free(p->indices);
  memset(p, 0, sizeof(*p));
}
/* Plan::fwd / Plan::inv, src/unordered.rs:826-839, :927-940 */
void or_uplan_fwd(const or_uplan *p, c64 *buf, c64 *scratch) {
  fwd_depth(buf, p->n, p->tw, p->base_algo, p->base_n, scratch);
}
void or_uplan_inv(const or_uplan *p, c64 *buf, c64 *scratch) {
  inv_depth(buf, p->n, p->tw_inv, p->n + p->base_n, p->base_algo, p->base_n, scratch);
}
/* Plan::fwd_monomial, src/unordered.rs:844-900 */
void or_uplan_fwd_monomial(const or_uplan *p, size_t degree, c64 *buf) {
  size_t n = p->n, mask = n - 1;
  size_t ratio = n / p->base_n;
  if (ratio == 1) {
    for (size_t i = 0; i < n; i++) buf[i] = p->mono[(i * degree) & mask];
  } else if (ratio == 2) {
    for (size_t i = 0; i < n / 2; i++) {
      buf[i] = p->mono[((2 * i) * degree) & mask];
      buf[n / 2 + i] = p->mono[((2 * i + 1) * degree) & mask];
    }
  } else {
    for (size_t i = 0; i < n; i++) buf[i] = p->mono[(p->indices[i] * degree) & mask];
  }
}

/* ---- batch helpers ---- */
int or_unordered_fwd_batch(size_t n, int base_algo, size_t base_n, c64 *buf, size_t batch) {
  or_uplan p;
  int rc = or_uplan_new(&p, n, base_algo, base_n);
  if (rc) return rc;
  c64 *scratch = (c64 *)malloc(base_n * sizeof(c64));
  for (size_t b = 0; b < batch; b++) or_uplan_fwd(&p, buf + b * n, scratch);
  free(scratch);
  or_uplan_free(&p);
  return 0;
}
int or_unordered_inv_batch(size_t n, int base_algo, size_t base_n, c64 *buf, size_t batch) {
  or_uplan p;
  int rc = or_uplan_new(&p, n, base_algo, base_n);
  if (rc) return rc;
  c64 *scratch = (c64 *)malloc(base_n * sizeof(c64));
  for (size_t b = 0; b < batch; b++) or_uplan_inv(&p, buf + b * n, scratch);
  free(scratch);
  or_uplan_free(&p);
  return 0;
}
static int ordered_batch_max(size_t n, int algo, c64 *buf, size_t batch, int fwd, size_t nmax) {
  if (n == 0 || (n & (n - 1)) != 0 || n > nmax || or_algo_radix(algo) == 0) return -1;
  c64 *tw = (c64 *)calloc(2 * n, sizeof(c64));
  c64 *twi = (c64 *)calloc(2 * n, sizeof(c64));
  c64 *scratch = (c64 *)malloc(n * sizeof(c64));
  or_init_wt(or_algo_radix(algo), n, tw, twi);
  for (size_t b = 0; b < batch; b++) {
    if (fwd) or_ordered_fwd(algo, n, buf + b * n, scratch, tw);
    else or_ordered_inv(algo, n, buf + b * n, scratch, twi);
  }
  free(tw); free(twi); free(scratch);
  return 0;
}
/* Plan::new asserts n < 2^11 (src/ordered.rs:243-244) */
int or_ordered_fwd_batch(size_t n, int algo, c64 *buf, size_t batch) { return ordered_batch_max(n, algo, buf, batch, 1, 1024); }
int or_ordered_inv_batch(size_t n, int algo, c64 *buf, size_t batch) { return ordered_batch_max(n, algo, buf, batch, 0, 1024); }
/* The same recursion and twiddle tables without that assert, up to n = 2^16
 * (the plan's documented range, src/ordered.rs:185).  No output of the
 * reference pins these sizes: they are checked against the closed-form DFT
 * and by round trips. */
int or_ordered_fwd_batch_large(size_t n, int algo, c64 *buf, size_t batch) { return ordered_batch_max(n, algo, buf, batch, 1, 65536); }
int or_ordered_inv_batch_large(size_t n, int algo, c64 *buf, size_t batch) { return ordered_batch_max(n, algo, buf, batch, 0, 65536); }
