/*
 * polymul.c -- oracle for the caller-side negacyclic product built on the
 * reference's unordered::Plan, the exact u64 schoolbook, the torus rounding
 * and the TFHE external product.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Compile with -ffp-contract=off.
 *
 * The twist / fold / torus conventions live in the reference's caller
 * (TFHE-rs, not in the tree); the reference shows the fold layout in its
 * fft128 tests (src/fft128/mod.rs:1996-2016: re = a[0..N/2], im =
 * a[N/2..N]) and the 2/N normalisation (:2033-2047).  The convention
 * restated here, and implemented by the library, is:
 *   n = N/2, zeta_j = sincospi64(j/N) -> (cos, sin)(pi j / N), j < n
 *   torus poly a (u64):  x_j = ((i64)a_j * 2^-64, (i64)a_{j+n} * 2^-64)
 *   integer poly b (i64): x_j = ((double)b_j, (double)b_{j+n})
 *   fold+twist           z_j = cmul(zeta_j, x_j)
 *   forward              unordered::Plan::fwd (src/unordered.rs:826)
 *   pointwise            Z_k = cmul(Za_k, Zb_k)    (torus operand first)
 *   inverse              unordered::Plan::inv (src/unordered.rs:927)
 *   untwist              c_j = cmul(u_j, z_j), u_j = (cos/n, -sin/n)
 *   torus rounding       from_torus(c_j.re) -> C_j, from_torus(c_j.im) -> C_{j+n}
 * with cmul the reference's scalar complex product (src/fft_simd.rs:220-233).
 * The exact oracle is the u64 wrapping schoolbook: C = A*B mod (X^N+1, 2^64).
 */
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

typedef or_c64 c64;

static inline c64 cmul(c64 ab, c64 xy) {
  c64 r;
  r.re = fma(ab.re, xy.re, -(ab.im * xy.im));
  r.im = fma(ab.re, xy.im, ab.im * xy.re);
  return r;
}

void or_twist_table(size_t N, c64 *twist) {
  size_t n = N / 2;
  for (size_t j = 0; j < n; j++) {
    double s, c;
    or_sincospi64((double)j / (double)N, &s, &c);
    twist[j].re = c;
    twist[j].im = s;
  }
}

/* from_torus: nearest multiple of 2^-64 of frac(x), as a u64.  fract in
 * [-1/2, 1/2]; the value 2^63 (fract == 1/2) wraps to the same torus element
 * 0x8000000000000000 as -2^63. */
/* CPU features for the SIMD restatement (fft64_simd.c) */
int or_cpu_has_isa(int isa) {
  __builtin_cpu_init();
  if (isa == 3) return __builtin_cpu_supports("avx2") && __builtin_cpu_supports("fma");
  if (isa == 4) return __builtin_cpu_supports("avx512f") && __builtin_cpu_supports("avx512dq");
  return 0;
}

uint64_t or_from_torus(double x) {
  double fract = x - round(x);
  double v = round(fract * 18446744073709551616.0);
  if (v >= 9223372036854775808.0) return 0x8000000000000000ull;
  return (uint64_t)(int64_t)v;
}

static const double TWO_M64 = 5.421010862427522e-20; /* 2^-64, exact */

typedef struct {
  size_t N;
  const or_uplan *plan;
  const c64 *twist;
  const c64 *untwist;
  const uint64_t *a;
  const int64_t *b;
  uint64_t *c;
  size_t lo, hi;
} pm_job;

static void polymul_one(const pm_job *J, size_t idx, c64 *za, c64 *zb, c64 *scratch) {
  size_t N = J->N, n = N / 2;
  const uint64_t *a = J->a + idx * N;
  const int64_t *b = J->b + idx * N;
  uint64_t *c = J->c + idx * N;
  for (size_t j = 0; j < n; j++) {
    c64 xa = {(double)(int64_t)a[j] * TWO_M64, (double)(int64_t)a[j + n] * TWO_M64};
    c64 xb = {(double)b[j], (double)b[j + n]};
    za[j] = cmul(J->twist[j], xa);
    zb[j] = cmul(J->twist[j], xb);
  }
  or_uplan_fwd(J->plan, za, scratch);
  or_uplan_fwd(J->plan, zb, scratch);
  for (size_t k = 0; k < n; k++) za[k] = cmul(za[k], zb[k]);
  or_uplan_inv(J->plan, za, scratch);
  for (size_t j = 0; j < n; j++) {
    c64 v = cmul(J->untwist[j], za[j]);
    c[j] = or_from_torus(v.re);
    c[j + n] = or_from_torus(v.im);
  }
}

static void *polymul_worker(void *arg) {
  pm_job *J = (pm_job *)arg;
  size_t n = J->N / 2;
  c64 *za = (c64 *)malloc(n * sizeof(c64));
  c64 *zb = (c64 *)malloc(n * sizeof(c64));
  c64 *scratch = (c64 *)malloc(J->plan->base_n * sizeof(c64));
  for (size_t i = J->lo; i < J->hi; i++) polymul_one(J, i, za, zb, scratch);
  free(za); free(zb); free(scratch);
  return NULL;
}

static void make_untwist(size_t N, const c64 *tw, c64 *u) {
  size_t n = N / 2;
  double inv_n = 1.0 / (double)n;
  for (size_t j = 0; j < n; j++) {
    u[j].re = tw[j].re * inv_n;
    u[j].im = (-tw[j].im) * inv_n;
  }
}

int or_negacyclic_polymul_f64(size_t N, int base_algo, size_t base_n, const uint64_t *a,
                              const int64_t *b, uint64_t *c, size_t batch, int threads) {
  if (N < 2 || (N & (N - 1))) return -1;
  size_t n = N / 2;
  or_uplan plan;
  int rc = or_uplan_new(&plan, n, base_algo, base_n);
  if (rc) return rc;
  c64 *tw = (c64 *)malloc(n * sizeof(c64));
  c64 *u = (c64 *)malloc(n * sizeof(c64));
  or_twist_table(N, tw);
  make_untwist(N, tw, u);
  if (threads < 1) threads = 1;
  if ((size_t)threads > batch) threads = (int)(batch ? batch : 1);
  pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)threads);
  pm_job *jobs = (pm_job *)malloc(sizeof(pm_job) * (size_t)threads);
  for (int t = 0; t < threads; t++) {
    pm_job J = {N, &plan, tw, u, a, b, c, batch * (size_t)t / (size_t)threads,
                batch * (size_t)(t + 1) / (size_t)threads};
    jobs[t] = J;
    pthread_create(&th[t], NULL, polymul_worker, &jobs[t]);
  }
  for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
  free(th); free(jobs); free(tw); free(u);
  or_uplan_free(&plan);
  return 0;
}

double or_time_polymul(size_t N, int base_algo, size_t base_n, const uint64_t *a,
                       const int64_t *b, uint64_t *c, size_t batch, int threads) {
  struct timespec t0, t1;
  clock_gettime(CLOCK_MONOTONIC, &t0);
  int rc = or_negacyclic_polymul_f64(N, base_algo, base_n, a, b, c, batch, threads);
  clock_gettime(CLOCK_MONOTONIC, &t1);
  if (rc) return -1.0;
  return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}

/* C = A*B mod (X^N + 1) with coefficients mod 2^64 (wrapping u64) */
void or_schoolbook_negacyclic_u64(size_t N, const uint64_t *a, const uint64_t *b, uint64_t *c) {
  for (size_t k = 0; k < N; k++) c[k] = 0;
  for (size_t i = 0; i < N; i++) {
    uint64_t ai = a[i];
    for (size_t j = 0; j < N - i; j++) c[i + j] += ai * b[j];
    for (size_t j = N - i; j < N; j++) c[i + j - N] -= ai * b[j];
  }
}

typedef struct {
  size_t N;
  const uint64_t *a, *b;
  uint64_t *c;
  size_t lo, hi;
} sb_job;
static void *sb_worker(void *arg) {
  sb_job *J = (sb_job *)arg;
  for (size_t i = J->lo; i < J->hi; i++)
    or_schoolbook_negacyclic_u64(J->N, J->a + i * J->N, J->b + i * J->N, J->c + i * J->N);
  return NULL;
}
void or_schoolbook_negacyclic_u64_batch(size_t N, const uint64_t *a, const uint64_t *b,
                                        uint64_t *c, size_t batch, int threads) {
  if (threads < 1) threads = 1;
  if ((size_t)threads > batch) threads = (int)(batch ? batch : 1);
  pthread_t th[256];
  sb_job jobs[256];
  if (threads > 256) threads = 256;
  for (int t = 0; t < threads; t++) {
    sb_job J = {N, a, b, c, batch * (size_t)t / (size_t)threads, batch * (size_t)(t + 1) / (size_t)threads};
    jobs[t] = J;
    pthread_create(&th[t], NULL, sb_worker, &jobs[t]);
  }
  for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
}

/* Gadget decomposition with base 2^base_log and `levels` levels (TFHE's
 * signed decomposition, defined here): round x to the nearest multiple of
 * 2^(64 - levels*base_log) (ties up), then peel balanced digits in
 * [-2^(B-1), 2^(B-1)) from the least significant level with carry.
 * out[i*N + j] is the level-i digit of x_j (level 0 = most significant,
 * weight 2^(64 - (i+1)*B)). */
void or_gadget_decompose(size_t N, const uint64_t *in, int levels, int base_log, int64_t *out) {
  int rep = levels * base_log;
  int shift = 64 - rep;
  for (size_t j = 0; j < N; j++) {
    uint64_t x = in[j];
    uint64_t st;
    if (shift == 0) st = x;
    else st = (x >> shift) + ((x >> (shift - 1)) & 1);
    if (rep < 64) st &= (((uint64_t)1 << rep) - 1);
    uint64_t mask = ((uint64_t)1 << base_log) - 1;
    int64_t half = (int64_t)1 << (base_log - 1);
    for (int i = levels - 1; i >= 0; i--) {
      int64_t d = (int64_t)(st & mask);
      st >>= base_log;
      if (d >= half) { d -= (int64_t)1 << base_log; st += 1; }
      out[(size_t)i * N + j] = d;
    }
  }
}

/* f64 negacyclic round trip of real polynomials (the uneven-batch job):
 * fold + twist + forward transform -> F (n c64, plan order), then inverse +
 * untwist (with 1/n) -> out (N f64).  x, out: count x N; F: count x n. */
int or_negacyclic_roundtrip_f64(size_t N, int base_algo, size_t base_n, const double *x,
                                c64 *F, double *out, size_t count) {
  if (N < 2 || (N & (N - 1))) return -1;
  size_t n = N / 2;
  or_uplan plan;
  int rc = or_uplan_new(&plan, n, base_algo, base_n);
  if (rc) return rc;
  c64 *tw = (c64 *)malloc(n * sizeof(c64));
  c64 *u = (c64 *)malloc(n * sizeof(c64));
  c64 *z = (c64 *)malloc(n * sizeof(c64));
  c64 *scratch = (c64 *)malloc(base_n * sizeof(c64));
  or_twist_table(N, tw);
  make_untwist(N, tw, u);
  for (size_t p = 0; p < count; p++) {
    const double *xp = x + p * N;
    for (size_t j = 0; j < n; j++) {
      c64 v = {xp[j], xp[j + n]};
      z[j] = cmul(tw[j], v);
    }
    or_uplan_fwd(&plan, z, scratch);
    for (size_t k = 0; k < n; k++) F[p * n + k] = z[k];
    or_uplan_inv(&plan, z, scratch);
    double *o = out + p * N;
    for (size_t j = 0; j < n; j++) {
      c64 v = cmul(u[j], z[j]);
      o[j] = v.re;
      o[j + n] = v.im;
    }
  }
  free(tw); free(u); free(z); free(scratch);
  or_uplan_free(&plan);
  return 0;
}

/* Torus polynomials -> Fourier domain of the unordered plan: fold,
 * (i64)x * 2^-64, twist, forward transform (the A-operand path of the
 * product above).  in: npolys x N u64, out: npolys x N/2 c64. */
int or_torus_to_fourier(size_t N, int base_algo, size_t base_n, const uint64_t *in, c64 *out,
                        size_t npolys) {
  size_t n = N / 2;
  or_uplan plan;
  int rc = or_uplan_new(&plan, n, base_algo, base_n);
  if (rc) return rc;
  c64 *tw = (c64 *)malloc(n * sizeof(c64));
  c64 *scratch = (c64 *)malloc(base_n * sizeof(c64));
  or_twist_table(N, tw);
  for (size_t p = 0; p < npolys; p++) {
    const uint64_t *a = in + p * N;
    c64 *z = out + p * n;
    for (size_t j = 0; j < n; j++) {
      c64 x = {(double)(int64_t)a[j] * TWO_M64, (double)(int64_t)a[j + n] * TWO_M64};
      z[j] = cmul(tw[j], x);
    }
    or_uplan_fwd(&plan, z, scratch);
  }
  free(tw); free(scratch);
  or_uplan_free(&plan);
  return 0;
}

/* External product (GGSW in the Fourier domain x GLWE):
 *   glwe: k1 polys of N u64 (k1 = k+1), decomposed into k1*levels digit
 *   polys (row = p*levels + i);  key_fourier: [row][col][n] c64 in the
 *   plan's permuted order;  out[col] = sum_row digit_row * key[row][col],
 *   accumulated in the Fourier domain as acc = acc + cmul(key, F(digit))
 *   from acc = 0, then inverse, untwist and torus rounding as above. */
int or_external_product(size_t N, int k1, int levels, int base_log, const c64 *key_fourier,
                        int base_algo, size_t base_n, const uint64_t *glwe, uint64_t *out) {
  size_t n = N / 2;
  or_uplan plan;
  int rc = or_uplan_new(&plan, n, base_algo, base_n);
  if (rc) return rc;
  c64 *tw = (c64 *)malloc(n * sizeof(c64));
  c64 *u = (c64 *)malloc(n * sizeof(c64));
  or_twist_table(N, tw);
  make_untwist(N, tw, u);
  int rows = k1 * levels;
  int64_t *dig = (int64_t *)malloc((size_t)rows * N * sizeof(int64_t));
  for (int p = 0; p < k1; p++) or_gadget_decompose(N, glwe + (size_t)p * N, levels, base_log, dig + (size_t)p * levels * N);
  c64 *fd = (c64 *)malloc((size_t)rows * n * sizeof(c64));
  c64 *scratch = (c64 *)malloc(base_n * sizeof(c64));
  for (int r = 0; r < rows; r++) {
    const int64_t *d = dig + (size_t)r * N;
    c64 *z = fd + (size_t)r * n;
    for (size_t j = 0; j < n; j++) {
      c64 x = {(double)d[j], (double)d[j + n]};
      z[j] = cmul(tw[j], x);
    }
    or_uplan_fwd(&plan, z, scratch);
  }
  c64 *acc = (c64 *)malloc(n * sizeof(c64));
  for (int col = 0; col < k1; col++) {
    for (size_t k = 0; k < n; k++) { acc[k].re = 0.0; acc[k].im = 0.0; }
    for (int r = 0; r < rows; r++) {
      const c64 *key = key_fourier + ((size_t)r * (size_t)k1 + (size_t)col) * n;
      const c64 *z = fd + (size_t)r * n;
      for (size_t k = 0; k < n; k++) {
        c64 pr = cmul(key[k], z[k]);
        acc[k].re = acc[k].re + pr.re;
        acc[k].im = acc[k].im + pr.im;
      }
    }
    or_uplan_inv(&plan, acc, scratch);
    uint64_t *o = out + (size_t)col * N;
    for (size_t j = 0; j < n; j++) {
      c64 v = cmul(u[j], acc[j]);
      o[j] = or_from_torus(v.re);
      o[j + n] = or_from_torus(v.im);
    }
  }
  free(acc); free(scratch); free(fd); free(dig); free(tw); free(u);
  or_uplan_free(&plan);
  return 0;
}
