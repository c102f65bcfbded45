/*
 * fft64_simd.c -- the reference's x86 SIMD f64 path restated with intrinsics,
 * for the CPU baseline.  TEST INFRASTRUCTURE ONLY (see oracle.h): bench.py's
 * cpu_baseline leg times it and tests/ check it; the library never links it.
 *
 * Compiled twice (oracle/Makefile):
 *   OR_SIMD_W = 2: pulp V3 (AVX2 + FMA), c64x2 in one __m256d  -> *_v3
 *   OR_SIMD_W = 4: pulp V4 (AVX-512F),   c64x4 in one __m512d  -> *_v4
 * Restates (src/x86.rs:4-155, src/fft_simd.rs:106-233):
 *   FftSimd::mul      aa = unpacklo(a, a), bb = unpackhi(a, a),
 *                     yx = swap_re_im(b), fmaddsub(aa, b, bb * yx)
 *                     (src/x86.rs:51-58; the twiddle is operand a)
 *   mul_j, conj       sign flips by xor, re/im swap by permute
 *   cores             stockham_core_generic over the contiguous inner index
 *                     for stride s >= W (src/dif4.rs:118-166), and, for
 *                     s < W, the outer index in the lanes with twiddles from a
 *                     transposed table and a W x W transpose on store
 *                     (stockham_core_1x2 / 1x4, src/dif4.rs:24-116;
 *                     transpose src/x86.rs:126-154)
 *   depth passes      fwd/inv_process_x{2,4,8} over the contiguous index j
 *                     (src/unordered.rs:222-293) with the twiddles transposed
 *                     per lane as init_twiddles does for complex_per_reg > 1
 *                     (src/unordered.rs:335-389)
 * Every lane performs the scalar oracle's operation sequence on the same
 * operands, so results are bit-identical to fft64.c (the reference's f64
 * output is ISA-independent, src/unordered.rs:1201-1395 asserts one vector on
 * every ISA); tests/test_oracle_simd.py checks the golden vector and the
 * scalar oracle.  DIF base algorithms only (Dif2/4/8/16; the DIT forms stay
 * scalar), n >= 32.
 */
#include "oracle.h"

#include <immintrin.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

typedef or_c64 c64;

#if OR_SIMD_W == 2
#define W 2
#define SFX(name) name##_v3
typedef __m256d vd;
static inline vd LD(const c64 *p) { return _mm256_loadu_pd((const double *)p); }
static inline void ST(c64 *p, vd v) { _mm256_storeu_pd((double *)p, v); }
static inline vd ADD(vd a, vd b) { return _mm256_add_pd(a, b); }
static inline vd SUB(vd a, vd b) { return _mm256_sub_pd(a, b); }
static inline vd MULR(vd a, double c) { return _mm256_mul_pd(a, _mm256_set1_pd(c)); }
static inline vd SPLAT(c64 z) { return _mm256_setr_pd(z.re, z.im, z.re, z.im); }
static inline vd SWAP(vd z) { return _mm256_permute_pd(z, 0x5); }
static inline vd CONJ(vd z) { return _mm256_xor_pd(z, _mm256_setr_pd(0.0, -0.0, 0.0, -0.0)); }
static inline vd CMUL(vd a, vd b) {
  const vd aa = _mm256_unpacklo_pd(a, a), bb = _mm256_unpackhi_pd(a, a);
  return _mm256_fmaddsub_pd(aa, b, _mm256_mul_pd(bb, SWAP(b)));
}
/* rows[j] = (v[0] lane j, v[1] lane j) */
static inline void TRANSPOSE(const vd *v, vd *rows) {
  rows[0] = _mm256_permute2f128_pd(v[0], v[1], 0x20);
  rows[1] = _mm256_permute2f128_pd(v[0], v[1], 0x31);
}
#elif OR_SIMD_W == 4
#define W 4
#define SFX(name) name##_v4
typedef __m512d vd;
static inline vd LD(const c64 *p) { return _mm512_loadu_pd((const double *)p); }
static inline void ST(c64 *p, vd v) { _mm512_storeu_pd((double *)p, v); }
static inline vd ADD(vd a, vd b) { return _mm512_add_pd(a, b); }
static inline vd SUB(vd a, vd b) { return _mm512_sub_pd(a, b); }
static inline vd MULR(vd a, double c) { return _mm512_mul_pd(a, _mm512_set1_pd(c)); }
static inline vd SPLAT(c64 z) {
  return _mm512_setr_pd(z.re, z.im, z.re, z.im, z.re, z.im, z.re, z.im);
}
static inline vd SWAP(vd z) { return _mm512_permute_pd(z, 0x55); }
static inline vd CONJ(vd z) {
  const __m512i m = _mm512_setr_epi64(0, (long long)0x8000000000000000ull, 0,
                                      (long long)0x8000000000000000ull, 0,
                                      (long long)0x8000000000000000ull, 0,
                                      (long long)0x8000000000000000ull);
  return _mm512_castsi512_pd(_mm512_xor_si512(_mm512_castpd_si512(z), m));
}
static inline vd CMUL(vd a, vd b) {
  const vd aa = _mm512_unpacklo_pd(a, a), bb = _mm512_unpackhi_pd(a, a);
  return _mm512_fmaddsub_pd(aa, b, _mm512_mul_pd(bb, SWAP(b)));
}
static inline void TRANSPOSE(const vd *v, vd *rows) {
  const vd t0 = _mm512_shuffle_f64x2(v[0], v[1], 0x88), t1 = _mm512_shuffle_f64x2(v[0], v[1], 0xDD);
  const vd t2 = _mm512_shuffle_f64x2(v[2], v[3], 0x88), t3 = _mm512_shuffle_f64x2(v[2], v[3], 0xDD);
  rows[0] = _mm512_shuffle_f64x2(t0, t2, 0x88);
  rows[2] = _mm512_shuffle_f64x2(t0, t2, 0xDD);
  rows[1] = _mm512_shuffle_f64x2(t1, t3, 0x88);
  rows[3] = _mm512_shuffle_f64x2(t1, t3, 0xDD);
}
#else
#error "OR_SIMD_W must be 2 or 4"
#endif

static const double H1X = 0.9238795325112867;    /* cos(-pi/8) */
static const double H1Y = -0.38268343236508984;  /* sin(-pi/8) */
static const double FRAC_1_SQRT_2 = 0.7071067811865476;

/* FftSimdExt::mul_j: swap_re_im(conj(z)) forward, conj(swap_re_im(z)) inverse */
static inline vd MULJ(int fwd, vd z) { return fwd ? SWAP(CONJ(z)) : CONJ(SWAP(z)); }
static inline vd MUL_EXP_PI_OVER_8(int fwd, vd z) { return MULR(ADD(z, MULJ(fwd, z)), FRAC_1_SQRT_2); }
static inline vd MUL_EXP_NEG_PI_OVER_8(int fwd, vd z) { return MUL_EXP_PI_OVER_8(!fwd, z); }
static inline vd MUL_EXP_PI_OVER_16(int fwd, vd z) {
  const c64 w = {H1X, fwd ? H1Y : -H1Y};
  return CMUL(SPLAT(w), z);
}
static inline vd MUL_EXP_17PI_OVER_16(int fwd, vd z) {
  const c64 w = {-H1Y, fwd ? -H1X : H1X};
  return CMUL(SPLAT(w), z);
}

/* ---- radix-r butterflies: the scalar oracle's DAGs (src/dif{2,4,8,16}.rs) ---- */
static inline void bf2(vd *v) {
  const vd a = v[0], b = v[1];
  v[0] = ADD(a, b);
  v[1] = SUB(a, b);
}
static inline void bf4(int fwd, vd *v) {
  const vd apc = ADD(v[0], v[2]), amc = SUB(v[0], v[2]);
  const vd bpd = ADD(v[1], v[3]), jbmd = MULJ(fwd, SUB(v[1], v[3]));
  v[0] = ADD(apc, bpd);
  v[1] = SUB(amc, jbmd);
  v[2] = SUB(apc, bpd);
  v[3] = ADD(amc, jbmd);
}
static inline void bf8(int fwd, vd *x) {
  const vd a04 = ADD(x[0], x[4]), s04 = SUB(x[0], x[4]);
  const vd a26 = ADD(x[2], x[6]), js26 = MULJ(fwd, SUB(x[2], x[6]));
  const vd a15 = ADD(x[1], x[5]), s15 = SUB(x[1], x[5]);
  const vd a37 = ADD(x[3], x[7]), js37 = MULJ(fwd, SUB(x[3], x[7]));
  const vd t0 = ADD(a04, a26), t4 = SUB(s04, js26), t2 = SUB(a04, a26), t6 = ADD(s04, js26);
  const vd t1 = ADD(a15, a37);
  const vd t5 = MUL_EXP_NEG_PI_OVER_8(fwd, SUB(s15, js37));
  const vd t3 = MULJ(fwd, SUB(a15, a37));
  const vd t7 = MUL_EXP_PI_OVER_8(fwd, ADD(s15, js37));
  x[0] = ADD(t0, t1);
  x[1] = ADD(t4, t5);
  x[2] = SUB(t2, t3);
  x[3] = SUB(t6, t7);
  x[4] = SUB(t0, t1);
  x[5] = SUB(t4, t5);
  x[6] = ADD(t2, t3);
  x[7] = ADD(t6, t7);
}
static inline void bf16(int fwd, vd *x) {
  const vd a08 = ADD(x[0], x[8]), s08 = SUB(x[0], x[8]);
  const vd a4c = ADD(x[4], x[12]), s4c = SUB(x[4], x[12]);
  const vd a2a = ADD(x[2], x[10]), s2a = SUB(x[2], x[10]);
  const vd a6e = ADD(x[6], x[14]), s6e = SUB(x[6], x[14]);
  const vd a19 = ADD(x[1], x[9]), s19 = SUB(x[1], x[9]);
  const vd a5d = ADD(x[5], x[13]), s5d = SUB(x[5], x[13]);
  const vd a3b = ADD(x[3], x[11]), s3b = SUB(x[3], x[11]);
  const vd a7f = ADD(x[7], x[15]), s7f = SUB(x[7], x[15]);
  const vd js4c = MULJ(fwd, s4c), js6e = MULJ(fwd, s6e), js5d = MULJ(fwd, s5d), js7f = MULJ(fwd, s7f);
  const vd e0 = ADD(a08, a4c), e1 = SUB(s08, js4c), e2 = SUB(a08, a4c), e3 = ADD(s08, js4c);
  const vd f0 = ADD(a2a, a6e), f1 = SUB(s2a, js6e), f2 = SUB(a2a, a6e), f3 = ADD(s2a, js6e);
  const vd g0 = ADD(a19, a5d), g1 = SUB(s19, js5d), g2 = SUB(a19, a5d), g3 = ADD(s19, js5d);
  const vd h0 = ADD(a3b, a7f), h1 = SUB(s3b, js7f), h2 = SUB(a3b, a7f), h3 = ADD(s3b, js7f);
  const vd wf1 = MUL_EXP_NEG_PI_OVER_8(fwd, f1), jf2 = MULJ(fwd, f2), vf3 = MUL_EXP_PI_OVER_8(fwd, f3);
  const vd A0 = ADD(e0, f0), A1 = ADD(e1, wf1), A2 = SUB(e2, jf2), A3 = SUB(e3, vf3);
  const vd A4 = SUB(e0, f0), A5 = SUB(e1, wf1), A6 = ADD(e2, jf2), A7 = ADD(e3, vf3);
  const vd wh1 = MUL_EXP_NEG_PI_OVER_8(fwd, h1), jh2 = MULJ(fwd, h2), vh3 = MUL_EXP_PI_OVER_8(fwd, h3);
  const vd B0 = ADD(g0, h0), B1 = ADD(g1, wh1), B2 = SUB(g2, jh2), B3 = SUB(g3, vh3);
  const vd B4 = SUB(g0, h0), B5 = SUB(g1, wh1), B6 = ADD(g2, jh2), B7 = ADD(g3, vh3);
  const vd m1 = MUL_EXP_PI_OVER_16(fwd, B1);
  const vd m2 = MUL_EXP_NEG_PI_OVER_8(fwd, B2);
  const vd m3 = MUL_EXP_17PI_OVER_16(fwd, B3);
  const vd m4 = MULJ(fwd, B4);
  const vd m5 = MUL_EXP_17PI_OVER_16(!fwd, B5);
  const vd m6 = MUL_EXP_PI_OVER_8(fwd, B6);
  const vd m7 = MUL_EXP_PI_OVER_16(!fwd, B7);
  x[0] = ADD(A0, B0);
  x[1] = ADD(A1, m1);
  x[2] = ADD(A2, m2);
  x[3] = ADD(A3, m3);
  x[4] = SUB(A4, m4);
  x[5] = SUB(A5, m5);
  x[6] = SUB(A6, m6);
  x[7] = SUB(A7, m7);
  x[8] = SUB(A0, B0);
  x[9] = SUB(A1, m1);
  x[10] = SUB(A2, m2);
  x[11] = SUB(A3, m3);
  x[12] = ADD(A4, m4);
  x[13] = ADD(A5, m5);
  x[14] = ADD(A6, m6);
  x[15] = ADD(A7, m7);
}
static inline void bf(int r, int fwd, vd *v) {
  switch (r) {
    case 2: bf2(v); break;
    case 4: bf4(fwd, v); break;
    case 8: bf8(fwd, v); break;
    default: bf16(fwd, v); break;
  }
}

static int ilog2z(size_t n) {
  int l = 0;
  while (((size_t)1 << l) < n) l++;
  return l;
}
static int bitrev_small(int r, int k) {
  const int lr = ilog2z((size_t)r);
  int o = 0;
  for (int b = 0; b < lr; b++) o |= ((k >> b) & 1) << (lr - 1 - b);
  return o;
}

/* ---- plan: the scalar plan's tables plus the transposed ones ---- */
#define MAX_LEVELS 8
typedef struct {
  or_uplan sp;
  int r;                           /* base radix */
  int nlev;                        /* depth levels */
  int lev_r[MAX_LEVELS];
  size_t lev_n[MAX_LEVELS];
  c64 *lev_fwd[MAX_LEVELS];        /* [k-1][j] = w[(r-1) j + k-1], forward */
  c64 *lev_inv[MAX_LEVELS];        /* same for the inverse table */
  const c64 *base_fwd, *base_inv;  /* 2 base_n tables passed to the base transform */
  /* base core passes with stride s < W: [k][p s + q] = w[r p s + k] */
  size_t small_s[4];
  int nsmall;
  c64 *small_fwd[4], *small_inv[4];
} simd_plan;

static int depth_radix(size_t n, size_t base_n) {
  if (n == 2 * base_n) return 2;
  if (n == 4 * base_n) return 4;
  return 8;
}

static c64 *transpose_depth(const c64 *w, int r, size_t m) {
  c64 *t = (c64 *)malloc((size_t)(r - 1) * m * sizeof(c64));
  for (size_t j = 0; j < m; j++)
    for (int k = 1; k < r; k++) t[(size_t)(k - 1) * m + j] = w[(size_t)(r - 1) * j + (size_t)(k - 1)];
  return t;
}
static c64 *transpose_core(const c64 *w /* 2n table + n */, int r, size_t n, size_t s) {
  const size_t nr = n / (size_t)r;
  c64 *t = (c64 *)calloc((size_t)r * nr, sizeof(c64));
  for (size_t p = 0; p < nr / s; p++)
    for (size_t q = 0; q < s; q++)
      for (int k = 1; k < r; k++) t[(size_t)k * nr + p * s + q] = w[(size_t)r * p * s + (size_t)k];
  return t;
}

static void simd_plan_free(simd_plan *P) {
  for (int i = 0; i < P->nlev; i++) { free(P->lev_fwd[i]); free(P->lev_inv[i]); }
  for (int i = 0; i < P->nsmall; i++) { free(P->small_fwd[i]); free(P->small_inv[i]); }
  or_uplan_free(&P->sp);
}

static int simd_plan_new(simd_plan *P, size_t n, int base_algo, size_t base_n) {
  memset(P, 0, sizeof(*P));
  if (base_algo != OR_DIF2 && base_algo != OR_DIF4 && base_algo != OR_DIF8 && base_algo != OR_DIF16)
    return -10;
  if (n < 32 || base_n < 32) return -11;
  /* every vector covers W points of one row: n / r of the cores and the
     end pass's n / r_end must hold W */
  {
    const int r = or_algo_radix(base_algo), lr = ilog2z((size_t)r), L = ilog2z(base_n);
    const int lend = L % lr == 0 ? lr : L % lr;
    if (base_n / (size_t)r < W || base_n >> lend < W) return -11;
  }
  int rc = or_uplan_new(&P->sp, n, base_algo, base_n);
  if (rc) return rc;
  P->r = or_algo_radix(base_algo);
  /* forward depth levels (fwd_depth, src/unordered.rs:391-439) */
  const c64 *w = P->sp.tw;
  size_t nn = n;
  while (nn != base_n) {
    const int r = depth_radix(nn, base_n);
    const size_t m = nn / (size_t)r;
    P->lev_r[P->nlev] = r;
    P->lev_n[P->nlev] = nn;
    P->lev_fwd[P->nlev] = transpose_depth(w, r, m);
    P->nlev++;
    w += (size_t)(r - 1) * m;
    nn = m;
  }
  P->base_fwd = w;
  /* inverse depth levels (inv_depth, src/unordered.rs:442-489) */
  size_t wlen = n + base_n;
  nn = n;
  for (int i = 0; i < P->nlev; i++) {
    const int r = P->lev_r[i];
    const size_t m = nn / (size_t)r, tail = (size_t)(r - 1) * m;
    P->lev_inv[i] = transpose_depth(P->sp.tw_inv + (wlen - tail), r, m);
    wlen -= tail;
    nn = m;
  }
  P->base_inv = P->sp.tw_inv;
  /* base core passes with s < W (dif recursion: s = 1, r, r^2, ...) */
  const int lr = ilog2z((size_t)P->r);
  int L = ilog2z(base_n);
  size_t s = 1;
  while (L > lr) {
    if (s < W) {
      P->small_s[P->nsmall] = s;
      P->small_fwd[P->nsmall] = transpose_core(P->base_fwd + base_n, P->r, base_n, s);
      P->small_inv[P->nsmall] = transpose_core(P->base_inv + base_n, P->r, base_n, s);
      P->nsmall++;
    }
    s *= (size_t)P->r;
    L -= lr;
  }
  return 0;
}

/* ---- Stockham DIF core: y[q + s (r p + k)] = w[r p s + k] * bf_k(x[k n/r + p s + q]) ---- */
static void dif_core(int r, int fwd, size_t n, size_t s, const c64 *x, c64 *y, const c64 *w,
                     const c64 *wt) {
  const size_t nr = n / (size_t)r;
  vd v[16];
  if (s >= W) {
    for (size_t p = 0; p < nr / s; p++) {
      vd tw[16];
      for (int k = 1; k < r; k++) tw[k] = SPLAT(w[(size_t)r * p * s + (size_t)k]);
      for (size_t q = 0; q < s; q += W) {
        for (int k = 0; k < r; k++) v[k] = LD(x + (size_t)k * nr + p * s + q);
        bf(r, fwd, v);
        ST(y + q + s * ((size_t)r * p), v[0]);
        for (int k = 1; k < r; k++) ST(y + q + s * ((size_t)r * p + (size_t)k), CMUL(tw[k], v[k]));
      }
    }
    return;
  }
  /* s < W: lanes hold W / s consecutive p's with s q's each */
  for (size_t i0 = 0; i0 < nr; i0 += W) {
    for (int k = 0; k < r; k++) v[k] = LD(x + (size_t)k * nr + i0);
    bf(r, fwd, v);
    for (int k = 1; k < r; k++) v[k] = CMUL(LD(wt + (size_t)k * nr + i0), v[k]);
    const size_t p0 = i0 / s;
    if (s == 1 && r >= W) {
      /* y[r p + k]: transpose blocks of W k's, store W consecutive k's per p */
      for (int kb = 0; kb < r; kb += W) {
        vd rows[W];
        TRANSPOSE(v + kb, rows);
        for (int j = 0; j < W; j++) ST(y + (size_t)r * (p0 + (size_t)j) + (size_t)kb, rows[j]);
      }
    } else {
      c64 tmp[W];
      for (int k = 0; k < r; k++) {
        ST(tmp, v[k]);
        for (int l = 0; l < W; l++) {
          const size_t p = p0 + (size_t)l / s, q = (size_t)l % s;
          y[q + s * ((size_t)r * p + (size_t)k)] = tmp[l];
        }
      }
    }
  }
}

/* end pass: s = n / r, butterflies of x[q + k s] without twiddles */
static void end_pass(int r, int fwd, size_t n, const c64 *src, c64 *dst) {
  const size_t s = n / (size_t)r;
  vd v[16];
  for (size_t q = 0; q < s; q += W) {
    for (int k = 0; k < r; k++) v[k] = LD(src + q + (size_t)k * s);
    bf(r, fwd, v);
    for (int k = 0; k < r; k++) ST(dst + q + (size_t)k * s, v[k]);
  }
}

/* DIF recursion (src/dif4.rs:276-303) */
static void dif_rec(const simd_plan *P, int fwd, int write_to_x, size_t n, size_t s, c64 *x,
                    c64 *y, const c64 *w, int L, int pass) {
  const int r = P->r, lr = ilog2z((size_t)r);
  if (L <= lr) {
    const int re = 1 << L;
    end_pass(re, fwd, n, x, write_to_x ? x : y);
  } else {
    const c64 *wt = NULL;
    if (s < W) wt = fwd ? P->small_fwd[pass] : P->small_inv[pass];
    dif_core(r, fwd, n, s, x, y, w, wt);
    dif_rec(P, fwd, !write_to_x, n, s * (size_t)r, y, x, w, L - lr, pass + 1);
  }
}
static void base_run(const simd_plan *P, int fwd, c64 *buf, c64 *scratch) {
  const size_t bn = P->sp.base_n;
  const c64 *w = (fwd ? P->base_fwd : P->base_inv) + bn;
  dif_rec(P, fwd, 1, bn, 1, buf, scratch, w, ilog2z(bn), 0);
}

/* depth passes: fwd_process_x{2,4,8} / inv_process_x{2,4,8} */
static void fwd_process(int r, c64 *z, size_t n, const c64 *wt) {
  const size_t m = n / (size_t)r;
  vd v[16];
  for (size_t j = 0; j < m; j += W) {
    for (int k = 0; k < r; k++) v[k] = LD(z + (size_t)k * m + j);
    bf(r, 1, v);
    for (int k = 1; k < r; k++) v[k] = CMUL(LD(wt + (size_t)(k - 1) * m + j), v[k]);
    for (int k = 0; k < r; k++) ST(z + (size_t)bitrev_small(r, k) * m + j, v[k]);
  }
}
static void inv_process(int r, c64 *z, size_t n, const c64 *wt) {
  const size_t m = n / (size_t)r;
  vd v[16];
  for (size_t j = 0; j < m; j += W) {
    for (int k = 0; k < r; k++) v[k] = LD(z + (size_t)bitrev_small(r, k) * m + j);
    for (int k = 1; k < r; k++) v[k] = CMUL(LD(wt + (size_t)(k - 1) * m + j), v[k]);
    bf(r, 0, v);
    for (int k = 0; k < r; k++) ST(z + (size_t)k * m + j, v[k]);
  }
}
static void fwd_depth(const simd_plan *P, c64 *z, int lev, c64 *scratch) {
  if (lev == P->nlev) {
    base_run(P, 1, z, scratch);
    return;
  }
  const int r = P->lev_r[lev];
  const size_t nn = P->lev_n[lev], m = nn / (size_t)r;
  fwd_process(r, z, nn, P->lev_fwd[lev]);
  for (int c = 0; c < r; c++) fwd_depth(P, z + (size_t)c * m, lev + 1, scratch);
}
static void inv_depth(const simd_plan *P, c64 *z, int lev, c64 *scratch) {
  if (lev == P->nlev) {
    base_run(P, 0, z, scratch);
    return;
  }
  const int r = P->lev_r[lev];
  const size_t nn = P->lev_n[lev], m = nn / (size_t)r;
  for (int c = 0; c < r; c++) inv_depth(P, z + (size_t)c * m, lev + 1, scratch);
  inv_process(r, z, nn, P->lev_inv[lev]);
}

/* ---- batched transforms (the golden-vector and scalar-oracle checks) ---- */
static int unordered_batch(size_t n, int base_algo, size_t base_n, c64 *buf, size_t batch, int fwd) {
  simd_plan P;
  int rc = simd_plan_new(&P, n, base_algo, base_n);
  if (rc) return rc;
  c64 *scratch = (c64 *)malloc(base_n * sizeof(c64));
  for (size_t b = 0; b < batch; b++) {
    if (fwd) fwd_depth(&P, buf + b * n, 0, scratch);
    else inv_depth(&P, buf + b * n, 0, scratch);
  }
  free(scratch);
  simd_plan_free(&P);
  return 0;
}
int SFX(or_simd_unordered_fwd_batch)(size_t n, int base_algo, size_t base_n, c64 *buf, size_t batch) {
  return unordered_batch(n, base_algo, base_n, buf, batch, 1);
}
int SFX(or_simd_unordered_inv_batch)(size_t n, int base_algo, size_t base_n, c64 *buf, size_t batch) {
  return unordered_batch(n, base_algo, base_n, buf, batch, 0);
}

/* ---- negacyclic product (the caller-side convention of polymul.c) ---- */
static const double TWO_M64 = 5.421010862427522e-20; /* 2^-64, exact */

/* 2 W torus values -> u64 (the device kernels' rounding formula, equal to
 * or_from_torus for every finite input: tests/test_oracle_products.py) */
#if W == 2
static inline void from_torus_store(vd v, uint64_t *lo_dst, uint64_t *hi_dst) {
  /* (re0, im0, re1, im1) -> (re0, re1, im0, im1) */
  const __m256d x = _mm256_permute4x64_pd(v, 0xD8);
  const __m256d sign = _mm256_set1_pd(-0.0);
  const __m256d t = _mm256_sub_pd(x, _mm256_round_pd(x, _MM_FROUND_TO_NEAREST_INT | _MM_FROUND_NO_EXC));
  const __m256d T = _mm256_mul_pd(t, _mm256_set1_pd(18446744073709551616.0));
  const __m256d h = _mm256_or_pd(_mm256_and_pd(T, sign), _mm256_set1_pd(0.49999999999999994));
  const __m256d Tr = _mm256_round_pd(_mm256_add_pd(T, h), _MM_FROUND_TO_ZERO | _MM_FROUND_NO_EXC);
  const __m256d hi = _mm256_round_pd(_mm256_mul_pd(Tr, _mm256_set1_pd(2.3283064365386963e-10)),
                                     _MM_FROUND_TO_NEG_INF | _MM_FROUND_NO_EXC);
  const __m256d lo = _mm256_fmadd_pd(hi, _mm256_set1_pd(-4294967296.0), Tr);
  const __m256i hb = _mm256_castpd_si256(_mm256_add_pd(hi, _mm256_set1_pd(6755399441055744.0)));
  const __m256i lb = _mm256_castpd_si256(_mm256_add_pd(lo, _mm256_set1_pd(4503599627370496.0)));
  const __m256i u = _mm256_or_si256(_mm256_slli_epi64(hb, 32),
                                    _mm256_and_si256(lb, _mm256_set1_epi64x(0xFFFFFFFFll)));
  _mm_storeu_si128((__m128i *)lo_dst, _mm256_castsi256_si128(u));
  _mm_storeu_si128((__m128i *)hi_dst, _mm256_extracti128_si256(u, 1));
}
#else
static inline void from_torus_store(vd v, uint64_t *lo_dst, uint64_t *hi_dst) {
  const __m512d x = _mm512_permutexvar_pd(_mm512_setr_epi64(0, 2, 4, 6, 1, 3, 5, 7), v);
  const __m512i sign = _mm512_set1_epi64((long long)0x8000000000000000ull);
  const __m512d t = _mm512_sub_pd(x, _mm512_roundscale_pd(x, _MM_FROUND_TO_NEAREST_INT | _MM_FROUND_NO_EXC));
  const __m512d T = _mm512_mul_pd(t, _mm512_set1_pd(18446744073709551616.0));
  const __m512d h = _mm512_castsi512_pd(_mm512_or_si512(
      _mm512_and_si512(_mm512_castpd_si512(T), sign),
      _mm512_castpd_si512(_mm512_set1_pd(0.49999999999999994))));
  const __m512d Tr = _mm512_roundscale_pd(_mm512_add_pd(T, h), _MM_FROUND_TO_ZERO | _MM_FROUND_NO_EXC);
  const __m512d hi = _mm512_roundscale_pd(_mm512_mul_pd(Tr, _mm512_set1_pd(2.3283064365386963e-10)),
                                          _MM_FROUND_TO_NEG_INF | _MM_FROUND_NO_EXC);
  const __m512d lo = _mm512_fmadd_pd(hi, _mm512_set1_pd(-4294967296.0), Tr);
  const __m512i hb = _mm512_castpd_si512(_mm512_add_pd(hi, _mm512_set1_pd(6755399441055744.0)));
  const __m512i lb = _mm512_castpd_si512(_mm512_add_pd(lo, _mm512_set1_pd(4503599627370496.0)));
  const __m512i u = _mm512_or_si512(_mm512_slli_epi64(hb, 32),
                                    _mm512_and_si512(lb, _mm512_set1_epi64(0xFFFFFFFFll)));
  _mm256_storeu_si256((__m256i *)lo_dst, _mm512_castsi512_si256(u));
  _mm256_storeu_si256((__m256i *)hi_dst, _mm512_extracti64x4_epi64(u, 1));
}
#endif

typedef struct {
  size_t N;
  const simd_plan *plan;
  const c64 *twist, *untwist;
  const uint64_t *a;
  const int64_t *b;
  uint64_t *c;
  size_t lo, hi;
} pm_job;

static void polymul_one(const pm_job *J, size_t idx, c64 *za, c64 *zb, c64 *scratch) {
  const size_t N = J->N, n = N / 2;
  const uint64_t *a = J->a + idx * N;
  const int64_t *b = J->b + idx * N;
  uint64_t *c = J->c + idx * N;
  for (size_t j = 0; j < n; j++) {
    za[j].re = (double)(int64_t)a[j] * TWO_M64;
    za[j].im = (double)(int64_t)a[j + n] * TWO_M64;
    zb[j].re = (double)b[j];
    zb[j].im = (double)b[j + n];
  }
  for (size_t j = 0; j < n; j += W) {
    const vd z = LD(J->twist + j);
    ST(za + j, CMUL(z, LD(za + j)));
    ST(zb + j, CMUL(z, LD(zb + j)));
  }
  fwd_depth(J->plan, za, 0, scratch);
  fwd_depth(J->plan, zb, 0, scratch);
  for (size_t k = 0; k < n; k += W) ST(za + k, CMUL(LD(za + k), LD(zb + k)));
  inv_depth(J->plan, za, 0, scratch);
  for (size_t j = 0; j < n; j += W) from_torus_store(CMUL(LD(J->untwist + j), LD(za + j)), c + j, c + j + n);
}

static void *polymul_worker(void *arg) {
  const pm_job *J = (const pm_job *)arg;
  const size_t n = J->N / 2;
  c64 *za = (c64 *)aligned_alloc(64, n * sizeof(c64));
  c64 *zb = (c64 *)aligned_alloc(64, n * sizeof(c64));
  c64 *scratch = (c64 *)aligned_alloc(64, J->plan->sp.base_n * sizeof(c64));
  for (size_t i = J->lo; i < J->hi; i++) polymul_one(J, i, za, zb, scratch);
  free(za);
  free(zb);
  free(scratch);
  return NULL;
}

int SFX(or_simd_polymul_f64)(size_t N, int base_algo, size_t base_n, const uint64_t *a,
                             const int64_t *b, uint64_t *c, size_t batch, int threads) {
  if (N < 64 || (N & (N - 1))) return -1;
  const size_t n = N / 2;
  simd_plan P;
  int rc = simd_plan_new(&P, n, base_algo, base_n);
  if (rc) return rc;
  c64 *tw = (c64 *)aligned_alloc(64, n * sizeof(c64));
  c64 *u = (c64 *)aligned_alloc(64, n * sizeof(c64));
  or_twist_table(N, tw);
  const double inv_n = 1.0 / (double)n;
  for (size_t j = 0; j < n; j++) {
    u[j].re = tw[j].re * inv_n;
    u[j].im = (-tw[j].im) * inv_n;
  }
  if (threads < 1) threads = 1;
  if ((size_t)threads > batch) threads = (int)(batch ? batch : 1);
  pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)threads);
  pm_job *jobs = (pm_job *)malloc(sizeof(pm_job) * (size_t)threads);
  for (int t = 0; t < threads; t++) {
    const pm_job J = {N, &P, tw, u, a, b, c, batch * (size_t)t / (size_t)threads,
                      batch * (size_t)(t + 1) / (size_t)threads};
    jobs[t] = J;
    pthread_create(&th[t], NULL, polymul_worker, &jobs[t]);
  }
  for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
  free(th);
  free(jobs);
  free(tw);
  free(u);
  simd_plan_free(&P);
  return 0;
}

double SFX(or_simd_time_polymul)(size_t N, int base_algo, size_t base_n, const uint64_t *a,
                                 const int64_t *b, uint64_t *c, size_t batch, int threads) {
  struct timespec t0, t1;
  clock_gettime(CLOCK_MONOTONIC, &t0);
  const int rc = SFX(or_simd_polymul_f64)(N, base_algo, base_n, a, b, c, batch, threads);
  clock_gettime(CLOCK_MONOTONIC, &t1);
  if (rc) return -1.0;
  return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}
