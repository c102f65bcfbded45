/*
 * oracle.h -- CPU restatement of zama-ai/concrete-fft (v0.5.1) used as the
 * correctness oracle and the CPU baseline.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library includes, links or
 * calls this code: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load liboracle.so, and only as the checker (or as the
 * timed CPU comparison).  Every function cites the reference file:line whose
 * arithmetic it restates.
 *
 * Floating point: this file must be compiled with -ffp-contract=off and
 * without -ffast-math; every fused multiply-add of the reference is written
 * as an explicit fma() call, every other operation is a plain IEEE-754
 * double operation in the reference's operand order.
 *
 * Parity pins (what checks each part; see tests/test_oracle_*.py):
 *   sincospi64 / init_wt / init_twiddles  -- golden vector (bit-exact) and
 *                                            mpmath within 1 ulp
 *   Dif4 + depth x8 + permutation          -- reference golden vector,
 *                                            src/unordered.rs:1176-1395
 *   other algos / base sizes               -- brute-force DFT (numpy /
 *                                            mpmath), roundtrip, cross-plan
 *                                            serde permutation
 *   fwd_monomial                           -- invariant fwd_monomial(d) ==
 *                                            fwd(e_d) (src/unordered.rs:1108)
 *   negacyclic f64 product                 -- u64 wrapping schoolbook
 *                                            (exact mod 2^64), torus bound
 *   fft128 (f128 ops, sincospi, fwd/inv)   -- mpmath 2^-104 / 2^-101 op
 *                                            bounds, f128 schoolbook < 1e-30 N
 *   u64 schoolbook                         -- brute force on tiny inputs,
 *                                            X^N = -1 shift invariant
 */
#ifndef CFFT_ORACLE_H
#define CFFT_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { double re, im; } or_c64;

/* ---- reference f64 path (src/fft_simd.rs, src/dif*.rs, src/dit*.rs,
 *      src/ordered.rs, src/unordered.rs) ---- */

/* FftAlgo discriminants, src/ordered.rs:26-45 (same order). */
enum { OR_DIF2 = 0, OR_DIT2, OR_DIF4, OR_DIT4, OR_DIF8, OR_DIT8, OR_DIF16, OR_DIT16 };

void or_sincospi64(double a, double *s, double *c);
void or_init_wt(int r, size_t n, or_c64 *w, or_c64 *w_inv);
int or_algo_radix(int algo);

/* ordered plan: tw / tw_inv are the 2n-entry tables from or_init_wt */
void or_ordered_fwd(int algo, size_t n, or_c64 *buf, or_c64 *scratch, const or_c64 *tw);
void or_ordered_inv(int algo, size_t n, or_c64 *buf, or_c64 *scratch, const or_c64 *tw_inv);

typedef struct {
  size_t n, base_n;
  int base_algo;
  or_c64 *tw, *tw_inv;   /* n + base_n each (init_twiddles layout, 1 complex per reg) */
  or_c64 *mono;          /* monomial twiddles, n */
  size_t *indices;       /* bit_rev_twice_inv, n */
} or_uplan;

int or_uplan_new(or_uplan *p, size_t n, int base_algo, size_t base_n);
void or_uplan_free(or_uplan *p);
void or_uplan_fwd(const or_uplan *p, or_c64 *buf, or_c64 *scratch);
void or_uplan_inv(const or_uplan *p, or_c64 *buf, or_c64 *scratch);
void or_uplan_fwd_monomial(const or_uplan *p, size_t degree, or_c64 *buf);
size_t or_bit_rev_twice(unsigned nbits, unsigned base_nbits, size_t i);
size_t or_bit_rev_twice_inv(unsigned nbits, unsigned base_nbits, size_t i);

/* batched helpers over contiguous transforms (used by tests and the CPU baseline) */
int or_unordered_fwd_batch(size_t n, int base_algo, size_t base_n, or_c64 *buf, size_t batch);
int or_unordered_inv_batch(size_t n, int base_algo, size_t base_n, or_c64 *buf, size_t batch);
int or_ordered_fwd_batch(size_t n, int algo, or_c64 *buf, size_t batch);
int or_ordered_inv_batch(size_t n, int algo, or_c64 *buf, size_t batch);
int or_ordered_fwd_batch_large(size_t n, int algo, or_c64 *buf, size_t batch);
int or_ordered_inv_batch_large(size_t n, int algo, or_c64 *buf, size_t batch);

/* ---- rand 0.8 StdRng (ChaCha12, rand_chacha 0.3) seeded with
 *      rand_core 0.6 seed_from_u64 (PCG32 expansion); gen_range(0.0..1.0) ---- */
typedef struct {
  uint32_t key[8];
  uint64_t counter;
  uint32_t buf[16];
  int idx;
} or_stdrng;
void or_stdrng_seed_from_u64(or_stdrng *r, uint64_t seed);
uint32_t or_stdrng_next_u32(or_stdrng *r);
uint64_t or_stdrng_next_u64(or_stdrng *r);
double or_stdrng_gen_f64_01(or_stdrng *r);
void or_stdrng_fill_f64_01(uint64_t seed, double *out, size_t count);

/* ---- negacyclic polynomial products (caller-side convention, see
 *      polymul.c header) ---- */
void or_twist_table(size_t N, or_c64 *twist);
int or_negacyclic_polymul_f64(size_t N, int base_algo, size_t base_n, const uint64_t *a,
                              const int64_t *b, uint64_t *c, size_t batch, int threads);
void or_schoolbook_negacyclic_u64(size_t N, const uint64_t *a, const uint64_t *b, uint64_t *c);
void or_schoolbook_negacyclic_u64_batch(size_t N, const uint64_t *a, const uint64_t *b,
                                        uint64_t *c, size_t batch, int threads);
uint64_t or_from_torus(double x);

/* ---- fft128 (src/fft128/mod.rs, src/fft128/f128_ops.rs) ---- */
typedef struct { double hi, lo; } or_f128;
or_f128 or_f128_add(or_f128 a, or_f128 b);
or_f128 or_f128_sub(or_f128 a, or_f128 b);
or_f128 or_f128_add_estimate(or_f128 a, or_f128 b);
or_f128 or_f128_sub_estimate(or_f128 a, or_f128 b);
or_f128 or_f128_mul(or_f128 a, or_f128 b);
or_f128 or_f128_mul_fma(or_f128 a, or_f128 b);
or_f128 or_f128_div(or_f128 a, or_f128 b);
void or_f128_sincospi(or_f128 x, or_f128 *s, or_f128 *c);
void or_fft128_init_twiddles(size_t n, double *re0, double *re1, double *im0, double *im1);
/* mul_variant 0: scalar path (p2 + (a0*b1 + a1*b0)), 1: fma path (AVX) */
void or_fft128_fwd(size_t n, double *re0, double *re1, double *im0, double *im1,
                   const double *t_re0, const double *t_re1, const double *t_im0,
                   const double *t_im1, int mul_variant);
void or_fft128_inv(size_t n, double *re0, double *re1, double *im0, double *im1,
                   const double *t_re0, const double *t_re1, const double *t_im0,
                   const double *t_im1, int mul_variant);

/* ---- external product (TFHE GGSW x GLWE, caller-side; see polymul.c) ---- */
void or_gadget_decompose(size_t N, const uint64_t *in, int levels, int base_log, int64_t *out);
int or_negacyclic_roundtrip_f64(size_t N, int base_algo, size_t base_n, const double *x,
                                or_c64 *F, double *out, size_t count);
int or_torus_to_fourier(size_t N, int base_algo, size_t base_n, const uint64_t *in,
                        or_c64 *out, size_t npolys);
int or_external_product(size_t N, int k1, int levels, int base_log, const or_c64 *key_fourier,
                        int base_algo, size_t base_n, const uint64_t *glwe, uint64_t *out);

/* ---- SIMD restatement of the reference's x86 path (fft64_simd.c): V3 =
 *      AVX2 + FMA (c64x2), V4 = AVX-512F (c64x4); bit-identical to the scalar
 *      functions above; DIF base algorithms, n >= 32 ---- */
int or_cpu_has_isa(int isa); /* 3: AVX2 + FMA, 4: AVX-512F + AVX-512DQ */
int or_simd_unordered_fwd_batch_v3(size_t n, int base_algo, size_t base_n, or_c64 *buf, size_t batch);
int or_simd_unordered_inv_batch_v3(size_t n, int base_algo, size_t base_n, or_c64 *buf, size_t batch);
int or_simd_polymul_f64_v3(size_t N, int base_algo, size_t base_n, const uint64_t *a,
                           const int64_t *b, uint64_t *c, size_t batch, int threads);
double or_simd_time_polymul_v3(size_t N, int base_algo, size_t base_n, const uint64_t *a,
                               const int64_t *b, uint64_t *c, size_t batch, int threads);
int or_simd_unordered_fwd_batch_v4(size_t n, int base_algo, size_t base_n, or_c64 *buf, size_t batch);
int or_simd_unordered_inv_batch_v4(size_t n, int base_algo, size_t base_n, or_c64 *buf, size_t batch);
int or_simd_polymul_f64_v4(size_t N, int base_algo, size_t base_n, const uint64_t *a,
                           const int64_t *b, uint64_t *c, size_t batch, int threads);
double or_simd_time_polymul_v4(size_t N, int base_algo, size_t base_n, const uint64_t *a,
                               const int64_t *b, uint64_t *c, size_t batch, int threads);

/* ---- CPU baseline timing helper: runs the f64 negacyclic product over
 *      `batch` polynomials with `threads` pthreads (one plan per thread)
 *      and returns seconds ---- */
double or_time_polymul(size_t N, int base_algo, size_t base_n, const uint64_t *a,
                       const int64_t *b, uint64_t *c, size_t batch, int threads);

#ifdef __cplusplus
}
#endif
#endif
