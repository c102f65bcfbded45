/*
 * rng.c -- restatement of the input generator of the reference's golden test
 * (src/unordered.rs:1176-1188: `StdRng::seed_from_u64(0)`, then
 * `gen_range(0.0..1.0)` for re and im of each of 2048 values).
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * The generator lives in third-party crates that are not in the reference
 * tree (no lockfile; Cargo.toml:13-48 declares rand 0.8):
 *   rand 0.8 StdRng        = rand_chacha 0.3 ChaCha12Rng (ChaCha, 12 rounds,
 *                            256-bit key = seed, 64-bit block counter in
 *                            words 12-13, 64-bit stream id 0 in words 14-15,
 *                            output words consumed in order)
 *   rand_core 0.6          seed_from_u64: 8 x PCG32 (XSH-RR) outputs,
 *                            state advanced before each output with
 *                            MUL = 6364136223846793005,
 *                            INC = 11634580027462260723
 *   rand_core BlockRng     next_u64 = lo word | hi word << 32
 *   rand 0.8 UniformFloat  sample_single(0.0, 1.0): (next_u64 >> 12) as the
 *                            mantissa of a value in [1, 2), minus 1.0, times
 *                            scale 1.0, plus low 0.0
 * These are the published algorithms of those crates; the reference's golden
 * vector pins this restatement bit-for-bit.
 */
#include "oracle.h"

#include <string.h>

static inline uint32_t rotl32(uint32_t x, int k) { return (x << k) | (x >> (32 - k)); }

#define QR(a, b, c, d)                 \
  do {                                 \
    a += b; d ^= a; d = rotl32(d, 16); \
    c += d; b ^= c; b = rotl32(b, 12); \
    a += b; d ^= a; d = rotl32(d, 8);  \
    c += d; b ^= c; b = rotl32(b, 7);  \
  } while (0)

static void chacha_block(const uint32_t key[8], uint64_t counter, int rounds, uint32_t out[16]) {
  uint32_t s[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u,
                    key[0], key[1], key[2], key[3], key[4], key[5], key[6], key[7],
                    (uint32_t)counter, (uint32_t)(counter >> 32), 0u, 0u};
  uint32_t x[16];
  memcpy(x, s, sizeof(x));
  for (int i = 0; i < rounds; i += 2) {
    QR(x[0], x[4], x[8], x[12]);
    QR(x[1], x[5], x[9], x[13]);
    QR(x[2], x[6], x[10], x[14]);
    QR(x[3], x[7], x[11], x[15]);
    QR(x[0], x[5], x[10], x[15]);
    QR(x[1], x[6], x[11], x[12]);
    QR(x[2], x[7], x[8], x[13]);
    QR(x[3], x[4], x[9], x[14]);
  }
  for (int i = 0; i < 16; i++) out[i] = x[i] + s[i];
}

static uint32_t pcg32_step(uint64_t *state) {
  const uint64_t MUL = 6364136223846793005ull;
  const uint64_t INC = 11634580027462260723ull;
  *state = *state * MUL + INC;
  uint64_t st = *state;
  uint32_t xorshifted = (uint32_t)(((st >> 18) ^ st) >> 27);
  uint32_t rot = (uint32_t)(st >> 59);
  return (xorshifted >> rot) | (xorshifted << ((32 - rot) & 31));
}

void or_stdrng_seed_from_u64(or_stdrng *r, uint64_t seed) {
  uint64_t st = seed;
  for (int i = 0; i < 8; i++) r->key[i] = pcg32_step(&st); /* to_le_bytes, read back LE */
  r->counter = 0;
  r->idx = 16;
}

uint32_t or_stdrng_next_u32(or_stdrng *r) {
  if (r->idx >= 16) {
    chacha_block(r->key, r->counter, 12, r->buf);
    r->counter++;
    r->idx = 0;
  }
  return r->buf[r->idx++];
}

uint64_t or_stdrng_next_u64(or_stdrng *r) {
  uint64_t lo = or_stdrng_next_u32(r);
  uint64_t hi = or_stdrng_next_u32(r);
  return lo | (hi << 32);
}

double or_stdrng_gen_f64_01(or_stdrng *r) {
  uint64_t bits = (or_stdrng_next_u64(r) >> 12) | (1023ull << 52);
  double v12;
  memcpy(&v12, &bits, 8);
  double v01 = v12 - 1.0;
  return v01 * 1.0 + 0.0;
}

void or_stdrng_fill_f64_01(uint64_t seed, double *out, size_t count) {
  or_stdrng r;
  or_stdrng_seed_from_u64(&r, seed);
  for (size_t i = 0; i < count; i++) out[i] = or_stdrng_gen_f64_01(&r);
}
