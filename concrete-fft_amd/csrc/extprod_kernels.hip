// extprod_kernels.hip -- TFHE external product (GGSW (x) GLWE) on a resident
// Fourier-domain key, and the key's transform into that domain.
//
// Conventions (caller-side in the reference; same as the negacyclic product):
// twist zeta_j, fold (re = x_j, im = x_{j+n}), the plan's unordered forward
// transform, cmul with the key / torus operand first, untwist conj(zeta)/n,
// torus rounding.  Gadget decomposition: round x to the nearest multiple of
// 2^(64 - levels*base_log) (ties up), then balanced digits in
// [-2^(B-1), 2^(B-1)) from the least significant level up, carrying.
// Row r = p * levels + i is the level-i digit polynomial of GLWE poly p
// (level 0 most significant).
#include <hip/hip_runtime.h>

#include "lds_attr.hpp"


#include "launch.hpp"
#include "r16_exec.hpp"
#include "stockham_exec.hpp"

using namespace cfft_dev;
using cfft_launch::dc64;
using cfft_prog::Program;

namespace {

// One level of the gadget decomposition, without branches or loops: the
// balanced digits of the rounded value st (rep = levels * base_log bits) are
// the plain base-2^B digits of st + (half in every digit position), each
// minus half -- the carry loop of the decomposition done by one addition
// (the same digits; the wrap-around past rep bits is the top digit's carry,
// which the decomposition drops as well).
struct Gadget {
  uint32_t shift;   // 64 - rep: the rounding position (0: no rounding)
  uint32_t down;    // base_log * (levels - 1 - level): this level's digit position
  uint64_t keep;    // the low rep bits
  uint64_t offset;  // half in every digit position
  uint64_t mask;    // one digit
  int64_t half;
};

__device__ __forceinline__ Gadget gadget_of(uint32_t level, uint32_t levels, uint32_t base_log) {
  Gadget g;
  const uint32_t rep = levels * base_log;
  g.shift = 64 - rep;
  g.down = base_log * (levels - 1 - level);
  g.keep = rep < 64 ? ((uint64_t)1 << rep) - 1 : ~(uint64_t)0;
  g.mask = ((uint64_t)1 << base_log) - 1;
  g.half = (int64_t)1 << (base_log - 1);
  g.offset = 0;
  for (uint32_t i = 0; i < levels; ++i) g.offset = (g.offset << base_log) | (uint64_t)g.half;
  return g;
}

__device__ __forceinline__ int64_t gadget_digit(const Gadget &g, uint64_t x) {
  const uint64_t up = g.shift ? (x >> (g.shift - 1)) & 1u : 0;  // ties round up
  const uint64_t st = ((x >> g.shift) + up) & g.keep;
  return (int64_t)(((st + g.offset) >> g.down) & g.mask) - g.half;
}

// key: npolys u64 torus polynomials -> Fourier (plan order), one per transform
template <int E>
__global__ void __launch_bounds__(512) k_torus_to_fourier(const uint64_t *__restrict__ key,
                                                          c64 *__restrict__ out, uint64_t npolys,
                                                          const c64 *__restrict__ tw,
                                                          const c64 *__restrict__ twist,
                                                          Program fwd) {
  extern __shared__ c64 smem[];
  const uint32_t n = 1u << fwd.log2n;
  const uint32_t T = n / E;
  const uint32_t G = blockDim.x / T;
  const uint32_t g = threadIdx.x / T, lt = threadIdx.x % T;
  const uint64_t t = (uint64_t)blockIdx.x * G + g;
  const bool active = t < npolys;
  c64 *lds = smem + g * lds_words(n);
  const uint64_t *src = key + t * 2ull * n;
  auto load = [&](uint32_t j) -> c64 {
    if (!active) return mk(0.0, 0.0);
    return cmul(twist[j], mk(torus_to_f64(src[j]), torus_to_f64(src[j + n])));
  };
  c64 *dst = out + t * n;
  auto store = [&](uint32_t pos, c64 v) {
    if (active) dst[pos] = v;
  };
  run_program<E>(fwd, lds, lt, T, tw, load, store);
}

// Split form: one transform group per (GLWE, row) transforms its digit
// polynomial into a Fourier scratch S[(p rows + r) n]; then one per (GLWE,
// column) accumulates sum_r key[r][c] * S[p, r] in ascending row order,
// inverse, torus.
template <int E>
__global__ void __launch_bounds__(512) k_extprod_rows(const uint64_t *__restrict__ glwe,
                                                      c64 *__restrict__ S, uint64_t nblocks,
                                                      uint32_t levels, uint32_t base_log,
                                                      const c64 *__restrict__ tw,
                                                      const c64 *__restrict__ twist,
                                                      Program fwd) {
  extern __shared__ c64 smem[];
  const uint32_t n = 1u << fwd.log2n;
  const uint32_t T = n / E;
  const uint32_t G = blockDim.x / T;
  const uint32_t g = threadIdx.x / T, lt = threadIdx.x % T;
  const uint64_t t = (uint64_t)blockIdx.x * G + g;  // (p * k1 + poly) * levels + lvl
  const bool active = t < nblocks;
  c64 *lds = smem + g * lds_words(n);
  const Gadget gd = gadget_of((uint32_t)(t % levels), levels, base_log);
  const uint64_t *src = glwe + (t / levels) * 2ull * n;
  auto load = [&](uint32_t j) -> c64 {
    if (!active) return mk(0.0, 0.0);
    return cmul(twist[j], mk((double)(int32_t)gadget_digit(gd, ld_stream(&src[j])),
                             (double)(int32_t)gadget_digit(gd, ld_stream(&src[j + n]))));
  };
  c64 *dst = S + t * n;
  auto store = [&](uint32_t pos, c64 v) {
    if (active) dst[pos] = v;
  };
  run_program<E>(fwd, lds, lt, T, tw, load, store);
}

template <int E>
__global__ void __launch_bounds__(512) k_extprod_cols(const c64 *__restrict__ keyf,
                                                      const c64 *__restrict__ S,
                                                      uint64_t *__restrict__ out,
                                                      uint64_t nblocks, uint32_t k1,
                                                      uint32_t rows, const c64 *__restrict__ tw,
                                                      const c64 *__restrict__ untwist,
                                                      Program inv) {
  extern __shared__ c64 smem[];
  const uint32_t n = 1u << inv.log2n;
  const uint32_t T = n / E;
  const uint32_t G = blockDim.x / T;
  const uint32_t g = threadIdx.x / T, lt = threadIdx.x % T;
  const uint64_t t = (uint64_t)blockIdx.x * G + g;  // p * k1 + c
  const bool active = t < nblocks;
  c64 *lds = smem + g * lds_words(n);
  const uint64_t p = t / k1;
  const uint32_t c = (uint32_t)(t % k1);
  const c64 *sp = S + p * (uint64_t)rows * n;
  auto load = [&](uint32_t pos) -> c64 {
    if (!active) return mk(0.0, 0.0);
    c64 acc = mk(0.0, 0.0);
    for (uint32_t r = 0; r < rows; ++r) {
      const c64 pr = cmul(keyf[((uint64_t)r * k1 + c) * n + pos], sp[(uint64_t)r * n + pos]);
      acc = mk(acc.re + pr.re, acc.im + pr.im);
    }
    return acc;
  };
  uint64_t *o = out + t * 2ull * n;
  auto store = [&](uint32_t j, c64 v) {
    if (!active) return;
    const c64 r = cmul(untwist[j], v);
    st_stream(from_torus(r.re), &o[j]);
    st_stream(from_torus(r.im), &o[j + n]);
  };
  run_program<E>(inv, lds, lt, T, tw, load, store);
}

// ---- split form on Dif16 plans, n = 512 (base 512), 1024 or 2048 (base 1024) ----
// One wavefront per row transform and per column (two, one per 32-lane
// half, at n = 512): the wave-local register-resident Dif16 transform
// (n = 512, 1024) or r16l_2048_transform
// (n = 2048: the plan's radix-2 depth pass in registers around its two
// 1024-point chunks), W = blockDim.x / 64 waves per workgroup sharing the
// 1024-point twiddle block in LDS.  The split form serves small batches:
// up to 8 rows share one workgroup, more spread one wave per workgroup and
// W grows only once every CU has a row (latency first).
// Same loads, accumulation order and stores as the generic split kernels
// above, so the same bits.
constexpr uint32_t kETw = 16 * 64 + 16 * 4;
constexpr size_t split_w_lds(uint32_t waves) {
  return (kETw + waves * R16<10>::lds_words) * sizeof(c64);
}

// copies the twiddle block into LDS with every load of a lane issued before
// its first LDS store (a one-wave workgroup would otherwise wait for 17
// dependent load round trips); callers barrier after
__device__ __forceinline__ void stage_block(const c64 *__restrict__ src, c64 *dst) {
  constexpr uint32_t kPer = (kETw + 63) / 64;
  c64 t[kPer];
#pragma unroll
  for (uint32_t k = 0; k < kPer; ++k) {
    const uint32_t i = threadIdx.x + k * blockDim.x;
    t[k] = src[i < kETw ? i : kETw - 1];
  }
#pragma unroll
  for (uint32_t k = 0; k < kPer; ++k) {
    const uint32_t i = threadIdx.x + k * blockDim.x;
    if (i < kETw) dst[i] = t[k];
  }
}

// glwe_tab (optional): the address of each GLWE of the batch (GLWE q at
// glwe_tab[q], its k1 polynomials contiguous), in place of the contiguous
// `glwe`; likewise out_tab in the column kernel.
template <int LOG2N>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_extprod_rows_w(const uint64_t *__restrict__ glwe, const uint64_t *__restrict__ glwe_tab,
                 c64 *__restrict__ S, uint64_t nrows, uint32_t k1, uint32_t levels,
                 uint32_t base_log, const c64 *__restrict__ tw, const c64 *__restrict__ tw_blk,
                 const c64 *__restrict__ twist) {
  using R = R16<10>;
  constexpr uint32_t n = 1u << LOG2N;
  extern __shared__ c64 smem[];
  c64 *ltw = smem;
  stage_block(tw_blk, ltw);
  __syncthreads();
  // units (rows) per wave: 64 / T for n <= 1024 (T = n/16 lanes each), 1 at 2048
  constexpr uint32_t T = LOG2N <= 10 ? (1u << LOG2N) / 16 : 64;
  constexpr uint32_t G = 64 / T;
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint64_t t0 = ((uint64_t)blockIdx.x * (blockDim.x >> 6) + wave) * G;
  if (t0 >= nrows) return;  // the whole wave is past the end
  const uint32_t half = (threadIdx.x & 63u) / T, l = threadIdx.x % T;
  const uint64_t t = t0 + half;  // (p k1 + poly) levels + lvl
  const bool active = t < nrows;
  c64 *x = smem + kETw + wave * R::lds_words + half * R16<(LOG2N <= 10 ? LOG2N : 10)>::lds_words;
  const Gadget gd = gadget_of((uint32_t)(t % levels), levels, base_log);
  const uint64_t q = active ? t / levels : 0;  // polynomial p k1 + poly
  const uint64_t *src =
      glwe_tab ? reinterpret_cast<const uint64_t *>(glwe_tab[q / k1]) + (q % k1) * 2ull * n
               : glwe + q * 2ull * n;
  c64 *dst = S + t * n;
  auto digit_pt = [&](uint32_t j) -> c64 {
    return cmul(twist[j], mk((double)(int32_t)gadget_digit(gd, ld_stream(&src[j])),
                             (double)(int32_t)gadget_digit(gd, ld_stream(&src[j + n]))));
  };
  if constexpr (LOG2N <= 10) {
    using L = R16<LOG2N>;
    c64 v[16];
#pragma unroll
    for (uint32_t m = 0; m < 16; ++m) v[m] = active ? digit_pt(l + T * m) : mk(0.0, 0.0);
    r16l_transform<LOG2N, true>(v, v, x, l, ltw);
    if (active) {
#pragma unroll
      for (uint32_t s = 0; s < 16; ++s) dst[l + T * L::end_m(s)] = v[s];
    }
  } else {
    c64 va[16], vb[16];
    // each half: its coefficient loads first, then digits and twist four
    // points at a time
#pragma unroll
    for (uint32_t h = 0; h < 2; ++h) {
      c64 *v = h == 0 ? va : vb;
      uint64_t r[32];
#pragma unroll
      for (uint32_t m = 0; m < 16; ++m) {
        const uint32_t j = h * 1024 + l + 64 * m;
        r[2 * m] = ld_stream(&src[j]);
        r[2 * m + 1] = ld_stream(&src[j + n]);
      }
      sched_fence();
#pragma unroll
      for (uint32_t m0 = 0; m0 < 16; m0 += 4) {
#pragma unroll
        for (uint32_t m = m0; m < m0 + 4; ++m)
          v[m] = cmul(twist[h * 1024 + l + 64 * m],
                      mk((double)(int32_t)gadget_digit(gd, r[2 * m]),
                         (double)(int32_t)gadget_digit(gd, r[2 * m + 1])));
        sched_fence();
      }
    }
    r16l_2048_transform<true>(va, vb, x, l, ltw, tw);
#pragma unroll
    for (uint32_t s = 0; s < 16; ++s) {
      const uint32_t q = l + 64 * R::end_m(s);
      dst[q] = va[s];
      dst[1024 + q] = vb[s];
    }
  }
}

template <int LOG2N>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_extprod_cols_w(const c64 *__restrict__ keyf, const c64 *__restrict__ S,
                 uint64_t *__restrict__ out, const uint64_t *__restrict__ out_tab,
                 uint64_t ncols, uint32_t k1, uint32_t rows, const c64 *__restrict__ tw,
                 const c64 *__restrict__ tw_blk, const c64 *__restrict__ untwist) {
  using R = R16<10>;
  constexpr uint32_t n = 1u << LOG2N;
  extern __shared__ c64 smem[];
  c64 *ltw = smem;
  stage_block(tw_blk, ltw);
  __syncthreads();
  constexpr uint32_t T = LOG2N <= 10 ? (1u << LOG2N) / 16 : 64;
  constexpr uint32_t G = 64 / T;
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint64_t t0 = ((uint64_t)blockIdx.x * (blockDim.x >> 6) + wave) * G;
  if (t0 >= ncols) return;  // the whole wave is past the end
  const uint32_t half = (threadIdx.x & 63u) / T, l = threadIdx.x % T;
  const uint64_t t = t0 + half;  // p * k1 + c
  const bool active = t < ncols;
  c64 *x = smem + kETw + wave * R::lds_words + half * R16<(LOG2N <= 10 ? LOG2N : 10)>::lds_words;
  const uint64_t p = active ? t / k1 : 0;
  const uint32_t c = (uint32_t)(t % k1);
  const c64 *sp = S + p * (uint64_t)rows * n;
  uint64_t *o = out_tab ? reinterpret_cast<uint64_t *>(out_tab[p]) + (uint64_t)c * 2ull * n
                        : out + (active ? t : 0) * 2ull * n;
  auto put = [&](uint32_t j, c64 v) {
    if (!active) return;
    const c64 r = cmul(untwist[j], v);
    st_stream(from_torus(r.re), &o[j]);
    st_stream(from_torus(r.im), &o[j + n]);
  };
  // Fourier multiply-accumulate row by row: each row's key and spectrum
  // points of the lane are loaded together (one memory round trip per row,
  // not per point); every point still sums its rows in ascending order
  auto mac_rows = [&](c64 *acc, uint32_t base) {
#pragma unroll
    for (uint32_t m = 0; m < 16; ++m) acc[m] = mk(0.0, 0.0);
    for (uint32_t r = 0; r < rows; ++r) {
      const c64 *kr = keyf + ((uint64_t)r * k1 + c) * n + base + l;
      const c64 *sr = sp + (uint64_t)r * n + base + l;
      c64 kv[16], sv[16];
#pragma unroll
      for (uint32_t m = 0; m < 16; ++m) {
        kv[m] = kr[T * m];
        sv[m] = sr[T * m];
      }
#pragma unroll
      for (uint32_t m = 0; m < 16; ++m) {
        const c64 pr = cmul(kv[m], sv[m]);
        acc[m] = mk(acc[m].re + pr.re, acc[m].im + pr.im);
      }
      sched_fence();  // one row's loads in flight at a time
    }
  };
  if constexpr (LOG2N <= 10) {
    using L = R16<LOG2N>;
    c64 v[16];
    mac_rows(v, 0);
    r16l_transform<LOG2N, false>(v, v, x, l, ltw);
#pragma unroll
    for (uint32_t s = 0; s < 16; ++s) put(l + T * L::end_m(s), v[s]);
  } else {
    c64 va[16], vb[16];
    mac_rows(va, 0);
    // the first half waits in the wave's exchange buffer while the second
    // accumulates (both halves' sums beside a row's loads would spill)
    wave_lds_fence();
#pragma unroll
    for (uint32_t m = 0; m < 16; ++m) x[m * 64 + l] = va[m];
    mac_rows(vb, 1024);
    wave_lds_fence();
#pragma unroll
    for (uint32_t m = 0; m < 16; ++m) va[m] = x[m * 64 + l];
    wave_lds_fence();  // read back before the transform reuses the buffer
    r16l_2048_transform<false>(va, vb, x, l, ltw, tw);
#pragma unroll
    for (uint32_t s = 0; s < 16; ++s) {
      const uint32_t q = l + 64 * R::end_m(s);
      put(q, va[s]);
      put(1024 + q, vb[s]);
    }
  }
}

template <int LOG2N>
hipError_t launch_split_w(const c64 *keyf, const uint64_t *glwe, const uint64_t *glwe_tab,
                          uint64_t *out, const uint64_t *out_tab, uint64_t batch, uint32_t k1,
                          uint32_t levels, uint32_t base_log, const c64 *tw, const c64 *tw_blk,
                          const c64 *twist, const c64 *untwist, c64 *S, hipStream_t s) {
  hipError_t e = cfft_launch::ensure_lds(reinterpret_cast<const void *>(k_extprod_rows_w<LOG2N>),
                                         split_w_lds(8));
  if (e == hipSuccess)
    e = cfft_launch::ensure_lds(reinterpret_cast<const void *>(k_extprod_cols_w<LOG2N>),
                                split_w_lds(8));
  if (e != hipSuccess) return e;
  // waves per workgroup: a few waves share one workgroup (one L2 fills its
  // tables); more spread one per workgroup until they outnumber the CUs
  // (256 on MI355X).  Measured: up to 8 row waves (decompose + forward) and
  // up to 4 column waves (6-row MAC + inverse) gain from sharing; 8 column
  // waves in one workgroup took 41 us against 32 us spread (N = 1024,
  // batch 8), while 6 row waves spread took 30 us against 28 us shared
  // (N = 2048, batch 1).
  auto waves_for = [](uint64_t units, uint64_t share) -> uint32_t {
    if (units <= share) return 8;
    uint32_t w = 1;
    while (w < 8 && units > 256ull * w) w *= 2;
    return w;
  };
  // units per wave: two at n = 512 (one per 32-lane half), else one
  constexpr uint32_t G = LOG2N == 9 ? 2 : 1;
  const uint64_t nrows = batch * k1 * levels, ncols = batch * k1;
  const uint64_t wrows = (nrows + G - 1) / G, wcols = (ncols + G - 1) / G;  // waves needed
  const uint32_t wr = waves_for(wrows, 8), wc = waves_for(wcols, 4);
  hipLaunchKernelGGL(k_extprod_rows_w<LOG2N>, dim3((uint32_t)((wrows + wr - 1) / wr)),
                     dim3(64 * wr), split_w_lds(wr), s, glwe, glwe_tab, S, nrows, k1, levels,
                     base_log, tw, tw_blk, twist);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_extprod_cols_w<LOG2N>, dim3((uint32_t)((wcols + wc - 1) / wc)),
                     dim3(64 * wc), split_w_lds(wc), s, keyf, S, out, out_tab, ncols, k1,
                     k1 * levels, tw, tw_blk, untwist);
  return hipGetLastError();
}

hipError_t set_lds(const void *kernel, size_t bytes) {
  return cfft_launch::ensure_lds(kernel, bytes);
}

struct Geom {
  int E;
  uint32_t T, G, threads;
};
Geom geom(uint32_t log2n, uint32_t words_per_group_limit, uint32_t groups_words) {
  Geom g;
  const uint32_t n = 1u << log2n;
  g.E = n >= 16 ? 16 : (int)n;
  g.T = n / (uint32_t)g.E;
  g.G = n >= 2048 ? 1 : 2048 / n;
  if (g.T * g.G > 512) g.G = 512 / g.T > 0 ? 512 / g.T : 1;
  while (g.G > 1 && (size_t)g.G * groups_words * sizeof(c64) > words_per_group_limit) g.G /= 2;
  g.threads = g.T * g.G;
  return g;
}

template <int E>
hipError_t launch_key(const Program &fwd, const uint64_t *key, c64 *out, uint64_t npolys,
                      const c64 *tw, const c64 *twist, hipStream_t s) {
  const uint32_t words = lds_words(1u << fwd.log2n);
  const Geom g = geom(fwd.log2n, 64 * 1024, words);
  const size_t lds = (size_t)g.G * words * sizeof(c64);
  hipError_t e = set_lds(reinterpret_cast<const void *>(k_torus_to_fourier<E>), lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_torus_to_fourier<E>, dim3((uint32_t)((npolys + g.G - 1) / g.G)),
                     dim3(g.threads), lds, s, key, out, npolys, tw, twist, fwd);
  return hipGetLastError();
}

template <int E>
hipError_t launch_split(const Program &fwd, const Program &inv, const c64 *keyf,
                        const uint64_t *glwe, uint64_t *out, uint64_t batch, uint32_t k1,
                        uint32_t levels, uint32_t base_log, const c64 *tw, const c64 *twist,
                        const c64 *untwist, c64 *S, hipStream_t s) {
  const uint32_t words = lds_words(1u << fwd.log2n);
  const Geom g = geom(fwd.log2n, 64 * 1024, words);
  const size_t lds = (size_t)g.G * words * sizeof(c64);
  hipError_t e = set_lds(reinterpret_cast<const void *>(k_extprod_rows<E>), lds);
  if (e == hipSuccess) e = set_lds(reinterpret_cast<const void *>(k_extprod_cols<E>), lds);
  if (e != hipSuccess) return e;
  const uint64_t nrows = batch * k1 * levels, ncols = batch * k1;
  hipLaunchKernelGGL(k_extprod_rows<E>, dim3((uint32_t)((nrows + g.G - 1) / g.G)),
                     dim3(g.threads), lds, s, glwe, S, nrows, levels, base_log, tw, twist, fwd);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_extprod_cols<E>, dim3((uint32_t)((ncols + g.G - 1) / g.G)),
                     dim3(g.threads), lds, s, keyf, S, out, ncols, k1, k1 * levels, tw, untwist,
                     inv);
  return hipGetLastError();
}

int pick_e(uint32_t log2n) {
  const uint32_t n = 1u << log2n;
  return n >= 16 ? 16 : (int)n;
}

}  // namespace

namespace cfft_launch {

hipError_t torus_to_fourier(const Program &fwd, const uint64_t *key, dc64 *out, uint64_t npolys,
                            const dc64 *tw, const dc64 *twist, hipStream_t s) {
  if (npolys == 0) return hipSuccess;
  c64 *o = reinterpret_cast<c64 *>(out);
  const c64 *t = reinterpret_cast<const c64 *>(tw), *z = reinterpret_cast<const c64 *>(twist);
  switch (pick_e(fwd.log2n)) {
    case 2: return launch_key<2>(fwd, key, o, npolys, t, z, s);
    case 4: return launch_key<4>(fwd, key, o, npolys, t, z, s);
    case 8: return launch_key<8>(fwd, key, o, npolys, t, z, s);
    default: return launch_key<16>(fwd, key, o, npolys, t, z, s);
  }
}

hipError_t external_product_split(const Program &fwd, const Program &inv, const dc64 *keyf,
                                  const uint64_t *glwe, uint64_t *out, uint64_t batch,
                                  uint32_t k1, uint32_t levels, uint32_t base_log,
                                  const dc64 *tw, const dc64 *twist, const dc64 *untwist,
                                  dc64 *scratch, hipStream_t s) {
  if (batch == 0) return hipSuccess;
  const c64 *kf = reinterpret_cast<const c64 *>(keyf);
  const c64 *t = reinterpret_cast<const c64 *>(tw), *z = reinterpret_cast<const c64 *>(twist);
  const c64 *u = reinterpret_cast<const c64 *>(untwist);
  c64 *S = reinterpret_cast<c64 *>(scratch);
  switch (pick_e(fwd.log2n)) {
    case 2: return launch_split<2>(fwd, inv, kf, glwe, out, batch, k1, levels, base_log, t, z, u, S, s);
    case 4: return launch_split<4>(fwd, inv, kf, glwe, out, batch, k1, levels, base_log, t, z, u, S, s);
    case 8: return launch_split<8>(fwd, inv, kf, glwe, out, batch, k1, levels, base_log, t, z, u, S, s);
    default:
      return launch_split<16>(fwd, inv, kf, glwe, out, batch, k1, levels, base_log, t, z, u, S, s);
  }
}

// Split form on the wave-local kernels: Dif16 plans with n = 1024 (tw_blk =
// the plan's r16 block), n = 512 or n = 2048 and base_n = 1024 (the plan's
// 1024-point block).  glwe_tab / out_tab (optional, device-readable): one
// address per GLWE of the batch in place of the contiguous glwe / out.
hipError_t external_product_split_w(uint32_t log2n, const dc64 *keyf, const uint64_t *glwe,
                                    const uint64_t *glwe_tab, uint64_t *out,
                                    const uint64_t *out_tab, uint64_t batch, uint32_t k1,
                                    uint32_t levels, uint32_t base_log, const dc64 *tw,
                                    const dc64 *tw_blk, const dc64 *twist, const dc64 *untwist,
                                    dc64 *scratch, hipStream_t s) {
  if (batch == 0) return hipSuccess;
  const c64 *kf = reinterpret_cast<const c64 *>(keyf);
  const c64 *t = reinterpret_cast<const c64 *>(tw), *tb = reinterpret_cast<const c64 *>(tw_blk);
  const c64 *z = reinterpret_cast<const c64 *>(twist), *u = reinterpret_cast<const c64 *>(untwist);
  c64 *S = reinterpret_cast<c64 *>(scratch);
  if (log2n == 9)
    return launch_split_w<9>(kf, glwe, glwe_tab, out, out_tab, batch, k1, levels, base_log, t,
                                tb, z, u, S, s);
  if (log2n == 10)
    return launch_split_w<10>(kf, glwe, glwe_tab, out, out_tab, batch, k1, levels, base_log, t,
                                tb, z, u, S, s);
  if (log2n == 11)
    return launch_split_w<11>(kf, glwe, glwe_tab, out, out_tab, batch, k1, levels, base_log, t,
                                tb, z, u, S, s);
  return hipErrorInvalidValue;
}

}  // namespace cfft_launch
