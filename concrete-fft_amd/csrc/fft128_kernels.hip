// fft128_kernels.hip -- gfx950 kernels of the 128-bit (double-double)
// negacyclic FFT (fft128::Plan, src/fft128/mod.rs).
//
// Arithmetic: the reference's error-free transforms (src/fft128/f128_ops.rs:
// 3-40) and double-double estimates (add/sub_estimate :298-354, mul :394-400)
// in the same operation order, built without contraction; two_prod uses the
// hardware FP64 FMA.  MULV 0 reproduces the scalar product
// p2 + (a0*b1 + a1*b0), MULV 1 the AVX2/AVX-512 product
// fma(a0, b1, fma(a1, b0, p2)) (:836-841).
//
// Schedule: one workgroup per transform of n points (4 SoA arrays in LDS,
// 32 B per point).  The log2(n) radix-2 stages (CT forward
// src/fft128/mod.rs:351-402, GS inverse :1104-1155) are taken three at a
// time: each thread owns 8 points that form closed radix-2 networks over those
// stages, applies them in registers, and exchanges through LDS between
// rounds.  The first round reads global memory and the last writes it.
#include <hip/hip_runtime.h>

#include "lds_attr.hpp"


namespace {

struct dd {
  double hi, lo;
};

__device__ __forceinline__ void quick_two_sum(double a, double b, double &s, double &e) {
  s = a + b;
  e = b - (s - a);
}
__device__ __forceinline__ void two_sum(double a, double b, double &s, double &e) {
  s = a + b;
  const double bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}
__device__ __forceinline__ void two_diff(double a, double b, double &s, double &e) {
  s = a - b;
  const double bb = s - a;
  e = (a - (s - bb)) - (b + bb);
}
__device__ __forceinline__ dd add_est(dd a, dd b) {
  double s, e;
  two_sum(a.hi, b.hi, s, e);
  e = e + (a.lo + b.lo);
  dd r;
  quick_two_sum(s, e, r.hi, r.lo);
  return r;
}
__device__ __forceinline__ dd sub_est(dd a, dd b) {
  double s, e;
  two_diff(a.hi, b.hi, s, e);
  e = e + a.lo;
  e = e - b.lo;
  dd r;
  quick_two_sum(s, e, r.hi, r.lo);
  return r;
}
template <int MULV>
__device__ __forceinline__ dd mul(dd a, dd b) {
  const double p1 = a.hi * b.hi;
  double p2 = __fma_rn(a.hi, b.hi, -p1);
  if (MULV == 0) p2 = p2 + (a.hi * b.lo + a.lo * b.hi);
  else p2 = __fma_rn(a.hi, b.lo, __fma_rn(a.lo, b.hi, p2));
  dd r;
  quick_two_sum(p1, p2, r.hi, r.lo);
  return r;
}

struct cdd {
  dd re, im;
};

// cplx_mul (src/fft128/mod.rs:310-326) and cplx_mul_conj (:329-345)
template <int MULV>
__device__ __forceinline__ cdd cmul(cdd a, cdd b) {
  const dd rr = mul<MULV>(a.re, b.re), ri = mul<MULV>(a.re, b.im);
  const dd ir = mul<MULV>(a.im, b.re), ii = mul<MULV>(a.im, b.im);
  cdd r;
  r.re = sub_est(rr, ii);
  r.im = add_est(ir, ri);
  return r;
}
template <int MULV>
__device__ __forceinline__ cdd cmul_conj(cdd a, cdd b) {
  const dd rr = mul<MULV>(a.re, b.re), ri = mul<MULV>(a.re, b.im);
  const dd ir = mul<MULV>(a.im, b.re), ii = mul<MULV>(a.im, b.im);
  cdd r;
  r.re = add_est(rr, ii);
  r.im = sub_est(ir, ri);
  return r;
}

struct Tw {
  const double *re0, *re1, *im0, *im1;
};
__device__ __forceinline__ cdd load_tw(const Tw &t, uint32_t i) {
  cdd w;
  w.re.hi = t.re0[i];
  w.re.lo = t.re1[i];
  w.im.hi = t.im0[i];
  w.im.lo = t.im1[i];
  return w;
}

struct Bufs {
  double *re0, *re1, *im0, *im1;
};

__host__ __device__ __forceinline__ uint32_t pad128(uint32_t i) { return i + (i >> 5); }
// blocks of 2^L points per k_fft128 workgroup: enough to fill one wave
__host__ __device__ __forceinline__ uint32_t fft128_blocks_per_wg(uint32_t L) {
  return L >= 9 ? 1u : 1u << (9 - L);
}

// Position of point `u` (3-bit combination, most significant = first stage
// of the round) of group `g`, for a round whose top group bit is `ptop` and
// which spans r bits (ptop, ptop-1, ..., ptop-r+1).
template <int R>
__device__ __forceinline__ uint32_t group_pos(uint32_t g, uint32_t ptop, uint32_t u) {
  const uint32_t low = ptop + 1 - R;  // bits below the group bits
  const uint32_t base = ((g >> low) << (low + R)) | (g & ((1u << low) - 1));
  uint32_t off = 0;
#pragma unroll
  for (int b = 0; b < R; ++b) off |= ((u >> (R - 1 - b)) & 1u) << (ptop - b);
  return base + off;
}

// One round of R stages on NG = 8 / R-point groups per thread.
// FWD: stages st0, st0+1, .. (bit ptop = Lg-1-st0 downwards), CT butterflies.
// INV: stages processed from the low bit up, GS butterflies.
// Positions are local to a block of 2^L points that starts at point jhi of a
// transform of 2^Lg points (jhi = 0, Lg = L for a whole transform); stage
// indices and twiddles are those of the whole transform.
template <int R, bool FWD, int MULV>
__device__ __forceinline__ void round_regs(cdd *z, uint32_t g, uint32_t ptop, uint32_t Lg,
                                           uint32_t jhi, const Tw &tw) {
  // z[u], u < 2^R, indexed with the bit for position ptop-b at bit (R-1-b)
  const uint32_t base = group_pos<R>(g, ptop, 0);
  if (FWD) {
#pragma unroll
    for (int b = 0; b < R; ++b) {
      const uint32_t p = ptop - (uint32_t)b;       // pair bit
      const uint32_t st = Lg - 1 - p;              // stage index, m = 2^st
#pragma unroll
      for (int u = 0; u < (1 << R); ++u) {
        if ((u >> (R - 1 - b)) & 1) continue;
        const int v = u | (1 << (R - 1 - b));
        uint32_t j = base;
#pragma unroll
        for (int c = 0; c < R; ++c) j |= ((uint32_t)(u >> (R - 1 - c)) & 1u) << (ptop - c);
        const uint32_t i = (jhi | j) >> (p + 1);
        const cdd w = load_tw(tw, (1u << st) + i);
        const cdd z1w = cmul<MULV>(z[v], w);
        const cdd z0 = z[u];
        z[u].re = add_est(z0.re, z1w.re);
        z[u].im = add_est(z0.im, z1w.im);
        z[v].re = sub_est(z0.re, z1w.re);
        z[v].im = sub_est(z0.im, z1w.im);
      }
    }
  } else {
#pragma unroll
    for (int b = R - 1; b >= 0; --b) {
      const uint32_t p = ptop - (uint32_t)b;
      const uint32_t st = Lg - 1 - p;
#pragma unroll
      for (int u = 0; u < (1 << R); ++u) {
        if ((u >> (R - 1 - b)) & 1) continue;
        const int v = u | (1 << (R - 1 - b));
        uint32_t j = base;
#pragma unroll
        for (int c = 0; c < R; ++c) j |= ((uint32_t)(u >> (R - 1 - c)) & 1u) << (ptop - c);
        const uint32_t i = (jhi | j) >> (p + 1);
        const cdd w = load_tw(tw, (1u << st) + i);
        const cdd z0 = z[u], z1 = z[v];
        cdd d;
        d.re = sub_est(z0.re, z1.re);
        d.im = sub_est(z0.im, z1.im);
        z[u].re = add_est(z0.re, z1.re);
        z[u].im = add_est(z0.im, z1.im);
        z[v] = cmul_conj<MULV>(d, w);
      }
    }
  }
}

template <int R, bool FWD, int MULV>
__device__ __forceinline__ void run_round(double *lds, const Bufs &gbuf, bool first, bool last,
                                          uint32_t ptop, uint32_t L, uint32_t Lg, uint32_t jhi,
                                          uint32_t lt, uint32_t T, const Tw &tw, bool active) {
  constexpr int P = 1 << R;
  constexpr int NG = 8 / P;
  const uint32_t n = 1u << L;
  const uint32_t pn = pad128(n);
  cdd z[NG][P];
#pragma unroll
  for (int q = 0; q < NG; ++q) {
    const uint32_t g = lt + T * (uint32_t)q;
#pragma unroll
    for (int u = 0; u < P; ++u) {
      const uint32_t j = group_pos<R>(g, ptop, (uint32_t)u);
      if (first) {
        if (active) {
          z[q][u].re.hi = gbuf.re0[j];
          z[q][u].re.lo = gbuf.re1[j];
          z[q][u].im.hi = gbuf.im0[j];
          z[q][u].im.lo = gbuf.im1[j];
        } else {
          z[q][u].re.hi = z[q][u].re.lo = z[q][u].im.hi = z[q][u].im.lo = 0.0;
        }
      } else {
        const uint32_t a = pad128(j);
        z[q][u].re.hi = lds[a];
        z[q][u].re.lo = lds[pn + a];
        z[q][u].im.hi = lds[2 * pn + a];
        z[q][u].im.lo = lds[3 * pn + a];
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NG; ++q) {
    const uint32_t g = lt + T * (uint32_t)q;
    round_regs<R, FWD, MULV>(z[q], g, ptop, Lg, jhi, tw);
#pragma unroll
    for (int u = 0; u < P; ++u) {
      const uint32_t j = group_pos<R>(g, ptop, (uint32_t)u);
      if (last) {
        if (active) {
          gbuf.re0[j] = z[q][u].re.hi;
          gbuf.re1[j] = z[q][u].re.lo;
          gbuf.im0[j] = z[q][u].im.hi;
          gbuf.im1[j] = z[q][u].im.lo;
        }
      } else {
        const uint32_t a = pad128(j);
        lds[a] = z[q][u].re.hi;
        lds[pn + a] = z[q][u].re.lo;
        lds[2 * pn + a] = z[q][u].im.hi;
        lds[3 * pn + a] = z[q][u].im.lo;
      }
    }
  }
  __syncthreads();
}

// Blocks of n = 2^L points (L >= 3), T = n / 8 threads each: the rounds over
// bits 0 .. L-1 of a transform of 2^Lg points (Lg >= L) whose blocks are
// contiguous; block b is block b % 2^(Lg-L) of transform b / 2^(Lg-L), at
// b * n.  A workgroup holds G = max(1, 64 / T) blocks (at least one full
// wave), each with its own LDS slice; a block past the last is idle but
// takes part in every barrier.
template <bool FWD, int MULV>
__global__ void __launch_bounds__(512) k_fft128(double *re0, double *re1, double *im0,
                                                double *im1, const double *__restrict__ t_re0,
                                                const double *__restrict__ t_re1,
                                                const double *__restrict__ t_im0,
                                                const double *__restrict__ t_im1, uint32_t L,
                                                uint32_t Lg, uint64_t blocks) {
  extern __shared__ double smem[];
  const uint32_t n = 1u << L;
  const uint32_t T = n / 8;
  const uint32_t G = fft128_blocks_per_wg(L);
  const uint32_t sub = threadIdx.x >> (L - 3), lt = threadIdx.x & (T - 1);
  const uint64_t tr = (uint64_t)blockIdx.x * G + sub;
  const bool active = tr < blocks;
  double *lds = smem + (size_t)sub * 4 * pad128(n);
  const uint32_t jhi = (uint32_t)(tr & ((1ull << (Lg - L)) - 1)) << L;
  Bufs g{re0 + tr * n, re1 + tr * n, im0 + tr * n, im1 + tr * n};
  Tw tw{t_re0, t_re1, t_im0, t_im1};
  const uint32_t rounds = (L + 2) / 3;
  for (uint32_t r = 0; r < rounds; ++r) {
    const bool first = r == 0, last = r + 1 == rounds;
    // forward: bits L-1 down to 0 in chunks of 3 from the top;
    // inverse: bits 0 up to L-1 in chunks of 3 from the bottom
    uint32_t lo_bit, nb;
    if (FWD) {
      const uint32_t top = L - 1 - 3 * r;
      nb = top + 1 >= 3 ? 3 : top + 1;
      lo_bit = top + 1 - nb;
    } else {
      lo_bit = 3 * r;
      nb = L - lo_bit >= 3 ? 3 : L - lo_bit;
    }
    const uint32_t ptop = lo_bit + nb - 1;
    switch (nb) {
      case 3: run_round<3, FWD, MULV>(lds, g, first, last, ptop, L, Lg, jhi, lt, T, tw, active); break;
      case 2: run_round<2, FWD, MULV>(lds, g, first, last, ptop, L, Lg, jhi, lt, T, tw, active); break;
      default: run_round<1, FWD, MULV>(lds, g, first, last, ptop, L, Lg, jhi, lt, T, tw, active); break;
    }
  }
}

// One round of R stages (bits ptop .. ptop-R+1, all above the LDS block)
// over global memory: a thread per R-bit group of 2^R points, loaded, run
// through the round's butterflies in registers and stored in place.
// Consecutive threads take consecutive groups, which sit at consecutive
// points (the round's bits are all >= R), so every access is coalesced.
template <int R, bool FWD, int MULV>
__global__ void __launch_bounds__(256) k_fft128_global(double *re0, double *re1, double *im0,
                                                       double *im1, const double *__restrict__ t_re0,
                                                       const double *__restrict__ t_re1,
                                                       const double *__restrict__ t_im0,
                                                       const double *__restrict__ t_im1,
                                                       uint32_t Lg, uint32_t ptop, uint64_t groups) {
  constexpr int P = 1 << R;
  const uint64_t gid = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (gid >= groups) return;
  const uint64_t tr = gid >> (Lg - R);
  const uint32_t g = (uint32_t)(gid & ((1ull << (Lg - R)) - 1));
  const uint64_t off = tr << Lg;
  Tw tw{t_re0, t_re1, t_im0, t_im1};
  cdd z[P];
#pragma unroll
  for (int u = 0; u < P; ++u) {
    const uint64_t j = off + group_pos<R>(g, ptop, (uint32_t)u);
    z[u].re.hi = re0[j];
    z[u].re.lo = re1[j];
    z[u].im.hi = im0[j];
    z[u].im.lo = im1[j];
  }
  round_regs<R, FWD, MULV>(z, g, ptop, Lg, 0u, tw);
#pragma unroll
  for (int u = 0; u < P; ++u) {
    const uint64_t j = off + group_pos<R>(g, ptop, (uint32_t)u);
    re0[j] = z[u].re.hi;
    re1[j] = z[u].re.lo;
    im0[j] = z[u].im.hi;
    im1[j] = z[u].im.lo;
  }
}

// ---- exact u64 product over the double-double transform ----
//
// C = A * B mod (X^N + 1, 2^64) with A = a0 + 2^32 a1 (balanced 32-bit limbs,
// likewise B): C = a0 (x) b0 + 2^32 (a0 (x) b1 + a1 (x) b0) mod 2^64.  Each
// limb product has |coefficients| <= N 2^62 <= 2^76 for N <= 8192, inside the
// 106-bit double-double mantissa, so rounding the transform result to the
// nearest integer recovers it exactly while the transform error stays below
// 1/2 (the largest distance seen is reported as the margin).  The fold, the
// (2/N) factor and the layout follow the reference's fft128 product test
// (src/fft128/mod.rs:1996-2058).
//
// Scratch S holds 4 SoA arrays of 4*batch transforms of n points:
// S[arr][(t * batch + p) * n + j], t = a0, a1, b0, b1; after the pointwise
// step t = 0 holds a0 b0 and t = 1 holds a0 b1 + a1 b0.

__device__ __forceinline__ void limbs32(uint64_t v, double &lo, double &hi) {
  int64_t d0 = (int64_t)(v & 0xffffffffull);
  const int64_t carry = d0 >= ((int64_t)1 << 31) ? 1 : 0;
  d0 -= carry << 32;
  // the top limb only matters mod 2^32: balance it too
  int64_t d1 = (int64_t)(v >> 32) + carry;
  d1 = ((d1 + ((int64_t)1 << 31)) & 0xffffffffll) - ((int64_t)1 << 31);
  lo = (double)d0;
  hi = (double)d1;
}

__global__ void __launch_bounds__(256) k_exact128_split(const uint64_t *__restrict__ A,
                                                        const uint64_t *__restrict__ B,
                                                        double *__restrict__ S, double *margin,
                                                        uint64_t batch, uint32_t L) {
  const uint32_t n = 1u << L;
  const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= batch * n) return;
  const uint64_t p = gid >> L;
  const uint32_t j = (uint32_t)(gid & (n - 1));
  if (margin && j == 0) margin[p] = 0.0;
  const uint64_t plane = 4 * batch * (uint64_t)n;  // one SoA array
  double *re0 = S, *re1 = S + plane, *im0 = S + 2 * plane, *im1 = S + 3 * plane;
  const uint64_t *ab[2] = {A + p * 2 * n, B + p * 2 * n};
#pragma unroll
  for (int o = 0; o < 2; ++o) {
    double r_lo, r_hi, i_lo, i_hi;
    limbs32(ab[o][j], r_lo, r_hi);
    limbs32(ab[o][j + n], i_lo, i_hi);
    const uint64_t t0 = ((uint64_t)(2 * o) * batch + p) * n + j;
    const uint64_t t1 = ((uint64_t)(2 * o + 1) * batch + p) * n + j;
    re0[t0] = r_lo;
    re1[t0] = 0.0;
    im0[t0] = i_lo;
    im1[t0] = 0.0;
    re0[t1] = r_hi;
    re1[t1] = 0.0;
    im0[t1] = i_hi;
    im1[t1] = 0.0;
  }
}

__global__ void __launch_bounds__(256) k_exact128_pointwise(double *__restrict__ S,
                                                            uint64_t batch, uint32_t L) {
  const uint32_t n = 1u << L;
  const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= batch * n) return;
  const uint64_t plane = 4 * batch * (uint64_t)n;
  double *re0 = S, *re1 = S + plane, *im0 = S + 2 * plane, *im1 = S + 3 * plane;
  cdd z[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const uint64_t i = (uint64_t)t * batch * n + gid;
    z[t].re = dd{re0[i], re1[i]};
    z[t].im = dd{im0[i], im1[i]};
  }
  const double f = 2.0 / (double)(2u * n);  // 1/n, exact
  const cdd p0 = cmul<1>(z[0], z[2]);
  const cdd q0 = cmul<1>(z[0], z[3]), q1 = cmul<1>(z[1], z[2]);
  cdd p1;
  p1.re = add_est(q0.re, q1.re);
  p1.im = add_est(q0.im, q1.im);
  re0[gid] = p0.re.hi * f;
  re1[gid] = p0.re.lo * f;
  im0[gid] = p0.im.hi * f;
  im1[gid] = p0.im.lo * f;
  const uint64_t i1 = batch * n + gid;
  re0[i1] = p1.re.hi * f;
  re1[i1] = p1.re.lo * f;
  im0[i1] = p1.im.hi * f;
  im1[i1] = p1.im.lo * f;
}

// integer-valued double -> its value mod 2^64 (any magnitude)
__device__ __forceinline__ uint64_t int_double_mod64(double x) {
  if (fabs(x) < 9223372036854775808.0) return (uint64_t)(int64_t)x;
  const uint64_t bits = (uint64_t)__double_as_longlong(x);
  const uint32_t ex = (uint32_t)(bits >> 52) & 0x7ffu;  // >= 1086 here
  const uint64_t m = (bits & 0x000fffffffffffffull) | 0x0010000000000000ull;
  const uint32_t sh = ex - 1075u;
  const uint64_t r = sh >= 64u ? 0ull : (m << sh);
  return (bits >> 63) ? (0ull - r) : r;
}

// nearest integer of hi + lo, mod 2^64, and its distance
__device__ __forceinline__ uint64_t dd_round_mod64(double hi, double lo, double &dist) {
  const double rh = round(hi);
  const double f = (hi - rh) + lo;  // hi - rh is exact
  const double rl = round(f);
  dist = fabs(f - rl);
  return int_double_mod64(rh) + (uint64_t)(int64_t)rl;
}

__global__ void __launch_bounds__(256) k_exact128_combine(const double *__restrict__ S,
                                                          uint64_t *__restrict__ C,
                                                          double *margin, uint64_t batch,
                                                          uint32_t L) {
  const uint32_t n = 1u << L;
  const uint64_t total = batch * n;
  const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = gid < total;
  const uint64_t g = active ? gid : total - 1;
  const uint64_t p = g >> L;
  const uint32_t j = (uint32_t)(g & (n - 1));
  const uint64_t plane = 4 * batch * (uint64_t)n;
  const double *re0 = S, *re1 = S + plane, *im0 = S + 2 * plane, *im1 = S + 3 * plane;
  const uint64_t i1 = batch * n + g;
  double d0, d1, d2, d3;
  const uint64_t c_re = dd_round_mod64(re0[g], re1[g], d0) +
                        (dd_round_mod64(re0[i1], re1[i1], d1) << 32);
  const uint64_t c_im = dd_round_mod64(im0[g], im1[g], d2) +
                        (dd_round_mod64(im0[i1], im1[i1], d3) << 32);
  if (active) {
    C[p * 2 * n + j] = c_re;
    C[p * 2 * n + j + n] = c_im;
  }
  if (margin) {
    // one atomic per product per wave: max over the product's lanes first
    // (non-negative doubles order like their bit patterns)
    double w = active ? fmax(fmax(d0, d1), fmax(d2, d3)) : 0.0;
    const uint32_t seg = n < 64 ? n : 64;
    for (uint32_t off = 1; off < seg; off <<= 1) w = fmax(w, __shfl_xor(w, (int)off));
    if (active && (threadIdx.x & (seg - 1)) == 0)
      atomicMax(reinterpret_cast<unsigned long long *>(margin + p),
                (unsigned long long)__double_as_longlong(w));
  }
}

hipError_t set_lds(const void *kernel, size_t bytes) {
  return cfft_launch::ensure_lds(kernel, bytes);
}

}  // namespace

namespace cfft_launch {

// fft128 block held in one workgroup: 2048 points = 68 KiB of LDS, so two
// workgroups share a CU (a 4096-point block, 136 KiB, leaves one per CU
// waiting at every barrier; measured, batched n = 8192 forward 2.0 -> 2.7 M/s,
// n = 16384 0.92 -> 1.19 M/s, n = 4096 forward 4.4 -> 5.2 M/s), except the
// n = 4096 inverse, which keeps the whole transform in one block (5.2 against
// 5.0 M/s).  Larger transforms (up to 2^kFft128MaxLog2) run their stages
// above the block over global memory: forward first (CT runs from the top
// bit down), inverse last (GS runs from the bottom bit up).
constexpr uint32_t kFft128BlockLog2 = 11;
constexpr uint32_t kFft128InvWholeLog2 = 12;
constexpr uint32_t kFft128MaxLog2 = 20;

namespace {

template <bool FWD, int MULV>
hipError_t fft128_global_round(uint32_t R, uint32_t Lg, uint32_t ptop, double *re0, double *re1,
                               double *im0, double *im1, const double *t_re0,
                               const double *t_re1, const double *t_im0, const double *t_im1,
                               uint64_t batch, hipStream_t s) {
  const uint64_t groups = batch << (Lg - R);
  const dim3 grid((uint32_t)((groups + 255) / 256)), block(256);
  switch (R) {
    case 3:
      hipLaunchKernelGGL((k_fft128_global<3, FWD, MULV>), grid, block, 0, s, re0, re1, im0, im1,
                         t_re0, t_re1, t_im0, t_im1, Lg, ptop, groups);
      break;
    case 2:
      hipLaunchKernelGGL((k_fft128_global<2, FWD, MULV>), grid, block, 0, s, re0, re1, im0, im1,
                         t_re0, t_re1, t_im0, t_im1, Lg, ptop, groups);
      break;
    default:
      hipLaunchKernelGGL((k_fft128_global<1, FWD, MULV>), grid, block, 0, s, re0, re1, im0, im1,
                         t_re0, t_re1, t_im0, t_im1, Lg, ptop, groups);
      break;
  }
  return hipGetLastError();
}

template <bool FWD, int MULV>
hipError_t fft128_blocks(uint32_t L, uint32_t Lg, double *re0, double *re1, double *im0,
                         double *im1, const double *t_re0, const double *t_re1,
                         const double *t_im0, const double *t_im1, uint64_t batch, hipStream_t s) {
  const uint32_t n = 1u << L;
  const uint64_t blocks = batch << (Lg - L);
  const uint32_t G = fft128_blocks_per_wg(L);
  const size_t lds = (size_t)G * 4 * pad128(n) * sizeof(double);
  const void *k = reinterpret_cast<const void *>(k_fft128<FWD, MULV>);
  hipError_t e = set_lds(k, lds);
  if (e != hipSuccess) return e;
  const uint64_t wgs = (blocks + G - 1) / G;
  hipLaunchKernelGGL((k_fft128<FWD, MULV>), dim3((uint32_t)wgs), dim3(G * (n / 8)), lds, s, re0,
                     re1, im0, im1, t_re0, t_re1, t_im0, t_im1, L, Lg, blocks);
  return hipGetLastError();
}

template <bool FWD, int MULV>
hipError_t fft128_any(uint32_t Lg, double *re0, double *re1, double *im0, double *im1,
                      const double *t_re0, const double *t_re1, const double *t_im0,
                      const double *t_im1, uint64_t batch, hipStream_t s) {
  const uint32_t L = !FWD && Lg == kFft128InvWholeLog2 ? Lg
                     : Lg < kFft128BlockLog2             ? Lg
                                                         : kFft128BlockLog2;
  hipError_t e = hipSuccess;
  if (FWD) {
    // bits Lg-1 .. L in rounds of up to 3 from the top, then the blocks
    for (uint32_t top = Lg - 1; top + 1 > L && e == hipSuccess;) {
      const uint32_t R = top + 1 - L >= 3 ? 3 : top + 1 - L;
      e = fft128_global_round<FWD, MULV>(R, Lg, top, re0, re1, im0, im1, t_re0, t_re1, t_im0,
                                         t_im1, batch, s);
      top -= R;
    }
    if (e == hipSuccess)
      e = fft128_blocks<FWD, MULV>(L, Lg, re0, re1, im0, im1, t_re0, t_re1, t_im0, t_im1, batch, s);
  } else {
    // the blocks, then bits L .. Lg-1 in rounds of up to 3 from the bottom
    e = fft128_blocks<FWD, MULV>(L, Lg, re0, re1, im0, im1, t_re0, t_re1, t_im0, t_im1, batch, s);
    for (uint32_t lo = L; lo < Lg && e == hipSuccess;) {
      const uint32_t R = Lg - lo >= 3 ? 3 : Lg - lo;
      e = fft128_global_round<FWD, MULV>(R, Lg, lo + R - 1, re0, re1, im0, im1, t_re0, t_re1,
                                         t_im0, t_im1, batch, s);
      lo += R;
    }
  }
  return e;
}

}  // namespace

hipError_t fft128_run(bool fwd, int mulv, uint32_t L, double *re0, double *re1, double *im0,
                      double *im1, const double *t_re0, const double *t_re1,
                      const double *t_im0, const double *t_im1, uint64_t batch, hipStream_t s) {
  if (batch == 0) return hipSuccess;
  if (L < 3 || L > kFft128MaxLog2) return hipErrorInvalidValue;
  if (fwd)
    return mulv ? fft128_any<true, 1>(L, re0, re1, im0, im1, t_re0, t_re1, t_im0, t_im1, batch, s)
                : fft128_any<true, 0>(L, re0, re1, im0, im1, t_re0, t_re1, t_im0, t_im1, batch, s);
  return mulv ? fft128_any<false, 1>(L, re0, re1, im0, im1, t_re0, t_re1, t_im0, t_im1, batch, s)
              : fft128_any<false, 0>(L, re0, re1, im0, im1, t_re0, t_re1, t_im0, t_im1, batch, s);
}

// S: 16 * batch * n doubles; margin (batch doubles) may be null.  twiddles
// t: 4 arrays of n.
hipError_t exact128_polymul(uint32_t L, const uint64_t *a, const uint64_t *b, uint64_t *c,
                            double *margin, uint64_t batch, const double *t, double *S,
                            hipStream_t s) {
  if (batch == 0) return hipSuccess;
  const uint32_t n = 1u << L;
  const uint64_t total = batch * n;
  const dim3 grid((uint32_t)((total + 255) / 256)), block(256);
  const uint64_t plane = 4 * batch * (uint64_t)n;
  hipLaunchKernelGGL(k_exact128_split, grid, block, 0, s, a, b, S, margin, batch, L);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  e = fft128_run(true, 1, L, S, S + plane, S + 2 * plane, S + 3 * plane, t, t + n, t + 2 * n,
                 t + 3 * n, 4 * batch, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_exact128_pointwise, grid, block, 0, s, S, batch, L);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  e = fft128_run(false, 1, L, S, S + plane, S + 2 * plane, S + 3 * plane, t, t + n, t + 2 * n,
                 t + 3 * n, 2 * batch, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_exact128_combine, grid, block, 0, s, S, c, margin, batch, L);
  return hipGetLastError();
}

}  // namespace cfft_launch
