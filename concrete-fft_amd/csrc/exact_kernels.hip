// exact_kernels.hip -- exact u64 negacyclic product via 16-bit limbs.
//
// C = A * B mod (X^N + 1, 2^64).  Each operand is split into four balanced
// limbs in [-2^15, 2^15) (A = sum_i a_i 2^(16 i) mod 2^64), so that
//   C = sum_{k<4} 2^(16 k) D_k  mod 2^64,   D_k = sum_{i+j=k} a_i (x) b_j
// (terms with i + j >= 4 vanish mod 2^64).  |D_k| <= (k+1) N 2^30 < 2^45 for
// N <= 4096, so an f64 negacyclic transform product rounds to D_k exactly as
// long as its error stays below 1/2; the largest distance to the nearest
// integer seen while rounding a product is reported as its margin.
//
// Kernel 1: the 8 limb polynomials of every product are folded, twisted and
// transformed (any f64 plan) into a Fourier scratch, transform t = 8 p + l
// (l < 4: a_l, l >= 4: b_{l-4}).
// Kernel 2: per product and weight k, the inverse program's first pass reads
// sum_{i+j=k} F[a_i] F[b_j] (i ascending), the last pass untwists, rounds to
// an integer and accumulates 2^(16 k) D_k into C (owned by one workgroup).
// Dif16 plans at N = 2048 and 4096 have register-resident forms below (one
// fused kernel at N = 2048).
#include <hip/hip_runtime.h>

#include "lds_attr.hpp"

#include "launch.hpp"
#include "r16_exec.hpp"
#include "stockham_exec.hpp"

using namespace cfft_dev;
using cfft_launch::dc64;
using cfft_prog::Program;

namespace {

// balanced 16-bit limb `li` of v: the digits d_i in [-2^15, 2^15) with
// v = sum d_i 2^(16 i) mod 2^64 (the carry out of the top limb vanishes).
// With e_i = d_i + 2^15 in [0, 2^16), sum e_i 2^(16 i) = v + 0x8000...8000
// mod 2^64, so the e_i are the base-2^16 digits of that sum (the digits are
// unique); the same as propagating carries limb by limb.
__device__ __forceinline__ int64_t limb16(uint64_t v, uint32_t li) {
  const uint64_t e = v + 0x8000800080008000ull;
  return (int64_t)((e >> (16 * li)) & 0xffffu) - 32768;
}

// Nearest integer of x (|x| < 2^51) as a two's-complement u64, and the
// distance |x - that integer|: x + 1.5 2^52 lies in [2^52, 2^53), where the
// spacing is 1, so the addition rounds x to an integer (half to even) in the
// low mantissa bits, and the two bit patterns differ by exactly that
// integer.  (Differs from round-half-away only at exact half-integers, where
// the distance is 1/2 and the product's margin reports it.)
__device__ __forceinline__ uint64_t nearest_u64(double x, double &dist) {
  constexpr double kMagic = 6755399441055744.0;  // 1.5 2^52
  const double xm = x + kMagic;
  dist = fabs(x - (xm - kMagic));
  return __builtin_bit_cast(uint64_t, xm) - __builtin_bit_cast(uint64_t, kMagic);
}

template <int E>
__global__ void __launch_bounds__(512) k_exact_fwd(const uint64_t *__restrict__ A,
                                                   const uint64_t *__restrict__ B,
                                                   c64 *__restrict__ F, double *margin,
                                                   uint64_t batch, const c64 *__restrict__ tw,
                                                   const c64 *__restrict__ twist, Program fwd) {
  extern __shared__ c64 smem[];
  const uint32_t n = 1u << fwd.log2n;
  const uint32_t T = n / E;
  const uint32_t G = blockDim.x / T;
  const uint32_t g = threadIdx.x / T, lt = threadIdx.x % T;
  const uint64_t t = (uint64_t)blockIdx.x * G + g;  // transform index
  const bool active = t < batch * 8;
  const uint64_t p = t >> 3;
  const uint32_t li = (uint32_t)(t & 7);
  c64 *lds = smem + g * lds_words(n);
  const uint64_t *src = (li < 4 ? A : B) + p * (2ull * n);
  const uint32_t lb = li & 3;
  if (active && li == 0 && lt == 0 && margin) margin[p] = 0.0;
  auto load = [&](uint32_t j) -> c64 {
    if (!active) return mk(0.0, 0.0);
    return cmul(twist[j], mk((double)(int32_t)limb16(src[j], lb), (double)(int32_t)limb16(src[j + n], lb)));
  };
  c64 *out = F + t * n;
  auto store = [&](uint32_t pos, c64 v) {
    if (active) out[pos] = v;
  };
  run_program<E>(fwd, lds, lt, T, tw, load, store);
}

// One block per product, one transform group per weight k (4 groups): group
// k runs the inverse of sum_{i+j=k} F[a_i] F[b_j], rounds, shifts by 16 k
// and leaves its 2n words in its own LDS region (the last pass's LDS reads
// are complete by then); after the final barrier all threads sum the four
// regions mod 2^64 and store C.
template <int E>
__global__ void __launch_bounds__(512) k_exact_inv(const c64 *__restrict__ F,
                                                   uint64_t *__restrict__ C, double *margin,
                                                   uint64_t batch, const c64 *__restrict__ tw,
                                                   const c64 *__restrict__ untwist, Program inv) {
  extern __shared__ c64 smem[];
  const uint32_t n = 1u << inv.log2n;
  const uint32_t T = n / E;
  const uint32_t k = threadIdx.x / T, lt = threadIdx.x % T;
  const uint64_t p = blockIdx.x;
  const uint32_t words = lds_words(n);
  c64 *lds = smem + k * words;
  uint64_t *part = reinterpret_cast<uint64_t *>(lds);  // 2n words fit in `words` points
  const c64 *fp = F + p * 8ull * n;
  double worst = 0.0;
  auto load = [&](uint32_t pos) -> c64 {
    c64 acc = cmul(fp[0 * n + pos], fp[(4 + k) * (size_t)n + pos]);
    for (uint32_t i = 1; i <= k; ++i) {
      const c64 pr = cmul(fp[(size_t)i * n + pos], fp[(size_t)(4 + k - i) * n + pos]);
      acc = cadd(acc, pr);
    }
    return acc;
  };
  auto store = [&](uint32_t pos, c64 v) {
    const c64 r = cmul(untwist[pos], v);
    double d0, d1;
    const uint64_t x0 = nearest_u64(r.re, d0), x1 = nearest_u64(r.im, d1);
    worst = fmax(worst, fmax(d0, d1));
    part[pos] = x0 << (16 * k);
    part[n + pos] = x1 << (16 * k);
  };
  run_program<E>(inv, lds, lt, T, tw, load, store);  // ends with a barrier
  const uint64_t *p0 = reinterpret_cast<const uint64_t *>(smem);
  const uint64_t *p1 = reinterpret_cast<const uint64_t *>(smem + words);
  const uint64_t *p2 = reinterpret_cast<const uint64_t *>(smem + 2 * words);
  const uint64_t *p3 = reinterpret_cast<const uint64_t *>(smem + 3 * words);
  uint64_t *c = C + p * (2ull * n);
  for (uint32_t q = threadIdx.x; q < 2 * n; q += blockDim.x)
    st_stream(p0[q] + p1[q] + p2[q] + p3[q], &c[q]);
  // one atomic per transform group per wave: max over the group's lanes first
  // (non-negative doubles order like their bit patterns)
  if (margin) {
    const uint32_t seg = T < 64 ? T : 64;
    for (uint32_t off = 1; off < seg; off <<= 1) worst = fmax(worst, __shfl_xor(worst, (int)off));
    if ((threadIdx.x & (seg - 1)) == 0)
      atomicMax(reinterpret_cast<unsigned long long *>(margin + p),
                (unsigned long long)__double_as_longlong(worst));
  }
}

// ---- register-resident form for Dif16 plans with n = 1024 (N = 2048) ----
// One fused kernel, two products per 8-wave workgroup sharing the twiddle
// block staged in LDS.  Forward: one wave per (product, limb index i)
// transforms a_i and b_i together (wave-local Dif16 schedule, twiddles
// shared).  The product's four waves then meet in their exchange buffers,
// half of the slots at a time, and each lane forms the four weight spectra
// W_k = sum_{i<=k} F[a_i] F[b_{k-i}] (i ascending) of one point of each
// product, keeping them (two halves x two products x four weights) in
// registers until every limb spectrum is read; then it writes them into the
// exchange buffers, W_k of product pp into wave 4 pp + k's buffer at its
// point index.  Inverse: wave 4 pp + k transforms W_k back from LDS, rounds,
// shifts by 16 k and leaves its parts in its buffer; after a barrier the
// product's first wave sums the four buffers mod 2^64.  No scratch.
constexpr uint32_t kXW = 8;
constexpr uint32_t kXTw = 16 * 64 + 16 * 4;
constexpr size_t kXLds = (kXTw + kXW * R16<10>::lds_words) * sizeof(c64);

__global__ void __launch_bounds__(64 * kXW) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_exact_r16(const uint64_t *__restrict__ A, const uint64_t *__restrict__ B,
            uint64_t *__restrict__ C, double *margin, uint64_t batch,
            const c64 *__restrict__ tw_r16, const c64 *__restrict__ twist) {
  using S = R16<10>;
  constexpr uint32_t n = S::n;
  extern __shared__ c64 smem[];
  c64 *ltw = smem;
  for (uint32_t i = threadIdx.x; i < kXTw; i += blockDim.x) ltw[i] = tw_r16[i];
  double *s_worst = reinterpret_cast<double *>(smem + kXTw + kXW * S::lds_words);
  __syncthreads();
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint32_t pp = wave >> 2, li = wave & 3;  // product in the workgroup; limb, then weight
  const uint64_t p = (uint64_t)blockIdx.x * 2 + pp;
  const bool active = p < batch;  // every wave reaches the barriers below
  c64 *x = smem + kXTw + wave * S::lds_words;
  const uint32_t l = threadIdx.x & 63u;
  c64 fa[16], fb[16];
  if (active) {
    const uint64_t *a = A + p * (2 * n);
    const uint64_t *b = B + p * (2 * n);
#pragma unroll
    for (uint32_t m = 0; m < 16; ++m) {
      const uint32_t j = l + S::T * m;
      const c64 z = twist[j];
      fa[m] = cmul(z, mk((double)(int32_t)limb16(ld_stream(&a[j]), li),
                         (double)(int32_t)limb16(ld_stream(&a[j + n]), li)));
      fb[m] = cmul(z, mk((double)(int32_t)limb16(ld_stream(&b[j]), li),
                         (double)(int32_t)limb16(ld_stream(&b[j + n]), li)));
    }
    r16w_transform<true, 2>(fa, fb, x, l, ltw);
  }
  const c64 *xs = smem + kXTw;
  const uint32_t t = threadIdx.x;
  c64 wq[2][2][4];  // [half][product][weight] at this thread's point of the half
#pragma unroll
  for (uint32_t h = 0; h < 2; ++h) {
    wave_lds_fence();
    if (active) {
#pragma unroll
      for (uint32_t s = 0; s < 8; ++s) {
        x[s * 64 + l] = fa[8 * h + s];
        x[512 + s * 64 + l] = fb[8 * h + s];
      }
    }
    __syncthreads();
#pragma unroll
    for (uint32_t q2 = 0; q2 < 2; ++q2) {
      c64 wa[4], wb[4];
#pragma unroll
      for (uint32_t i = 0; i < 4; ++i) {
        wa[i] = xs[(4 * q2 + i) * S::lds_words + t];
        wb[i] = xs[(4 * q2 + i) * S::lds_words + 512 + t];
      }
#pragma unroll
      for (uint32_t k = 0; k < 4; ++k) {
        c64 d = cmul(wa[0], wb[k]);
#pragma unroll
        for (uint32_t i = 1; i <= k; ++i) d = cadd(d, cmul(wa[i], wb[k - i]));
        wq[h][q2][k] = d;
      }
    }
    __syncthreads();  // every limb spectrum of the half read
  }
  // W_k of product q2, point q, to wave 4 q2 + k's buffer at q
#pragma unroll
  for (uint32_t h = 0; h < 2; ++h) {
    const uint32_t q = (t & 63) + S::T * S::end_m(8 * h + (t >> 6));
#pragma unroll
    for (uint32_t q2 = 0; q2 < 2; ++q2)
#pragma unroll
      for (uint32_t k = 0; k < 4; ++k) smem[kXTw + (4 * q2 + k) * S::lds_words + q] = wq[h][q2][k];
  }
  __syncthreads();
  // inverse of W_k, k = li: 2^(16 k) D_k, rounded
  const uint32_t k = li;
  const double inv_n = 1.0 / (double)n;
  double worst = 0.0;
  uint64_t part[32];
  {
    c64 v[16];
#pragma unroll
    for (uint32_t m = 0; m < 16; ++m) v[m] = x[l + S::T * m];
    r16w_transform<false, 1>(v, v, x, l, ltw);
    const c64 *tws = launder(twist);
#pragma unroll
    for (uint32_t s = 0; s < 16; ++s) {
      const uint32_t j = l + S::T * S::end_m(s);
      const c64 tz = tws[j];
      const c64 r = cmul(mk(tz.re * inv_n, (-tz.im) * inv_n), v[s]);
      double d0, d1;
      const uint64_t x0 = nearest_u64(r.re, d0), x1 = nearest_u64(r.im, d1);
      worst = fmax(worst, fmax(d0, d1));
      part[2 * s] = x0 << (16 * k);
      part[2 * s + 1] = x1 << (16 * k);
    }
  }
  // the transform's last LDS reads are behind its final wave fence; reuse the
  // buffer for this wave's 32 words per lane
  uint64_t *xb = reinterpret_cast<uint64_t *>(x);
#pragma unroll
  for (uint32_t q = 0; q < 32; ++q) xb[q * 64 + l] = part[q];
  for (uint32_t off = 1; off < 64; off <<= 1) worst = fmax(worst, __shfl_xor(worst, (int)off));
  if (l == 0) s_worst[wave] = worst;
  __syncthreads();
  if (active && k == 0) {
    const uint64_t *b1 = reinterpret_cast<const uint64_t *>(x + 1 * S::lds_words);
    const uint64_t *b2 = reinterpret_cast<const uint64_t *>(x + 2 * S::lds_words);
    const uint64_t *b3 = reinterpret_cast<const uint64_t *>(x + 3 * S::lds_words);
    uint64_t *c = C + p * (2 * n);
#pragma unroll
    for (uint32_t s = 0; s < 16; ++s) {
      const uint32_t j = l + S::T * S::end_m(s);
      const uint32_t q0 = (2 * s) * 64 + l, q1 = (2 * s + 1) * 64 + l;
      st_stream(part[2 * s] + b1[q0] + b2[q0] + b3[q0], &c[j]);
      st_stream(part[2 * s + 1] + b1[q1] + b2[q1] + b3[q1], &c[j + n]);
    }
    if (margin && l == 0)
      margin[p] = fmax(fmax(s_worst[wave], s_worst[wave + 1]),
                       fmax(s_worst[wave + 2], s_worst[wave + 3]));
  }
}

// ---- wave-local form for Dif16 plans with n = 2048, base_n = 1024 (N = 4096) ----
// One 2048-point transform per wavefront (r16l_2048_transform: the plan's
// radix-2 depth pass in registers around two 1024-point base transforms run
// together).  Forward: one workgroup per product, one wave per limb
// transform; the eight limb spectra (same slot layout in every wave) meet in
// the waves' exchange buffers, one half at a time, and each lane forms the
// four weight spectra W_k = sum_{i<=k} F[a_i] F[b_{k-i}] (i ascending) of
// two points, so the scratch holds 4 spectra per product, written once and
// read once.  Inverse: one wave per (product, weight k) reading W_k; the
// rounded, shifted parts of the four weights meet in LDS, chunk 0's
// positions first, then chunk 1's, and each of the product's waves adds and
// stores a quarter of them.  (The forward keeps a few dozen VGPRs in
// scratch: the paired 1024-point transforms alone take 240 of the 256
// available.)
constexpr uint32_t kWN = 2048;

__global__ void __launch_bounds__(64 * kXW) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_exact_fwd_w2(const uint64_t *__restrict__ A, const uint64_t *__restrict__ B,
               c64 *__restrict__ F, double *margin, uint64_t batch, const c64 *__restrict__ tw,
               const c64 *__restrict__ tw_blk, const c64 *__restrict__ twist) {
  using S = R16<10>;
  constexpr uint32_t n = kWN;
  extern __shared__ c64 smem[];
  c64 *ltw = smem;
  for (uint32_t i = threadIdx.x; i < kXTw; i += blockDim.x) ltw[i] = tw_blk[i];
  __syncthreads();
  static_assert(kXW == 8, "one workgroup holds the 8 limb transforms of one product");
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint64_t u = (uint64_t)blockIdx.x * kXW + wave;  // 8 p + limb transform
  // the grid is one workgroup per product, so every wave is active; the
  // guard keeps every wave at the barriers below if that ever changes
  const bool active = u < batch * 8;
  c64 *x = smem + kXTw + wave * S::lds_words;
  const uint32_t l = threadIdx.x & 63u;
  const uint64_t p = active ? u >> 3 : 0;
  const uint32_t li = (uint32_t)(u & 7), lb = li & 3;
  if (active && li == 0 && l == 0 && margin) margin[p] = 0.0;
  const uint64_t *src = (li < 4 ? A : B) + p * (2ull * n);
  c64 va[16], vb[16];
  // each half (points j and 1024 + j): its coefficient loads first, then
  // limbs and twist four points at a time
#pragma unroll
  for (uint32_t h = 0; h < 2; ++h) {
    c64 *v = h == 0 ? va : vb;
    uint64_t r[32];
#pragma unroll
    for (uint32_t m = 0; m < 16; ++m) {
      const uint32_t j = h * 1024 + l + 64 * m;
      r[2 * m] = src[j];
      r[2 * m + 1] = src[j + n];
    }
    sched_fence();
#pragma unroll
    for (uint32_t m0 = 0; m0 < 16; m0 += 4) {
#pragma unroll
      for (uint32_t m = m0; m < m0 + 4; ++m)
        v[m] = cmul(twist[h * 1024 + l + 64 * m],
                    mk((double)(int32_t)limb16(r[2 * m], lb),
                       (double)(int32_t)limb16(r[2 * m + 1], lb)));
      sched_fence();
    }
  }
  r16l_2048_transform<true>(va, vb, x, l, ltw, tw);
  // weights: buffer w holds limb spectrum w of the product (a_w for w < 4,
  // b_{w-4} after), slot s of lane l at s * 64 + l; thread t forms the four
  // weights of buffer positions t and t + 512 (every wave of the workgroup
  // belongs to product p: the grid has one workgroup per product)
  const bool store = (uint64_t)blockIdx.x < batch;
  c64 *W = F + p * 4ull * n;
  const c64 *xs = smem + kXTw;
  const uint32_t t = threadIdx.x;
#pragma unroll
  for (uint32_t h = 0; h < 2; ++h) {
    const c64 *v = h == 0 ? va : vb;
    wave_lds_fence();
#pragma unroll
    for (uint32_t s = 0; s < 16; ++s) x[s * 64 + l] = v[s];
    __syncthreads();
#pragma unroll
    for (uint32_t e = 0; e < 2; ++e) {
      const uint32_t idx = t + 512 * e;
      const uint32_t q = h * 1024 + (idx & 63) + 64 * S::end_m(idx >> 6);
      c64 fa[4], fb[4];
#pragma unroll
      for (uint32_t i = 0; i < 4; ++i) {
        fa[i] = xs[i * S::lds_words + idx];
        fb[i] = xs[(4 + i) * S::lds_words + idx];
      }
#pragma unroll
      for (uint32_t k = 0; k < 4; ++k) {
        c64 d = cmul(fa[0], fb[k]);
#pragma unroll
        for (uint32_t i = 1; i <= k; ++i) d = cadd(d, cmul(fa[i], fb[k - i]));
        if (store) st_stream(d, &W[k * n + q]);
      }
    }
    __syncthreads();  // the buffers are rewritten for the second half
  }
}

__global__ void __launch_bounds__(64 * kXW) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_exact_inv_w2(const c64 *__restrict__ F, uint64_t *__restrict__ C, double *margin,
               uint64_t batch, const c64 *__restrict__ tw, const c64 *__restrict__ tw_blk,
               const c64 *__restrict__ untwist) {
  using S = R16<10>;
  constexpr uint32_t n = kWN;
  extern __shared__ c64 smem[];
  c64 *ltw = smem;
  for (uint32_t i = threadIdx.x; i < kXTw; i += blockDim.x) ltw[i] = tw_blk[i];
  __syncthreads();
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint64_t p = (uint64_t)blockIdx.x * (kXW / 4) + (wave >> 2);
  const uint32_t k = wave & 3;
  const bool active = p < batch;  // every wave reaches the barriers below
  c64 *x = smem + kXTw + wave * S::lds_words;
  const uint32_t l = threadIdx.x & 63u;
  const c64 *wk = F + (active ? p : 0) * 4ull * n + k * n;  // weight spectrum W_k
  c64 va[16], vb[16];
#pragma unroll
  for (uint32_t m = 0; m < 16; ++m) {
    const uint32_t q = l + 64 * m;
    va[m] = ld_stream(&wk[q]);
    vb[m] = ld_stream(&wk[1024 + q]);
  }
  r16l_2048_transform<false>(va, vb, x, l, ltw, tw);
  // the transform's last LDS reads are behind its final wave fence; the
  // wave's buffer now carries 32 words per lane, one half at a time, each
  // word written as soon as it is rounded
  uint64_t *xb = reinterpret_cast<uint64_t *>(x);
  uint64_t *cp = C + (active ? p : 0) * (2ull * n);
  double worst = 0.0;
#pragma unroll
  for (uint32_t h = 0; h < 2; ++h) {
    const c64 *v = h == 0 ? va : vb;
#pragma unroll
    for (uint32_t s = 0; s < 16; ++s) {
      const uint32_t q = h * 1024 + l + 64 * S::end_m(s);
      const c64 r = cmul(untwist[q], v[s]);
      double d0, d1;
      const uint64_t x0 = nearest_u64(r.re, d0), x1 = nearest_u64(r.im, d1);
      worst = fmax(worst, fmax(d0, d1));
      xb[(2 * s) * 64 + l] = x0 << (16 * k);
      xb[(2 * s + 1) * 64 + l] = x1 << (16 * k);
      if ((s & 3) == 3) sched_fence();
    }
    // the margin is reduced here, not sunk to its use at the end (which
    // would keep the half's rounding residues live, and spilled)
    asm volatile("" : "+v"(worst));
    __syncthreads();
    if (active) {
      // wave k adds and stores slots 4k .. 4k + 3 of the four weights' words
      // (lane index recomputed: store addresses kept live across the first
      // half would spill)
      const uint32_t lh = launder_u32(threadIdx.x) & 63u;
      const uint64_t *b0 = reinterpret_cast<const uint64_t *>(x - k * S::lds_words);
      const uint32_t stride = S::lds_words * 2;  // one buffer, in words
#pragma unroll
      for (uint32_t e = 0; e < 4; ++e) {
        const uint32_t s = 4 * k + e;
        const uint32_t q = h * 1024 + lh + 64 * S::end_m(s);
        const uint32_t w0 = (2 * s) * 64 + lh, w1 = (2 * s + 1) * 64 + lh;
        st_stream(b0[w0] + b0[stride + w0] + b0[2 * stride + w0] + b0[3 * stride + w0], &cp[q]);
        st_stream(b0[w1] + b0[stride + w1] + b0[2 * stride + w1] + b0[3 * stride + w1],
                  &cp[q + n]);
      }
    }
    __syncthreads();  // the buffers are rewritten for the second half
  }
  if (margin) {
    for (uint32_t off = 1; off < 64; off <<= 1) worst = fmax(worst, __shfl_xor(worst, (int)off));
    if (active && l == 0)
      atomicMax(reinterpret_cast<unsigned long long *>(margin + p),
                (unsigned long long)__double_as_longlong(worst));
  }
}

hipError_t set_lds(const void *kernel, size_t bytes) {
  return cfft_launch::ensure_lds(kernel, bytes);
}

struct Geom {
  int E;
  uint32_t T, G, threads;
  size_t lds;
};
Geom geom(uint32_t log2n) {
  Geom g;
  const uint32_t n = 1u << log2n;
  g.E = n >= 16 ? 16 : (int)n;
  g.T = n / (uint32_t)g.E;
  g.G = n >= 2048 ? 1 : 2048 / n;
  if (g.T * g.G > 512) g.G = 512 / g.T > 0 ? 512 / g.T : 1;
  g.threads = g.T * g.G;
  g.lds = (size_t)g.G * lds_words(n) * sizeof(c64);
  return g;
}

template <int E>
hipError_t launch_e(const Program &fwd, const Program &inv, const uint64_t *a, const uint64_t *b,
                    uint64_t *c, double *margin, uint64_t batch, const c64 *tw,
                    const c64 *twist, const c64 *untwist, c64 *F, hipStream_t s) {
  const Geom g = geom(fwd.log2n);
  const uint32_t n = 1u << fwd.log2n;
  const uint32_t T = n / (uint32_t)E;
  const size_t inv_lds = 4 * (size_t)lds_words(n) * sizeof(c64);
  hipError_t e = set_lds(reinterpret_cast<const void *>(k_exact_fwd<E>), g.lds);
  if (e == hipSuccess) e = set_lds(reinterpret_cast<const void *>(k_exact_inv<E>), inv_lds);
  if (e != hipSuccess) return e;
  const uint64_t nt = batch * 8;
  hipLaunchKernelGGL(k_exact_fwd<E>, dim3((uint32_t)((nt + g.G - 1) / g.G)), dim3(g.threads),
                     g.lds, s, a, b, F, margin, batch, tw, twist, fwd);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_exact_inv<E>, dim3((uint32_t)batch), dim3(4 * T), inv_lds, s, F, c, margin,
                     batch, tw, untwist, inv);
  return hipGetLastError();
}

}  // namespace

namespace cfft_launch {

// F: scratch of batch * 8 * n points.  margin may be null.
hipError_t exact_polymul(const Program &fwd, const Program &inv, const uint64_t *a,
                         const uint64_t *b, uint64_t *c, double *margin, uint64_t batch,
                         const dc64 *tw, const dc64 *twist, const dc64 *untwist, dc64 *F,
                         hipStream_t s) {
  if (batch == 0) return hipSuccess;
  const c64 *t = reinterpret_cast<const c64 *>(tw);
  const c64 *tws = reinterpret_cast<const c64 *>(twist);
  const c64 *utw = reinterpret_cast<const c64 *>(untwist);
  c64 *f = reinterpret_cast<c64 *>(F);
  switch (geom(fwd.log2n).E) {
    case 2: return launch_e<2>(fwd, inv, a, b, c, margin, batch, t, tws, utw, f, s);
    case 4: return launch_e<4>(fwd, inv, a, b, c, margin, batch, t, tws, utw, f, s);
    case 8: return launch_e<8>(fwd, inv, a, b, c, margin, batch, t, tws, utw, f, s);
    default: return launch_e<16>(fwd, inv, a, b, c, margin, batch, t, tws, utw, f, s);
  }
}

// Register-resident form: Dif16 plan with n = 1024; tw_r16 is the plan's
// lane-major twiddle block.  One fused kernel, two products per workgroup,
// no scratch.
hipError_t exact_polymul_r16(const uint64_t *a, const uint64_t *b, uint64_t *c, double *margin,
                             uint64_t batch, const dc64 *tw_r16, const dc64 *twist,
                             hipStream_t s) {
  if (batch == 0) return hipSuccess;
  constexpr size_t lds = kXLds + kXW * sizeof(double);  // + each wave's margin
  hipError_t e = cfft_launch::ensure_lds(reinterpret_cast<const void *>(k_exact_r16), lds);
  if (e != hipSuccess) return e;
  const c64 *t = reinterpret_cast<const c64 *>(tw_r16);
  const c64 *z = reinterpret_cast<const c64 *>(twist);
  const uint64_t per_wg = kXW / 4;  // products per workgroup (4 waves each)
  hipLaunchKernelGGL(k_exact_r16, dim3((uint32_t)((batch + per_wg - 1) / per_wg)), dim3(64 * kXW),
                     lds, s, a, b, c, margin, batch, t, z);
  return hipGetLastError();
}

// Wave-local form: Dif16 plan with n = 2048 and base_n = 1024; tw is the
// plan's full table, tw_blk its 1024-point base block.
hipError_t exact_polymul_w2(const uint64_t *a, const uint64_t *b, uint64_t *c, double *margin,
                            uint64_t batch, const dc64 *tw, const dc64 *tw_blk,
                            const dc64 *twist, const dc64 *untwist, dc64 *F, hipStream_t s) {
  if (batch == 0) return hipSuccess;
  hipError_t e = cfft_launch::ensure_lds(reinterpret_cast<const void *>(k_exact_fwd_w2), kXLds);
  if (e == hipSuccess)
    e = cfft_launch::ensure_lds(reinterpret_cast<const void *>(k_exact_inv_w2), kXLds);
  if (e != hipSuccess) return e;
  const c64 *t = reinterpret_cast<const c64 *>(tw);
  const c64 *tb = reinterpret_cast<const c64 *>(tw_blk);
  c64 *f = reinterpret_cast<c64 *>(F);
  const uint64_t transforms = batch * 8;
  hipLaunchKernelGGL(k_exact_fwd_w2, dim3((uint32_t)((transforms + kXW - 1) / kXW)),
                     dim3(64 * kXW), kXLds, s, a, b, f, margin, batch, t, tb,
                     reinterpret_cast<const c64 *>(twist));
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  const uint64_t per_wg = kXW / 4;  // products per workgroup (4 waves each)
  hipLaunchKernelGGL(k_exact_inv_w2, dim3((uint32_t)((batch + per_wg - 1) / per_wg)),
                     dim3(64 * kXW), kXLds, s, f, c, margin, batch, t, tb,
                     reinterpret_cast<const c64 *>(untwist));
  return hipGetLastError();
}

}  // namespace cfft_launch
