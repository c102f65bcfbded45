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
#include <hip/hip_runtime.h>

#include <mutex>
#include <set>

#include "launch.hpp"
#include "stockham_exec.hpp"

using namespace cfft_dev;
using cfft_launch::dc64;
using cfft_prog::Program;

namespace {

// balanced 16-bit limb `li` of v (carry from the lower limbs; the carry out of
// the top limb vanishes mod 2^64)
__device__ __forceinline__ int64_t limb16(uint64_t v, uint32_t li) {
  int64_t carry = 0, d = 0;
#pragma unroll
  for (uint32_t q = 0; q < 4; ++q) {
    d = (int64_t)((v >> (16 * q)) & 0xffffu) + carry;
    carry = d >= 32768 ? 1 : 0;
    d -= carry << 16;
    if (q == li) break;
  }
  return d;
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

template <int E>
__global__ void __launch_bounds__(512) k_exact_inv(const c64 *__restrict__ F,
                                                   uint64_t *__restrict__ C, double *margin,
                                                   uint64_t batch, const c64 *__restrict__ tw,
                                                   const c64 *__restrict__ untwist, Program inv) {
  extern __shared__ c64 smem[];
  const uint32_t n = 1u << inv.log2n;
  const uint32_t T = n / E;
  const uint32_t G = blockDim.x / T;
  const uint32_t g = threadIdx.x / T, lt = threadIdx.x % T;
  const uint64_t p = (uint64_t)blockIdx.x * G + g;
  const bool active = p < batch;
  c64 *lds = smem + g * lds_words(n);
  const c64 *fp = F + p * 8ull * n;
  uint64_t *c = C + p * (2ull * n);
  double worst = 0.0;
  for (uint32_t k = 0; k < 4; ++k) {
    auto load = [&](uint32_t pos) -> c64 {
      if (!active) return mk(0.0, 0.0);
      c64 acc = cmul(fp[0 * n + pos], fp[(4 + k) * (size_t)n + pos]);
      for (uint32_t i = 1; i <= k; ++i) {
        const c64 pr = cmul(fp[(size_t)i * n + pos], fp[(size_t)(4 + k - i) * n + pos]);
        acc = cadd(acc, pr);
      }
      return acc;
    };
    auto store = [&](uint32_t pos, c64 v) {
      if (!active) return;
      const c64 r = cmul(untwist[pos], v);
      const double x0 = round(r.re), x1 = round(r.im);
      worst = fmax(worst, fmax(fabs(r.re - x0), fabs(r.im - x1)));
      const uint64_t d0 = (uint64_t)(int64_t)x0 << (16 * k);
      const uint64_t d1 = (uint64_t)(int64_t)x1 << (16 * k);
      if (k == 0) {
        c[pos] = d0;
        c[pos + n] = d1;
      } else {
        c[pos] += d0;
        c[pos + n] += d1;
      }
    };
    run_program<E>(inv, lds, lt, T, tw, load, store);
  }
  // one atomic per transform group per wave: max over the group's lanes first
  // (non-negative doubles order like their bit patterns)
  if (margin) {
    const uint32_t seg = T < 64 ? T : 64;
    for (uint32_t off = 1; off < seg; off <<= 1) worst = fmax(worst, __shfl_xor(worst, (int)off));
    if (active && (threadIdx.x & (seg - 1)) == 0)
      atomicMax(reinterpret_cast<unsigned long long *>(margin + p),
                (unsigned long long)__double_as_longlong(worst));
  }
}

hipError_t set_lds(const void *kernel, size_t bytes) {
  static std::mutex mu;
  static std::set<const void *> done;
  if (bytes <= 64 * 1024) return hipSuccess;
  std::lock_guard<std::mutex> gd(mu);
  if (done.count(kernel)) return hipSuccess;
  int dev = 0, maxb = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&maxb, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
  if (e != hipSuccess) return e;
  if ((size_t)maxb < bytes) return hipErrorInvalidValue;
  e = hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, maxb);
  if (e != hipSuccess) return e;
  done.insert(kernel);
  return hipSuccess;
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
  hipError_t e = set_lds(reinterpret_cast<const void *>(k_exact_fwd<E>), g.lds);
  if (e == hipSuccess) e = set_lds(reinterpret_cast<const void *>(k_exact_inv<E>), g.lds);
  if (e != hipSuccess) return e;
  const uint64_t nt = batch * 8;
  hipLaunchKernelGGL(k_exact_fwd<E>, dim3((uint32_t)((nt + g.G - 1) / g.G)), dim3(g.threads),
                     g.lds, s, a, b, F, margin, batch, tw, twist, fwd);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_exact_inv<E>, dim3((uint32_t)((batch + g.G - 1) / g.G)), dim3(g.threads),
                     g.lds, s, F, c, margin, batch, tw, untwist, inv);
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

}  // namespace cfft_launch
