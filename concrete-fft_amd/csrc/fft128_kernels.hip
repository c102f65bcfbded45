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

#include <mutex>
#include <set>

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
// FWD: stages st0, st0+1, .. (bit ptop = L-1-st0 downwards), CT butterflies.
// INV: stages processed from the low bit up, GS butterflies.
template <int R, bool FWD, int MULV>
__device__ __forceinline__ void round_regs(cdd *z, uint32_t g, uint32_t ptop, uint32_t L,
                                           const Tw &tw) {
  // z[u], u < 2^R, indexed with the bit for position ptop-b at bit (R-1-b)
  const uint32_t base = group_pos<R>(g, ptop, 0);
  if (FWD) {
#pragma unroll
    for (int b = 0; b < R; ++b) {
      const uint32_t p = ptop - (uint32_t)b;       // pair bit
      const uint32_t st = L - 1 - p;               // stage index, m = 2^st
#pragma unroll
      for (int u = 0; u < (1 << R); ++u) {
        if ((u >> (R - 1 - b)) & 1) continue;
        const int v = u | (1 << (R - 1 - b));
        uint32_t j = base;
#pragma unroll
        for (int c = 0; c < R; ++c) j |= ((uint32_t)(u >> (R - 1 - c)) & 1u) << (ptop - c);
        const uint32_t i = j >> (p + 1);
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
      const uint32_t st = L - 1 - p;
#pragma unroll
      for (int u = 0; u < (1 << R); ++u) {
        if ((u >> (R - 1 - b)) & 1) continue;
        const int v = u | (1 << (R - 1 - b));
        uint32_t j = base;
#pragma unroll
        for (int c = 0; c < R; ++c) j |= ((uint32_t)(u >> (R - 1 - c)) & 1u) << (ptop - c);
        const uint32_t i = j >> (p + 1);
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
                                          uint32_t ptop, uint32_t L, uint32_t lt, uint32_t T,
                                          const Tw &tw, bool active) {
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
    round_regs<R, FWD, MULV>(z[q], g, ptop, L, tw);
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

// One workgroup per transform of n = 2^L points (L >= 3), T = n / 8 threads.
template <bool FWD, int MULV>
__global__ void __launch_bounds__(512) k_fft128(double *re0, double *re1, double *im0,
                                                double *im1, const double *__restrict__ t_re0,
                                                const double *__restrict__ t_re1,
                                                const double *__restrict__ t_im0,
                                                const double *__restrict__ t_im1, uint32_t L,
                                                uint64_t batch) {
  extern __shared__ double lds[];
  const uint32_t n = 1u << L;
  const uint32_t T = n / 8;
  const uint64_t tr = blockIdx.x;
  const bool active = tr < batch;
  const uint32_t lt = threadIdx.x;
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
      case 3: run_round<3, FWD, MULV>(lds, g, first, last, ptop, L, lt, T, tw, active); break;
      case 2: run_round<2, FWD, MULV>(lds, g, first, last, ptop, L, lt, T, tw, active); break;
      default: run_round<1, FWD, MULV>(lds, g, first, last, ptop, L, lt, T, tw, active); break;
    }
  }
}

hipError_t set_lds(const void *kernel, size_t bytes) {
  static std::mutex mu;
  static std::set<const void *> done;
  if (bytes <= 64 * 1024) return hipSuccess;
  std::lock_guard<std::mutex> g(mu);
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

}  // namespace

namespace cfft_launch {

// Largest fft128 transform held in one workgroup: 4096 points = 136 KiB LDS.
constexpr uint32_t kFft128MaxLog2 = 12;

hipError_t fft128_run(bool fwd, int mulv, uint32_t L, double *re0, double *re1, double *im0,
                      double *im1, const double *t_re0, const double *t_re1,
                      const double *t_im0, const double *t_im1, uint64_t batch, hipStream_t s) {
  if (batch == 0) return hipSuccess;
  if (L < 3 || L > kFft128MaxLog2) return hipErrorInvalidValue;
  const uint32_t n = 1u << L;
  const size_t lds = 4 * (size_t)pad128(n) * sizeof(double);
  const dim3 grid((uint32_t)batch), block(n / 8);
  const void *k;
  hipError_t e;
#define CFFT128_LAUNCH(F, M)                                                                   \
  k = reinterpret_cast<const void *>(k_fft128<F, M>);                                          \
  e = set_lds(k, lds);                                                                         \
  if (e != hipSuccess) return e;                                                               \
  hipLaunchKernelGGL((k_fft128<F, M>), grid, block, lds, s, re0, re1, im0, im1, t_re0, t_re1, \
                     t_im0, t_im1, L, batch);
  if (fwd) {
    if (mulv) { CFFT128_LAUNCH(true, 1) } else { CFFT128_LAUNCH(true, 0) }
  } else {
    if (mulv) { CFFT128_LAUNCH(false, 1) } else { CFFT128_LAUNCH(false, 0) }
  }
#undef CFFT128_LAUNCH
  return hipGetLastError();
}

}  // namespace cfft_launch
