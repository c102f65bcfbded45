// polymul_n2048_kernels.hip -- the fused negacyclic product at N = 2048
// (Dif16 plan with n = base_n = 1024): a persistent kernel with one product
// per wave.
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>

#include "launch.hpp"
#include "lds_attr.hpp"
#include "r16_exec.hpp"

using namespace cfft_dev;
using cfft_launch::dc64;

namespace {

// Persistent product kernel for n = 1024 (N = 2048): 8 waves per workgroup,
// each running its own stream of products.  LDS holds the r16 twiddle rows
// k = 1..15 (1020 points), the twist table (1024) and each wave's unpadded
// 1024-point exchange buffer: 163 776 bytes, one workgroup per CU.
//
// The kernel runs at 2 waves per SIMD (up to 256 VGPRs: a product's two
// spectra are live together) and, with HBM and the FP64 units both busy, at
// the device's power limit (shader clock measured 1.9-2.1 GHz against
// 2.4 GHz for its loads alone or its arithmetic alone), so work per product
// is what its time follows:
//  * twist and untwist factors come from LDS: no global load is issued after
//    the prefetch below, so its latency is never waited for early (vector
//    memory counters complete in order);
//  * the 1/n of the inverse transform is folded into A's conversion
//    (power-of-two scalings are exact, the untwist is then conj(zeta_j)
//    itself, the same bits);
//  * B of the wave's next product is loaded during the inverse transform;
//    past its last product a wave reads the L2-resident twist table instead
//    (unconditional loads keep the registers unspilled; no HBM traffic).
constexpr uint32_t kPW = 8;  // waves per workgroup
constexpr uint32_t kPmTw1 = 15 * 64, kPmTw2 = 15 * 4, kPmTwist = 1024, kPmX = 1024;
constexpr uint32_t kPmHead = kPmTw1 + kPmTw2 + kPmTwist;
constexpr size_t kPersistLds = (size_t)(kPmHead + kPW * kPmX) * sizeof(c64);

// (double)(int64)a * 2^-64 / 1024 with one rounding (2^-74 and 2^-42 scale
// exactly)
__device__ __forceinline__ double torus_to_f64_over_n1024(uint64_t a) {
  const double hi = (double)(int32_t)(a >> 32);
  const double lo = (double)(uint32_t)a * 0x1p-74;
  return __fma_rn(hi, 0x1p-42, lo);
}

__global__ void __launch_bounds__(64 * kPW) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_r16_polymul_persistent(const uint64_t *__restrict__ A, const int64_t *__restrict__ B,
                         uint64_t *__restrict__ C, uint64_t batch,
                         const c64 *__restrict__ tw_r16, const c64 *__restrict__ twist) {
  using S = R16<10>;
  constexpr uint32_t n = S::n;
  extern __shared__ c64 smem[];
  c64 *ltw1 = smem, *ltw2 = smem + kPmTw1, *ltz = smem + kPmTw1 + kPmTw2;
  for (uint32_t i = threadIdx.x; i < kPmTw1; i += blockDim.x) ltw1[i] = tw_r16[64 + i];
  for (uint32_t i = threadIdx.x; i < kPmTw2; i += blockDim.x) ltw2[i] = tw_r16[1024 + 4 + i];
  for (uint32_t i = threadIdx.x; i < kPmTwist; i += blockDim.x) ltz[i] = twist[i];
  __syncthreads();
  // uniform per wave: keeps product indices and base addresses in SGPRs
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  c64 *x = smem + kPmHead + wave * kPmX;
  const uint64_t step = (uint64_t)gridDim.x * kPW;
  uint64_t poly = (uint64_t)blockIdx.x * kPW + wave;
  uint64_t pb[32];  // raw words of B for the wave's current product
  if (poly < batch) {
    const uint32_t l = threadIdx.x & 63u;
    const int64_t *b = B + poly * (2 * n);
#pragma unroll
    for (uint32_t m = 0; m < 16; ++m) {
      pb[2 * m] = (uint64_t)ld_stream(&b[l + S::T * m]);
      pb[2 * m + 1] = (uint64_t)ld_stream(&b[l + S::T * m + n]);
    }
  }
  for (; poly < batch; poly += step) {
    // recomputed per product so per-lane addresses are not kept live across
    // the loop
    const uint32_t l = launder_u32(threadIdx.x) & 63u;
    const uint64_t *a = A + poly * (2 * n);
    uint64_t *c = C + poly * (2 * n);
    c64 fa[16], v[16];
    {
      uint64_t ra[32];
#pragma unroll
      for (uint32_t m = 0; m < 16; ++m) {
        const uint32_t j = l + S::T * m;
        ra[2 * m] = ld_stream(&a[j]);
        ra[2 * m + 1] = ld_stream(&a[j + n]);
      }
      sched_fence();
#pragma unroll
      for (uint32_t m = 0; m < 16; ++m) {
        const c64 z = ltz[l + S::T * m];
        fa[m] = cmul(z, mk(torus_to_f64_over_n1024(ra[2 * m]),
                           torus_to_f64_over_n1024(ra[2 * m + 1])));
        v[m] = cmul(z, mk(i64_to_f64((int64_t)pb[2 * m]), i64_to_f64((int64_t)pb[2 * m + 1])));
      }
    }
    r16x_transform<true, 2>(fa, v, x, l, ltw1, ltw2);
    c64 z[16];
#pragma unroll
    for (uint32_t s = 0; s < 16; ++s) fa[s] = cmul(fa[s], v[s]);
    r16_end_to_input<10>(fa, z);
    sched_fence();
    {
      const uint64_t nx = poly + step;
      const int64_t *bn = nx < batch ? B + nx * (2 * n) : reinterpret_cast<const int64_t *>(twist);
#pragma unroll
      for (uint32_t m = 0; m < 16; ++m) {
        pb[2 * m] = (uint64_t)ld_stream(&bn[l + S::T * m]);
        pb[2 * m + 1] = (uint64_t)ld_stream(&bn[l + S::T * m + n]);
      }
    }
    sched_fence();
    r16x_transform<false, 1>(z, z, x, l, ltw1, ltw2);
#pragma unroll
    for (uint32_t s = 0; s < 16; ++s) {
      const uint32_t j = l + S::T * S::end_m(s);
      const c64 r = cmul(conj(ltz[j]), z[s]);
      st_stream(from_torus(r.re), &c[j]);
      st_stream(from_torus(r.im), &c[j + n]);
      if ((s & 3) == 3) sched_fence();
    }
  }
}

hipError_t launch_polymul_persistent(const uint64_t *a, const int64_t *b, uint64_t *c,
                                    uint64_t batch, const dc64 *tw, const dc64 *twist,
                                    hipStream_t s) {
  // per device: compute-unit count and the kernel's LDS limit, set once
  static std::mutex mu;
  static std::map<int, int> cus_of;
  int dev = 0, cus = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  {
    std::lock_guard<std::mutex> g(mu);
    auto it = cus_of.find(dev);
    if (it == cus_of.end()) {
      e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
      if (e != hipSuccess) return e;
      if (cus <= 0) return hipErrorInvalidDevice;
      e = cfft_launch::ensure_lds(reinterpret_cast<const void *>(k_r16_polymul_persistent),
                                  kPersistLds);
      if (e != hipSuccess) return e;
      cus_of[dev] = cus;
    } else {
      cus = it->second;
    }
  }
  const uint64_t want = (batch + kPW - 1) / kPW;
  const uint32_t grid = (uint32_t)(want < (uint64_t)cus ? want : (uint64_t)cus);
  hipLaunchKernelGGL(k_r16_polymul_persistent, dim3(grid), dim3(64 * kPW), kPersistLds, s, a, b,
                     c, batch, reinterpret_cast<const c64 *>(tw), reinterpret_cast<const c64 *>(twist));
  return hipGetLastError();
}

}  // namespace

namespace cfft_launch {

hipError_t r16_polymul(uint32_t log2n, const uint64_t *a, const int64_t *b, uint64_t *c,
                       uint64_t batch, const dc64 *tw, const dc64 *twist, hipStream_t s) {
  if (batch == 0) return hipSuccess;
  // n = 1024: one product per wave, 8 waves per workgroup sharing the
  // twiddles staged in LDS (n = 512 runs k_wsub_polymul)
  if (log2n != 10) return hipErrorInvalidValue;
  return launch_polymul_persistent(a, b, c, batch, tw, twist, s);
}

}  // namespace cfft_launch
