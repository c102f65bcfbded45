// dit16_kernels.hip -- register-resident kernels for Dit16 plans with
// n = base_n = 1024 (r16dit_exec.hpp): batched forward / inverse transforms
// and the fused negacyclic product at N = 2048.
#include <hip/hip_runtime.h>

#include "launch.hpp"
#include "lds_attr.hpp"
#include "r16dit_exec.hpp"

using namespace cfft_dev;
using cfft_launch::dc64;

namespace {

constexpr uint32_t kWaves = 8;
constexpr uint32_t kTw1 = 15 * 64, kTw2 = 15 * 4;  // twiddle rows k = 1..15
constexpr uint32_t kX = 1024;                       // one wave's exchange buffer
constexpr size_t kTransformLds = (size_t)(kTw1 + kTw2 + kWaves * kX) * sizeof(c64);

__device__ __forceinline__ void stage_rows(const c64 *__restrict__ tw_r16, c64 *ltw1, c64 *ltw2) {
  for (uint32_t i = threadIdx.x; i < kTw1; i += blockDim.x) ltw1[i] = tw_r16[64 + i];
  for (uint32_t i = threadIdx.x; i < kTw2; i += blockDim.x) ltw2[i] = tw_r16[1024 + 4 + i];
}

// one transform per wave, 8 waves per workgroup; in place (each wave reads
// its transform before writing it)
template <bool FWD>
__global__ void __launch_bounds__(64 * kWaves) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_dit16_transform(c64 *__restrict__ data, uint64_t batch, uint64_t stride,
                  const c64 *__restrict__ tw_r16) {
  extern __shared__ c64 smem[];
  c64 *ltw1 = smem, *ltw2 = smem + kTw1;
  stage_rows(tw_r16, ltw1, ltw2);
  __syncthreads();
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint64_t t = (uint64_t)blockIdx.x * kWaves + wave;
  if (t >= batch) return;
  c64 *x = smem + kTw1 + kTw2 + wave * kX;
  const uint32_t l = threadIdx.x & 63u;
  c64 *d = data + t * stride;
  c64 v[16];
#pragma unroll
  for (uint32_t m = 0; m < 16; ++m) v[m] = ld_stream(&d[l + 64 * m]);
  r16dit_transform<FWD, 1>(v, v, x, l, ltw1, ltw2);
#pragma unroll
  for (uint32_t m = 0; m < 16; ++m) st_stream(v[m], &d[l + 64 * m]);
}

// (double)(int64)a * 2^-64 / 1024 with one rounding (2^-74 and 2^-42 scale
// exactly)
__device__ __forceinline__ double torus_over_n1024(uint64_t a) {
  const double hi = (double)(int32_t)(a >> 32);
  const double lo = (double)(uint32_t)a * 0x1p-74;
  return __fma_rn(hi, 0x1p-42, lo);
}

// Fused negacyclic product, N = 2048: persistent, one product per wave at a
// time, 8 waves per workgroup; LDS holds the twiddle rows, the twist table
// and each wave's exchange buffer.  Fold + twist, both forward transforms
// together, pointwise product, inverse, untwist (the 1/n folded into A's
// conversion, an exact power of two) and torus rounding -- the operations of
// the generic fused product for this plan.  B of the wave's next product is
// loaded during the inverse (past its last product: the L2-resident twist
// table, never used).
constexpr uint32_t kTwist = 1024;
constexpr size_t kProductLds = (size_t)(kTw1 + kTw2 + kTwist + kWaves * kX) * sizeof(c64);

__global__ void __launch_bounds__(64 * kWaves) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_dit16_polymul(const uint64_t *__restrict__ A, const int64_t *__restrict__ B,
                uint64_t *__restrict__ C, uint64_t batch, const c64 *__restrict__ tw_r16,
                const c64 *__restrict__ twist) {
  constexpr uint32_t n = 1024;
  extern __shared__ c64 smem[];
  c64 *ltw1 = smem, *ltw2 = smem + kTw1, *ltz = smem + kTw1 + kTw2;
  stage_rows(tw_r16, ltw1, ltw2);
  for (uint32_t i = threadIdx.x; i < kTwist; i += blockDim.x) ltz[i] = twist[i];
  __syncthreads();
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  c64 *x = smem + kTw1 + kTw2 + kTwist + wave * kX;
  const uint64_t step = (uint64_t)gridDim.x * kWaves;
  uint64_t poly = (uint64_t)blockIdx.x * kWaves + wave;
  uint64_t pb[32];  // raw words of B for the wave's current product
  if (poly < batch) {
    const uint32_t l = threadIdx.x & 63u;
    const int64_t *b = B + poly * (2 * n);
#pragma unroll
    for (uint32_t m = 0; m < 16; ++m) {
      pb[2 * m] = (uint64_t)ld_stream(&b[l + 64 * m]);
      pb[2 * m + 1] = (uint64_t)ld_stream(&b[l + 64 * m + n]);
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
        const uint32_t j = l + 64 * m;
        ra[2 * m] = ld_stream(&a[j]);
        ra[2 * m + 1] = ld_stream(&a[j + n]);
      }
      sched_fence();
#pragma unroll
      for (uint32_t m = 0; m < 16; ++m) {
        const c64 z = ltz[l + 64 * m];
        fa[m] = cmul(z, mk(torus_over_n1024(ra[2 * m]), torus_over_n1024(ra[2 * m + 1])));
        v[m] = cmul(z, mk(i64_to_f64((int64_t)pb[2 * m]), i64_to_f64((int64_t)pb[2 * m + 1])));
      }
    }
    r16dit_transform<true, 2>(fa, v, x, l, ltw1, ltw2);
#pragma unroll
    for (uint32_t s = 0; s < 16; ++s) fa[s] = cmul(fa[s], v[s]);
    sched_fence();
    {
      const uint64_t nx = poly + step;
      const int64_t *bn = nx < batch ? B + nx * (2 * n) : reinterpret_cast<const int64_t *>(twist);
#pragma unroll
      for (uint32_t m = 0; m < 16; ++m) {
        pb[2 * m] = (uint64_t)ld_stream(&bn[l + 64 * m]);
        pb[2 * m + 1] = (uint64_t)ld_stream(&bn[l + 64 * m + n]);
      }
    }
    sched_fence();
    r16dit_transform<false, 1>(fa, fa, x, l, ltw1, ltw2);
#pragma unroll
    for (uint32_t s = 0; s < 16; ++s) {
      const uint32_t j = l + 64 * s;
      const c64 r = cmul(conj(ltz[j]), fa[s]);
      st_stream(from_torus(r.re), &c[j]);
      st_stream(from_torus(r.im), &c[j + n]);
      if ((s & 3) == 3) sched_fence();
    }
  }
}

}  // namespace

namespace cfft_launch {

hipError_t dit16_transform(bool fwd, dc64 *data, uint64_t batch, uint64_t stride,
                           const dc64 *tw_r16, hipStream_t s) {
  if (batch == 0) return hipSuccess;
  if (stride < 1024) return hipErrorInvalidValue;
  const void *k = fwd ? reinterpret_cast<const void *>(k_dit16_transform<true>)
                      : reinterpret_cast<const void *>(k_dit16_transform<false>);
  hipError_t e = ensure_lds(k, kTransformLds);
  if (e != hipSuccess) return e;
  const uint64_t per_wg = kWaves;
  const dim3 grid((uint32_t)((batch + per_wg - 1) / per_wg)), block(64 * kWaves);
  c64 *d = reinterpret_cast<c64 *>(data);
  const c64 *t = reinterpret_cast<const c64 *>(tw_r16);
  if (fwd)
    hipLaunchKernelGGL(k_dit16_transform<true>, grid, block, kTransformLds, s, d, batch, stride, t);
  else
    hipLaunchKernelGGL(k_dit16_transform<false>, grid, block, kTransformLds, s, d, batch, stride, t);
  return hipGetLastError();
}

hipError_t dit16_polymul(const uint64_t *a, const int64_t *b, uint64_t *c, uint64_t batch,
                         const dc64 *tw_r16, const dc64 *twist, hipStream_t s) {
  if (batch == 0) return hipSuccess;
  hipError_t e = ensure_lds(reinterpret_cast<const void *>(k_dit16_polymul), kProductLds);
  if (e != hipSuccess) return e;
  int cus = 0;
  e = cu_count(&cus);
  if (e != hipSuccess) return e;
  const uint64_t want = (batch + kWaves - 1) / kWaves;
  const uint32_t grid = (uint32_t)(want < (uint64_t)cus ? want : (uint64_t)cus);
  hipLaunchKernelGGL(k_dit16_polymul, dim3(grid), dim3(64 * kWaves), kProductLds, s, a, b, c,
                     batch, reinterpret_cast<const c64 *>(tw_r16),
                     reinterpret_cast<const c64 *>(twist));
  return hipGetLastError();
}

}  // namespace cfft_launch
