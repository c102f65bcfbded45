// large_transform_kernels.hip -- Dif16 plans with base_n = 1024 at n = 8192
// and 16384 in one pass over HBM, 16 base transforms per 512-thread
// workgroup run in two halves of 8 (one 16384-point transform, or two of
// 8192 points).
//
// n = 16384:
// the plan's recursion (src/unordered.rs:391-489: radix 8 above 2 base_n, 2
// at 2 base_n) is a radix-8 depth pass over the 16384 points (stride 2048),
// a radix-2 depth pass inside each 2048-point sub-array (stride 1024), then
// 16 Dif16 base transforms of 1024 points.  Thread t holds the points
// j + 1024 m for j in {t, t + 512} and all m < 16: every butterfly of both
// depth passes reads and writes points of one thread, so both passes run in
// registers after the loads (forward) or before the stores (inverse).
// Between them and the base transforms the 16 chunks go through LDS, eight
// at a time (8 x 1088 padded points): wave w runs chunk w, then chunk 8 + w,
// each with its chunk's LDS region as the exchange buffer -- the layout and
// operations of the n = 4096 depth kernel and the uneven batch's depth
// tiles, so the plan's bits.
//
// Depth pass semantics (the global form's DEPTH_FWD / DEPTH_INV): forward
// butterfly j reads points k stride + j, its output k times T[k j] goes to
// sub-array bitrev(k) at j; the inverse reads sub-array bitrev(k) at j,
// multiplies by conj(T[k j]) before the butterfly and writes k stride + j.
// T is the 16384-point table for the radix-8 level and the 2048-point table
// (T_2048[j] = T_16384[8 j], the same values) for the radix-2 level.
#include <hip/hip_runtime.h>

#include "launch.hpp"
#include "lds_attr.hpp"
#include "r16_exec.hpp"

using namespace cfft_dev;
using cfft_launch::dc64;

namespace {

constexpr uint32_t kWaves = 8;
constexpr uint32_t kBlkWords = 16 * 64 + 16 * 4;  // the 1024-point r16 twiddle block
constexpr uint32_t kCh = R16<10>::lds_words;      // one padded chunk
constexpr size_t kLds = (size_t)(kBlkWords + 8 * kCh) * sizeof(c64);

// forward radix-8 level for the 8 points j2 + 2048 r of one thread, then
// sub-array index s = bitrev(k) of output k
__device__ __forceinline__ void depth8_fwd(c64 *v, const c64 *__restrict__ tw, uint32_t j2) {
  butterfly<8, true>(v);
#pragma unroll
  for (int k = 1; k < 8; ++k) v[k] = cmul(tw[(uint32_t)k * j2], v[k]);
}

template <bool FWD>
__global__ void __launch_bounds__(64 * kWaves) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_depth16k_transform(c64 *__restrict__ data, uint64_t stride, const c64 *__restrict__ tw,
                     const c64 *__restrict__ tw2, const c64 *__restrict__ tw_blk) {
  using S = R16<10>;
  extern __shared__ c64 smem[];
  c64 *ltw = smem;
  for (uint32_t i = threadIdx.x; i < kBlkWords; i += blockDim.x) ltw[i] = tw_blk[i];
  c64 *X = smem + kBlkWords;
  const uint32_t t = threadIdx.x;
  const uint32_t w = __builtin_amdgcn_readfirstlane(t >> 6), l = t & 63u;
  c64 *d = data + (uint64_t)blockIdx.x * stride;
  c64 *xw = X + w * kCh;
  c64 hold[2][8];  // the half of this thread's chunk points not in LDS
  if (FWD) {
#pragma unroll
    for (uint32_t h = 0; h < 2; ++h) {
      const uint32_t j = t + 512 * h;
      c64 lo[8], hi[8];  // radix-8 outputs at j and j + 1024
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        lo[r] = ld_stream(&d[j + 2048 * (uint32_t)r]);
        hi[r] = ld_stream(&d[j + 1024 + 2048 * (uint32_t)r]);
      }
      depth8_fwd(lo, tw, j);
      depth8_fwd(hi, tw, j + 1024);
      const c64 f = tw2[j];
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        // sub-array s = bitrev(k): radix 2 over its points j and j + 1024,
        // outputs to chunks 2 s and 2 s + 1
        const int s = bitrev_digit<8>(k);
        c64 p[2] = {lo[k], hi[k]};
        butterfly<2, true>(p);
        p[1] = cmul(f, p[1]);
#pragma unroll
        for (int e = 0; e < 2; ++e) {
          const int c = 2 * s + e;
          if (c < 8) X[(uint32_t)c * kCh + r16_pad(j)] = p[e];
          else hold[h][c - 8] = p[e];
        }
      }
    }
#pragma unroll
    for (uint32_t half = 0; half < 2; ++half) {
      __syncthreads();  // chunks of this half written (and, first, the twiddle block)
      c64 v[16];
#pragma unroll
      for (uint32_t m = 0; m < 16; ++m) v[m] = xw[r16_pad(l + 64 * m)];
      r16l_transform<10, true>(v, v, xw, l, ltw);
      const uint32_t c = 8 * half + w;
#pragma unroll
      for (uint32_t s = 0; s < 16; ++s) st_stream(v[s], &d[c * 1024 + l + 64 * S::end_m(s)]);
      if (half == 0) {
        __syncthreads();  // every chunk region read and done with
#pragma unroll
        for (uint32_t h = 0; h < 2; ++h)
#pragma unroll
          for (uint32_t c2 = 0; c2 < 8; ++c2) X[c2 * kCh + r16_pad(t + 512 * h)] = hold[h][c2];
      }
    }
  } else {
#pragma unroll
    for (uint32_t half = 0; half < 2; ++half) {
      {
        const uint32_t c = 8 * half + w;
        c64 v[16], u[16];
#pragma unroll
        for (uint32_t s = 0; s < 16; ++s) v[s] = ld_stream(&d[c * 1024 + l + 64 * S::end_m(s)]);
        // first half: twiddle block staged; second: every chunk region read
        __syncthreads();
        r16_end_to_input<10>(v, u);
        r16l_transform<10, false>(u, u, xw, l, ltw);
        wave_lds_fence();
#pragma unroll
        for (uint32_t s = 0; s < 16; ++s) xw[r16_pad(l + 64 * S::end_m(s))] = u[s];
      }
      if (half == 0) {
        __syncthreads();
#pragma unroll
        for (uint32_t h = 0; h < 2; ++h)
#pragma unroll
          for (uint32_t c2 = 0; c2 < 8; ++c2) hold[h][c2] = X[c2 * kCh + r16_pad(t + 512 * h)];
      }
    }
    __syncthreads();
#pragma unroll
    for (uint32_t h = 0; h < 2; ++h) {
      const uint32_t j = t + 512 * h;
      c64 lo[8], hi[8];  // sub-array points at j and j + 1024, by sub-array
      c64 f = tw2[j];
      f.im = -f.im;
#pragma unroll
      for (int s = 0; s < 8; ++s) {
        c64 p[2];
#pragma unroll
        for (int e = 0; e < 2; ++e) {
          const int c = 2 * s + e;
          p[e] = c < 8 ? hold[h][c] : X[(uint32_t)(c - 8) * kCh + r16_pad(j)];
        }
        p[1] = cmul(f, p[1]);
        butterfly<2, false>(p);
        lo[s] = p[0];
        hi[s] = p[1];
      }
#pragma unroll
      for (uint32_t m0 = 0; m0 < 2; ++m0) {
        const uint32_t j2 = j + 1024 * m0;
        c64 v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          const int s = bitrev_digit<8>(k);
          v[k] = m0 ? hi[s] : lo[s];
        }
#pragma unroll
        for (int k = 1; k < 8; ++k) {
          c64 g = tw[(uint32_t)k * j2];
          g.im = -g.im;
          v[k] = cmul(g, v[k]);
        }
        butterfly<8, false>(v);
#pragma unroll
        for (int k = 0; k < 8; ++k) st_stream(v[k], &d[j2 + 2048 * (uint32_t)k]);
      }
    }
  }
}

// n = 8192 (one radix-8 depth level, stride 1024, above 8 chunks): two
// transforms per workgroup in the same two halves.  Thread t runs the depth
// butterflies j in {t, t + 512} of both; the first transform's chunks go
// through LDS while the second's wait in registers (forward), and the other
// way round for the inverse.  Whether a workgroup has a second transform is
// uniform across it, so a lone last transform skips the second half and its
// barriers altogether.

// forward depth butterflies of one transform: chunk bitrev(k) at j of output k
__device__ __forceinline__ void depth8k_in(const c64 *dd, uint32_t j, const c64 *__restrict__ tw,
                                           c64 *out) {
  c64 v[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) v[k] = ld_stream(&dd[j + 1024 * (uint32_t)k]);
  depth8_fwd(v, tw, j);
#pragma unroll
  for (int k = 0; k < 8; ++k) out[bitrev_digit<8>(k)] = v[k];
}

// inverse depth butterflies of one transform from its chunk points at j
__device__ __forceinline__ void depth8k_out(c64 *dd, uint32_t j, const c64 *__restrict__ tw,
                                            const c64 *chunk) {
  c64 v[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) v[k] = chunk[bitrev_digit<8>(k)];
#pragma unroll
  for (int k = 1; k < 8; ++k) {
    c64 g = tw[(uint32_t)k * j];
    g.im = -g.im;
    v[k] = cmul(g, v[k]);
  }
  butterfly<8, false>(v);
#pragma unroll
  for (int k = 0; k < 8; ++k) st_stream(v[k], &dd[j + 1024 * (uint32_t)k]);
}

// wave w's forward base transform of chunk w (in LDS at xw), stored to dd
__device__ __forceinline__ void chunk8k_fwd(c64 *dd, c64 *xw, uint32_t w, uint32_t l,
                                            const c64 *ltw) {
  using S = R16<10>;
  c64 v[16];
#pragma unroll
  for (uint32_t m = 0; m < 16; ++m) v[m] = xw[r16_pad(l + 64 * m)];
  r16l_transform<10, true>(v, v, xw, l, ltw);
#pragma unroll
  for (uint32_t s = 0; s < 16; ++s) st_stream(v[s], &dd[w * 1024 + l + 64 * S::end_m(s)]);
}

// wave w's inverse base transform of chunk w of dd into LDS at xw; the
// barrier between the loads and the first LDS access is part of it
__device__ __forceinline__ void chunk8k_inv(const c64 *dd, c64 *xw, uint32_t w, uint32_t l,
                                            const c64 *ltw) {
  using S = R16<10>;
  c64 v[16], u[16];
#pragma unroll
  for (uint32_t s = 0; s < 16; ++s) v[s] = ld_stream(&dd[w * 1024 + l + 64 * S::end_m(s)]);
  __syncthreads();
  r16_end_to_input<10>(v, u);
  r16l_transform<10, false>(u, u, xw, l, ltw);
  wave_lds_fence();
#pragma unroll
  for (uint32_t s = 0; s < 16; ++s) xw[r16_pad(l + 64 * S::end_m(s))] = u[s];
}

template <bool FWD>
__global__ void __launch_bounds__(64 * kWaves) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_depth8k_pair(c64 *__restrict__ data, uint64_t batch, uint64_t stride,
               const c64 *__restrict__ tw, const c64 *__restrict__ tw_blk) {
  extern __shared__ c64 smem[];
  c64 *ltw = smem;
  for (uint32_t i = threadIdx.x; i < kBlkWords; i += blockDim.x) ltw[i] = tw_blk[i];
  c64 *X = smem + kBlkWords;
  const uint32_t t = threadIdx.x;
  const uint32_t w = __builtin_amdgcn_readfirstlane(t >> 6), l = t & 63u;
  const uint64_t tr0 = 2ull * blockIdx.x;
  const bool second = tr0 + 1 < batch;  // uniform across the workgroup
  c64 *d0 = data + tr0 * stride;
  c64 *d1 = d0 + stride;  // used only when `second`
  c64 *xw = X + w * kCh;
  c64 hold[2][8];  // the transform not in LDS: points j of its 8 chunks
  if (FWD) {
#pragma unroll
    for (uint32_t h = 0; h < 2; ++h) {
      const uint32_t j = t + 512 * h;
      c64 o[8];
      depth8k_in(d0, j, tw, o);
#pragma unroll
      for (uint32_t c = 0; c < 8; ++c) X[c * kCh + r16_pad(j)] = o[c];
    }
    if (second) {
#pragma unroll
      for (uint32_t h = 0; h < 2; ++h) depth8k_in(d1, t + 512 * h, tw, hold[h]);
    }
    __syncthreads();  // first transform's chunks and the twiddle block in LDS
    chunk8k_fwd(d0, xw, w, l, ltw);
    if (second) {
      __syncthreads();  // every chunk region read and done with
#pragma unroll
      for (uint32_t h = 0; h < 2; ++h)
#pragma unroll
        for (uint32_t c = 0; c < 8; ++c) X[c * kCh + r16_pad(t + 512 * h)] = hold[h][c];
      __syncthreads();
      chunk8k_fwd(d1, xw, w, l, ltw);
    }
  } else {
    chunk8k_inv(d0, xw, w, l, ltw);  // (its barrier: twiddle block staged)
    if (second) {
      __syncthreads();
#pragma unroll
      for (uint32_t h = 0; h < 2; ++h)
#pragma unroll
        for (uint32_t c = 0; c < 8; ++c) hold[h][c] = X[c * kCh + r16_pad(t + 512 * h)];
      chunk8k_inv(d1, xw, w, l, ltw);  // (its barrier: every chunk region read)
    }
    __syncthreads();
#pragma unroll
    for (uint32_t h = 0; h < 2; ++h) {
      const uint32_t j = t + 512 * h;
      c64 f[8];
#pragma unroll
      for (uint32_t c = 0; c < 8; ++c) f[c] = X[c * kCh + r16_pad(j)];
      if (second) {
        depth8k_out(d0, j, tw, hold[h]);
        depth8k_out(d1, j, tw, f);
      } else {
        depth8k_out(d0, j, tw, f);
      }
    }
  }
}

}  // namespace

namespace cfft_launch {

hipError_t depth16k_transform(bool fwd, dc64 *data, uint64_t batch, uint64_t stride,
                              const dc64 *tw, const dc64 *tw2, const dc64 *tw_blk,
                              hipStream_t s) {
  if (batch == 0) return hipSuccess;
  if (batch > 0x7fffffffull || stride < 16384) return hipErrorInvalidValue;
  const void *k = fwd ? reinterpret_cast<const void *>(k_depth16k_transform<true>)
                      : reinterpret_cast<const void *>(k_depth16k_transform<false>);
  hipError_t e = ensure_lds(k, kLds);
  if (e != hipSuccess) return e;
  c64 *d = reinterpret_cast<c64 *>(data);
  const c64 *t = reinterpret_cast<const c64 *>(tw);
  const c64 *t2 = reinterpret_cast<const c64 *>(tw2);
  const c64 *tb = reinterpret_cast<const c64 *>(tw_blk);
  const dim3 grid((uint32_t)batch), block(64 * kWaves);
  if (fwd)
    hipLaunchKernelGGL(k_depth16k_transform<true>, grid, block, kLds, s, d, stride, t, t2, tb);
  else
    hipLaunchKernelGGL(k_depth16k_transform<false>, grid, block, kLds, s, d, stride, t, t2, tb);
  return hipGetLastError();
}

hipError_t depth8k_transform(bool fwd, dc64 *data, uint64_t batch, uint64_t stride,
                             const dc64 *tw, const dc64 *tw_blk, hipStream_t s) {
  if (batch == 0) return hipSuccess;
  if (batch > 0xfffffffeull || stride < 8192) return hipErrorInvalidValue;
  const void *k = fwd ? reinterpret_cast<const void *>(k_depth8k_pair<true>)
                      : reinterpret_cast<const void *>(k_depth8k_pair<false>);
  hipError_t e = ensure_lds(k, kLds);
  if (e != hipSuccess) return e;
  c64 *d = reinterpret_cast<c64 *>(data);
  const c64 *t = reinterpret_cast<const c64 *>(tw);
  const c64 *tb = reinterpret_cast<const c64 *>(tw_blk);
  const dim3 grid((uint32_t)((batch + 1) / 2)), block(64 * kWaves);
  if (fwd)
    hipLaunchKernelGGL(k_depth8k_pair<true>, grid, block, kLds, s, d, batch, stride, t, tb);
  else
    hipLaunchKernelGGL(k_depth8k_pair<false>, grid, block, kLds, s, d, batch, stride, t, tb);
  return hipGetLastError();
}

}  // namespace cfft_launch
