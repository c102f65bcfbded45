// fft_kernels.hip -- gfx950 kernels for the f64 transforms of the ordered
// and unordered plans, the permuted-order helpers and the generic fused
// negacyclic product.
#include <hip/hip_runtime.h>

#include "lds_attr.hpp"


#include "fixed_exec.hpp"
#include "launch.hpp"
#include "stockham_exec.hpp"

using namespace cfft_dev;
using cfft_launch::dc64;
using cfft_prog::Pass;
using cfft_prog::Program;
using cfft_fixed::Geom;
using cfft_fixed::geom_for;
using cfft_fixed::max_radix;

namespace {

// One workgroup runs G = blockDim.x / T chunks of n = 2^log2n points with
// T = n / E threads each.  Inactive tail chunks still take part in the
// barriers but neither read nor write global memory.
template <int E>
__global__ void __launch_bounds__(512) k_fft_lds(c64 *__restrict__ data, uint64_t ntrans,
                                                 uint32_t chunks_log2, uint64_t stride,
                                                 const c64 *__restrict__ tw, Program prog) {
  extern __shared__ c64 smem[];
  const uint32_t log2n = prog.log2n;
  const uint32_t n = 1u << log2n;
  const uint32_t T = n / E;
  const uint32_t G = blockDim.x / T;
  const uint32_t g = threadIdx.x / T, lt = threadIdx.x % T;
  const uint64_t tr = (uint64_t)blockIdx.x * G + g;
  const bool active = tr < ntrans;
  c64 *lds = smem + g * lds_words(n);
  const uint64_t tb = tr >> chunks_log2, tc = tr & ((1ull << chunks_log2) - 1);
  c64 *gp = data + tb * stride + (tc << log2n);
  auto load = [&](uint32_t pos) -> c64 { return active ? ld_stream(&gp[pos]) : mk(0.0, 0.0); };
  auto store = [&](uint32_t pos, c64 v) {
    if (active) st_stream(v, &gp[pos]);
  };
  run_program<E>(prog, lds, lt, T, tw, load, store);
}

template <int R>
__global__ void __launch_bounds__(256) k_depth_global(c64 *__restrict__ data, uint64_t total,
                                                      uint32_t log2n, uint64_t stride,
                                                      const c64 *__restrict__ tw, Pass P) {
  const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= total) return;
  constexpr uint32_t lr = RadixLog<R>::value;
  const uint64_t tr = gid >> (log2n - lr);
  const uint32_t b = (uint32_t)(gid & ((1ull << (log2n - lr)) - 1));
  c64 *gp = data + tr * stride;
  uint32_t in[R], out[R], twi[R];
  pass_positions<R>(P, b, in, out, twi);
  c64 v[R];
#pragma unroll
  for (int k = 0; k < R; ++k) v[k] = gp[in[k]];
  apply_butterfly<R>(P, v, twi, tw);
#pragma unroll
  for (int k = 0; k < R; ++k) gp[out[k]] = v[k];
}

// Any pass (Stockham core, end butterfly or depth pass) of a transform too
// large for one workgroup's LDS, straight over global memory: one thread per
// butterfly, reading src and writing dst (which may be src for the passes
// whose outputs land where their inputs were: END and DEPTH_*).
template <int R>
__global__ void __launch_bounds__(256) k_pass_global(const c64 *src, uint64_t src_stride,
                                                     c64 *dst, uint64_t dst_stride,
                                                     uint64_t total, uint32_t log2n,
                                                     const c64 *__restrict__ tw, Pass P) {
  const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= total) return;
  constexpr uint32_t lr = RadixLog<R>::value;
  const uint64_t tr = gid >> (log2n - lr);
  const uint32_t b = (uint32_t)(gid & ((1ull << (log2n - lr)) - 1));
  const c64 *sp = src + tr * src_stride;
  c64 *dp = dst + tr * dst_stride;
  uint32_t in[R], out[R], twi[R];
  pass_positions<R>(P, b, in, out, twi);
  c64 v[R];
#pragma unroll
  for (int k = 0; k < R; ++k) v[k] = sp[in[k]];
  apply_butterfly<R>(P, v, twi, tw);
#pragma unroll
  for (int k = 0; k < R; ++k) dp[out[k]] = v[k];
}

__global__ void k_monomial(c64 *__restrict__ buf, uint32_t log2n, uint64_t degree,
                           const uint32_t *__restrict__ freq, const c64 *__restrict__ tw) {
  const uint32_t n = 1u << log2n;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t mask = n - 1;
  buf[i] = tw[((uint64_t)freq[i] * (degree & mask)) & mask];
}

__global__ void k_gather(const c64 *__restrict__ src, c64 *__restrict__ dst,
                         const uint32_t *__restrict__ idx, uint32_t n, uint64_t total) {
  const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= total) return;
  const uint64_t b = gid / n;
  const uint32_t i = (uint32_t)(gid - b * n);
  dst[gid] = src[b * n + idx[i]];
}

__global__ void k_scatter(const c64 *__restrict__ src, c64 *__restrict__ dst,
                          const uint32_t *__restrict__ idx, uint32_t n, uint64_t total) {
  const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= total) return;
  const uint64_t b = gid / n;
  const uint32_t i = (uint32_t)(gid - b * n);
  dst[b * n + idx[i]] = src[gid];
}

// Generic fused negacyclic product, N = 2n: one chunk of LDS per
// polynomial.  A's transform is parked in registers while B's runs.
template <int E>
__global__ void __launch_bounds__(512) k_polymul_generic(
    const uint64_t *__restrict__ A, const int64_t *__restrict__ B, uint64_t *__restrict__ C,
    uint64_t batch, const c64 *__restrict__ tw, const c64 *__restrict__ twist,
    const c64 *__restrict__ untwist, Program fwd, Program inv) {
  extern __shared__ c64 smem[];
  const uint32_t log2n = fwd.log2n;
  const uint32_t n = 1u << log2n;
  const uint32_t T = n / E;
  const uint32_t G = blockDim.x / T;
  const uint32_t g = threadIdx.x / T, lt = threadIdx.x % T;
  const uint64_t poly = (uint64_t)blockIdx.x * G + g;
  const bool active = poly < batch;
  c64 *lds = smem + g * lds_words(n);
  const uint64_t N = 2ull * n;
  const uint64_t *a = A + poly * N;
  const int64_t *bb = B + poly * N;
  uint64_t *c = C + poly * N;

  auto load_a = [&](uint32_t j) -> c64 {
    if (!active) return mk(0.0, 0.0);
    return cmul(twist[j], mk(torus_to_f64(a[j]), torus_to_f64(a[j + n])));
  };
  auto load_b = [&](uint32_t j) -> c64 {
    if (!active) return mk(0.0, 0.0);
    return cmul(twist[j], mk(i64_to_f64(bb[j]), i64_to_f64(bb[j + n])));
  };
  auto load_lds = [&](uint32_t j) -> c64 { return lds[lds_pad(j)]; };
  auto store_lds = [&](uint32_t j, c64 v) { lds[lds_pad(j)] = v; };
  auto store_c = [&](uint32_t j, c64 v) {
    if (!active) return;
    c64 r = cmul(untwist[j], v);
    c[j] = from_torus(r.re);
    c[j + n] = from_torus(r.im);
  };

  run_program<E>(fwd, lds, lt, T, tw, load_a, store_lds);
  c64 fa[E];
#pragma unroll
  for (int e = 0; e < E; ++e) fa[e] = lds[lds_pad(lt + T * (uint32_t)e)];
  __syncthreads();
  run_program<E>(fwd, lds, lt, T, tw, load_b, store_lds);
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const uint32_t j = lds_pad(lt + T * (uint32_t)e);
    lds[j] = cmul(fa[e], lds[j]);
  }
  __syncthreads();
  run_program<E>(inv, lds, lt, T, tw, load_lds, store_c);
}

// Dynamic LDS above 64 KiB must be opted into per kernel.
hipError_t set_lds(const void *kernel, size_t bytes) {
  return cfft_launch::ensure_lds(kernel, bytes);
}

}  // namespace

namespace cfft_launch {

namespace {

bool fixed_fft(const Program &prog, dc64 *data, uint64_t ntrans, uint64_t stride,
               const dc64 *tw, hipStream_t s, hipError_t *e) {
  switch (prog.log2n) {
    case 8: return fixed_fft_n256(prog, data, ntrans, stride, tw, s, e);
    case 9: return fixed_fft_n512(prog, data, ntrans, stride, tw, s, e);
    case 10: return fixed_fft_n1024(prog, data, ntrans, stride, tw, s, e);
    case 11: return fixed_fft_n2048(prog, data, ntrans, stride, tw, s, e);
    case 12: return fixed_fft_n4096(prog, data, ntrans, stride, tw, s, e);
    case 13: return fixed_fft_n8192(prog, data, ntrans, stride, tw, s, e);
    default: return false;
  }
}

bool fixed_polymul(const Program &fwd, const Program &inv, const uint64_t *a, const int64_t *b,
                   uint64_t *c, uint64_t batch, const dc64 *tw, const dc64 *twist,
                   const dc64 *untwist, hipStream_t s, hipError_t *e) {
  switch (fwd.log2n) {
    case 8: return fixed_polymul_n256(fwd, inv, a, b, c, batch, tw, twist, untwist, s, e);
    case 9: return fixed_polymul_n512(fwd, inv, a, b, c, batch, tw, twist, untwist, s, e);
    case 10: return fixed_polymul_n1024(fwd, inv, a, b, c, batch, tw, twist, untwist, s, e);
    case 11: return fixed_polymul_n2048(fwd, inv, a, b, c, batch, tw, twist, untwist, s, e);
    case 12: return fixed_polymul_n4096(fwd, inv, a, b, c, batch, tw, twist, untwist, s, e);
    case 13: return fixed_polymul_n8192(fwd, inv, a, b, c, batch, tw, twist, untwist, s, e);
    default: return false;
  }
}

}  // namespace

hipError_t fft_lds(const Program &prog, dc64 *data, uint64_t ntrans, uint32_t chunks_log2,
                   uint64_t stride, const dc64 *tw, hipStream_t s) {
  if (ntrans == 0 || prog.npass == 0) return hipSuccess;
  if (chunks_log2 == 0) {
    hipError_t e = hipSuccess;
    if (fixed_fft(prog, data, ntrans, stride, tw, s, &e)) return e;
  }
  Geom g = geom_for(prog.log2n, 2048, max_radix(prog));
  const uint64_t blocks = (ntrans + g.G - 1) / g.G;
  c64 *d = reinterpret_cast<c64 *>(data);
  const c64 *t = reinterpret_cast<const c64 *>(tw);
  hipError_t e = hipSuccess;
  switch (g.E) {
    case 2:
      e = set_lds(reinterpret_cast<const void *>(k_fft_lds<2>), g.lds);
      if (e == hipSuccess)
        hipLaunchKernelGGL(k_fft_lds<2>, dim3((uint32_t)blocks), dim3(g.threads), g.lds, s, d,
                           ntrans, chunks_log2, stride, t, prog);
      break;
    case 4:
      e = set_lds(reinterpret_cast<const void *>(k_fft_lds<4>), g.lds);
      if (e == hipSuccess)
        hipLaunchKernelGGL(k_fft_lds<4>, dim3((uint32_t)blocks), dim3(g.threads), g.lds, s, d,
                           ntrans, chunks_log2, stride, t, prog);
      break;
    case 8:
      e = set_lds(reinterpret_cast<const void *>(k_fft_lds<8>), g.lds);
      if (e == hipSuccess)
        hipLaunchKernelGGL(k_fft_lds<8>, dim3((uint32_t)blocks), dim3(g.threads), g.lds, s, d,
                           ntrans, chunks_log2, stride, t, prog);
      break;
    default:
      e = set_lds(reinterpret_cast<const void *>(k_fft_lds<16>), g.lds);
      if (e == hipSuccess)
        hipLaunchKernelGGL(k_fft_lds<16>, dim3((uint32_t)blocks), dim3(g.threads), g.lds, s, d,
                           ntrans, chunks_log2, stride, t, prog);
      break;
  }
  if (e != hipSuccess) return e;
  return hipGetLastError();
}

hipError_t depth_global(const Pass &P, dc64 *data, uint64_t batch, uint32_t log2n,
                        uint64_t stride, const dc64 *tw, hipStream_t s) {
  const uint32_t lr = P.radix == 2 ? 1 : P.radix == 4 ? 2 : 3;
  const uint64_t total = batch << (log2n - lr);
  if (total == 0) return hipSuccess;
  const uint32_t threads = 256;
  const uint64_t blocks = (total + threads - 1) / threads;
  c64 *d = reinterpret_cast<c64 *>(data);
  const c64 *t = reinterpret_cast<const c64 *>(tw);
  switch (P.radix) {
    case 2:
      hipLaunchKernelGGL(k_depth_global<2>, dim3((uint32_t)blocks), dim3(threads), 0, s, d, total,
                         log2n, stride, t, P);
      break;
    case 4:
      hipLaunchKernelGGL(k_depth_global<4>, dim3((uint32_t)blocks), dim3(threads), 0, s, d, total,
                         log2n, stride, t, P);
      break;
    default:
      hipLaunchKernelGGL(k_depth_global<8>, dim3((uint32_t)blocks), dim3(threads), 0, s, d, total,
                         log2n, stride, t, P);
      break;
  }
  return hipGetLastError();
}

hipError_t pass_global(const Pass &P, const dc64 *src, uint64_t src_stride, dc64 *dst,
                       uint64_t dst_stride, uint64_t batch, uint32_t log2n, const dc64 *tw,
                       hipStream_t s) {
  const uint32_t lr = P.radix == 2 ? 1 : P.radix == 4 ? 2 : P.radix == 8 ? 3 : 4;
  const uint64_t total = batch << (log2n - lr);
  if (total == 0) return hipSuccess;
  const uint32_t threads = 256;
  const dim3 grid((uint32_t)((total + threads - 1) / threads));
  const c64 *a = reinterpret_cast<const c64 *>(src);
  c64 *d = reinterpret_cast<c64 *>(dst);
  const c64 *t = reinterpret_cast<const c64 *>(tw);
  switch (P.radix) {
    case 2:
      hipLaunchKernelGGL(k_pass_global<2>, grid, dim3(threads), 0, s, a, src_stride, d, dst_stride,
                         total, log2n, t, P);
      break;
    case 4:
      hipLaunchKernelGGL(k_pass_global<4>, grid, dim3(threads), 0, s, a, src_stride, d, dst_stride,
                         total, log2n, t, P);
      break;
    case 8:
      hipLaunchKernelGGL(k_pass_global<8>, grid, dim3(threads), 0, s, a, src_stride, d, dst_stride,
                         total, log2n, t, P);
      break;
    default:
      hipLaunchKernelGGL(k_pass_global<16>, grid, dim3(threads), 0, s, a, src_stride, d,
                         dst_stride, total, log2n, t, P);
      break;
  }
  return hipGetLastError();
}

hipError_t monomial(dc64 *buf, uint32_t log2n, uint64_t degree, const uint32_t *freq,
                    const dc64 *tw, hipStream_t s) {
  const uint32_t n = 1u << log2n;
  hipLaunchKernelGGL(k_monomial, dim3((n + 255) / 256), dim3(256), 0, s,
                     reinterpret_cast<c64 *>(buf), log2n, degree, freq,
                     reinterpret_cast<const c64 *>(tw));
  return hipGetLastError();
}

hipError_t gather(const dc64 *src, dc64 *dst, const uint32_t *idx, uint32_t n, uint64_t batch,
                  hipStream_t s) {
  const uint64_t total = (uint64_t)n * batch;
  if (total == 0) return hipSuccess;
  hipLaunchKernelGGL(k_gather, dim3((uint32_t)((total + 255) / 256)), dim3(256), 0, s,
                     reinterpret_cast<const c64 *>(src), reinterpret_cast<c64 *>(dst), idx, n,
                     total);
  return hipGetLastError();
}

hipError_t scatter(const dc64 *src, dc64 *dst, const uint32_t *idx, uint32_t n, uint64_t batch,
                   hipStream_t s) {
  const uint64_t total = (uint64_t)n * batch;
  if (total == 0) return hipSuccess;
  hipLaunchKernelGGL(k_scatter, dim3((uint32_t)((total + 255) / 256)), dim3(256), 0, s,
                     reinterpret_cast<const c64 *>(src), reinterpret_cast<c64 *>(dst), idx, n,
                     total);
  return hipGetLastError();
}

hipError_t polymul_generic(const Program &fwd, const Program &inv, const uint64_t *a,
                           const int64_t *b, uint64_t *c, uint64_t batch, const dc64 *tw,
                           const dc64 *twist, const dc64 *untwist, hipStream_t s) {
  if (batch == 0) return hipSuccess;
  {
    hipError_t e = hipSuccess;
    if (fixed_polymul(fwd, inv, a, b, c, batch, tw, twist, untwist, s, &e)) return e;
  }
  const uint32_t rf = max_radix(fwd), ri = max_radix(inv);
  Geom g = geom_for(fwd.log2n, 2048, rf > ri ? rf : ri);
  const uint64_t blocks = (batch + g.G - 1) / g.G;
  const c64 *t = reinterpret_cast<const c64 *>(tw);
  const c64 *tws = reinterpret_cast<const c64 *>(twist);
  const c64 *utw = reinterpret_cast<const c64 *>(untwist);
  hipError_t e = hipSuccess;
  switch (g.E) {
    case 2:
      e = set_lds(reinterpret_cast<const void *>(k_polymul_generic<2>), g.lds);
      if (e == hipSuccess)
        hipLaunchKernelGGL(k_polymul_generic<2>, dim3((uint32_t)blocks), dim3(g.threads), g.lds,
                           s, a, b, c, batch, t, tws, utw, fwd, inv);
      break;
    case 4:
      e = set_lds(reinterpret_cast<const void *>(k_polymul_generic<4>), g.lds);
      if (e == hipSuccess)
        hipLaunchKernelGGL(k_polymul_generic<4>, dim3((uint32_t)blocks), dim3(g.threads), g.lds,
                           s, a, b, c, batch, t, tws, utw, fwd, inv);
      break;
    case 8:
      e = set_lds(reinterpret_cast<const void *>(k_polymul_generic<8>), g.lds);
      if (e == hipSuccess)
        hipLaunchKernelGGL(k_polymul_generic<8>, dim3((uint32_t)blocks), dim3(g.threads), g.lds,
                           s, a, b, c, batch, t, tws, utw, fwd, inv);
      break;
    default:
      e = set_lds(reinterpret_cast<const void *>(k_polymul_generic<16>), g.lds);
      if (e == hipSuccess)
        hipLaunchKernelGGL(k_polymul_generic<16>, dim3((uint32_t)blocks), dim3(g.threads), g.lds,
                           s, a, b, c, batch, t, tws, utw, fwd, inv);
      break;
  }
  if (e != hipSuccess) return e;
  return hipGetLastError();
}

}  // namespace cfft_launch
