// fixed_exec.hpp -- kernels that run a plan's pass program fixed at compile
// time (cfft_prog::fixed_program, the builder the plans are made with; each
// launcher checks that the plan's program is the one it was instantiated
// for), and the workgroup geometry shared with the runtime executor's
// kernels.  With every pass
// field a constant the passes unroll, positions and twiddle indices fold and
// no pass descriptor is read at run time; the operations, and so the bits,
// are the runtime executor's.  fixed_n256.hip ... fixed_n8192.hip instantiate
// them for the plans listed there.
#pragma once

#include <hip/hip_runtime.h>

#include "launch.hpp"
#include "lds_attr.hpp"
#include "stockham_exec.hpp"

namespace cfft_fixed {

using cfft_dev::c64;
using cfft_prog::Program;

// workgroup geometry of the LDS kernels: E points per thread, about 34 KiB of
// LDS per workgroup (several workgroups per CU) and at most 512 threads.
// E = 16 only for programs with a radix-16 pass: the 16-point form of the
// executor needs about twice the registers of the 8-point one.
struct Geom {
  int E;
  uint32_t T, G, threads;
  size_t lds;
};
inline uint32_t max_radix(const Program &prog) {
  uint32_t r = 2;
  for (uint32_t i = 0; i < prog.npass; ++i) r = prog.p[i].radix > r ? prog.p[i].radix : r;
  return r;
}
inline Geom geom_for(uint32_t log2n, uint32_t per_wg_points, uint32_t radix) {
  Geom g;
  const uint32_t n = 1u << log2n;
  // (8 points per thread up to 4096 points: at most 512 threads per transform)
  const uint32_t e_max = radix > 8 || n > 4096 ? 16u : 8u;
  g.E = (int)(n >= e_max ? e_max : n);
  g.T = n / (uint32_t)g.E;
  g.G = per_wg_points > n ? per_wg_points / n : 1;
  if (g.T * g.G > 512) g.G = 512 / g.T > 0 ? 512 / g.T : 1;
  g.threads = g.T * g.G;
  g.lds = (size_t)g.G * cfft_dev::lds_words(n) * sizeof(c64);
  return g;
}

}  // namespace cfft_fixed

namespace {

using namespace cfft_dev;
using cfft_prog::Program;

// The generic fused product (k_polymul_generic) for one (algo, n, base_n)
// fixed at compile time: the forward and inverse programs are
// cfft_prog::fixed_program's, unrolled with constant pass fields.
template <int ALGO, int LOG2N, int LOG2BN, bool FWD>
struct FixedProg {
  static constexpr Program prog = cfft_prog::fixed_program(ALGO, LOG2N, LOG2BN, FWD);
};

template <int E, int ALGO, int LOG2N, int LOG2BN>
__global__ void __launch_bounds__(512) k_polymul_fixed(
    const uint64_t *__restrict__ A, const int64_t *__restrict__ B, uint64_t *__restrict__ C,
    uint64_t batch, const c64 *__restrict__ tw, const c64 *__restrict__ twist,
    const c64 *__restrict__ untwist) {
  extern __shared__ c64 smem[];
  constexpr uint32_t n = 1u << LOG2N;
  constexpr uint32_t T = n / E;
  const uint32_t G = blockDim.x / T;
  const uint32_t g = threadIdx.x / T, lt = threadIdx.x % T;
  const uint64_t poly = (uint64_t)blockIdx.x * G + g;
  const bool active = poly < batch;
  c64 *lds = smem + g * lds_words(n);
  constexpr uint64_t N = 2ull * n;
  const uint64_t *a = A + poly * N;
  const int64_t *bb = B + poly * N;
  uint64_t *c = C + poly * N;
  using Fwd = FixedProg<ALGO, LOG2N, LOG2BN, true>;
  using Inv = FixedProg<ALGO, LOG2N, LOG2BN, false>;

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

  run_fixed_program<Fwd, 0, E>(lds, lt, T, tw, load_a, store_lds);
  c64 fa[E];
#pragma unroll
  for (int e = 0; e < E; ++e) fa[e] = lds[lds_pad(lt + T * (uint32_t)e)];
  __syncthreads();
  run_fixed_program<Fwd, 0, E>(lds, lt, T, tw, load_b, store_lds);
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const uint32_t j = lds_pad(lt + T * (uint32_t)e);
    lds[j] = cmul(fa[e], lds[j]);
  }
  __syncthreads();
  run_fixed_program<Inv, 0, E>(lds, lt, T, tw, load_lds, store_c);
}

// k_fft_lds for one (algo, n, base_n, direction) fixed at compile time (a
// whole-transform program: no chunk split).
template <int E, int ALGO, int LOG2N, int LOG2BN, bool FWD>
__global__ void __launch_bounds__(512) k_fft_fixed(c64 *__restrict__ data, uint64_t ntrans,
                                                   uint64_t stride, const c64 *__restrict__ tw) {
  extern __shared__ c64 smem[];
  constexpr uint32_t n = 1u << LOG2N;
  constexpr uint32_t T = n / E;
  const uint32_t G = blockDim.x / T;
  const uint32_t g = threadIdx.x / T, lt = threadIdx.x % T;
  const uint64_t tr = (uint64_t)blockIdx.x * G + g;
  const bool active = tr < ntrans;
  c64 *lds = smem + g * lds_words(n);
  c64 *gp = data + tr * stride;
  auto load = [&](uint32_t pos) -> c64 { return active ? ld_stream(&gp[pos]) : mk(0.0, 0.0); };
  auto store = [&](uint32_t pos, c64 v) {
    if (active) st_stream(v, &gp[pos]);
  };
  run_fixed_program<FixedProg<ALGO, LOG2N, LOG2BN, FWD>, 0, E>(lds, lt, T, tw, load, store);
}


// Launch the product / transform kernel for (ALGO, LOG2N, LOG2BN) if the
// plan's programs are that program; false if they are not (nothing launched).
template <int ALGO, int LOG2N, int LOG2BN>
bool fixed_polymul_try(const Program &fwd, const Program &inv, const uint64_t *a, const int64_t *b,
                       uint64_t *c, uint64_t batch, const c64 *t, const c64 *tws, const c64 *utw,
                       hipStream_t s, hipError_t *e) {
  constexpr Program pf = cfft_prog::fixed_program(ALGO, LOG2N, LOG2BN, true);
  constexpr Program pi = cfft_prog::fixed_program(ALGO, LOG2N, LOG2BN, false);
  if (!cfft_prog::same_passes(fwd, pf) || !cfft_prog::same_passes(inv, pi)) return false;
  const uint32_t rf = cfft_fixed::max_radix(pf), ri = cfft_fixed::max_radix(pi);
  const cfft_fixed::Geom g = cfft_fixed::geom_for(LOG2N, 2048, rf > ri ? rf : ri);
  // 16 points per thread for radix-16 programs and above 4096 points
  constexpr int E = ALGO >= 6 || LOG2N > 12 ? 16 : 8;
  if (g.E != E) return false;
  const uint64_t blocks = (batch + g.G - 1) / g.G;
  const void *k = reinterpret_cast<const void *>(k_polymul_fixed<E, ALGO, LOG2N, LOG2BN>);
  *e = cfft_launch::ensure_lds(k, g.lds);
  if (*e == hipSuccess)
    hipLaunchKernelGGL((k_polymul_fixed<E, ALGO, LOG2N, LOG2BN>), dim3((uint32_t)blocks),
                       dim3(g.threads), g.lds, s, a, b, c, batch, t, tws, utw);
  if (*e == hipSuccess) *e = hipGetLastError();
  return true;
}

template <int ALGO, int LOG2N, int LOG2BN>
bool fixed_fft_try(const Program &prog, c64 *d, uint64_t ntrans, uint64_t stride, const c64 *t,
                   hipStream_t s, hipError_t *e) {
  const bool fwd = prog.npass != 0 && prog.p[0].fwd != 0;
  const Program want = fwd ? cfft_prog::fixed_program(ALGO, LOG2N, LOG2BN, true)
                           : cfft_prog::fixed_program(ALGO, LOG2N, LOG2BN, false);
  if (!cfft_prog::same_passes(prog, want)) return false;
  const cfft_fixed::Geom g = cfft_fixed::geom_for(LOG2N, 2048, cfft_fixed::max_radix(prog));
  constexpr int E = ALGO >= 6 || LOG2N > 12 ? 16 : 8;
  if (g.E != E) return false;
  const uint64_t blocks = (ntrans + g.G - 1) / g.G;
  const void *k = fwd ? reinterpret_cast<const void *>(k_fft_fixed<E, ALGO, LOG2N, LOG2BN, true>)
                      : reinterpret_cast<const void *>(k_fft_fixed<E, ALGO, LOG2N, LOG2BN, false>);
  *e = cfft_launch::ensure_lds(k, g.lds);
  if (*e != hipSuccess) return true;
  if (fwd)
    hipLaunchKernelGGL((k_fft_fixed<E, ALGO, LOG2N, LOG2BN, true>), dim3((uint32_t)blocks),
                       dim3(g.threads), g.lds, s, d, ntrans, stride, t);
  else
    hipLaunchKernelGGL((k_fft_fixed<E, ALGO, LOG2N, LOG2BN, false>), dim3((uint32_t)blocks),
                       dim3(g.threads), g.lds, s, d, ntrans, stride, t);
  *e = hipGetLastError();
  return true;
}

}  // namespace
