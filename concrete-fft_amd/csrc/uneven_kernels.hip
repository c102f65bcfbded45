// uneven_kernels.hip -- many f64 negacyclic round-trip jobs of mixed N in
// one launch.
//
// A job is one real polynomial of N = 2n doubles: fold (z_j = x_j + i x_{j+n}),
// twist, forward transform of its size class's plan -> Fourier buffer (n c64,
// plan order), inverse of the same plan, untwist with 1/n -> N doubles.
//
// Jobs are grouped by size class into tiles of up to G jobs (G = 512 / T
// transform groups of T threads each).  The persistent kernel runs one
// 512-thread workgroup per CU; each workgroup takes the next tile from a
// device counter until the queue is empty, so large and small tiles balance
// across CUs (tiles are queued largest first).  The plain form runs the same
// tile body with one launch per class and one workgroup per tile.
//
// The tile body is the pass-program executor (any plan, 16 points per
// thread).  The register-resident Dif16 schedule is not used here: inlined
// beside the generic body in the persistent loop it pushes the kernel past
// 256 VGPRs (spills), and as a called function it needs a scratch stack.
#include <hip/hip_runtime.h>

#include "launch.hpp"
#include "lds_attr.hpp"
#include "stockham_exec.hpp"
#include "uneven.hpp"

using namespace cfft_dev;
using cfft_prog::Program;
using cfft_ub::UClass;
using cfft_ub::UJob;
using cfft_ub::UTile;

namespace {

constexpr uint32_t kThreads = cfft_ub::kThreads;

struct JobPtrs {
  const double *x;
  c64 *f;
  double *y;
};

__device__ __forceinline__ JobPtrs job_ptrs(const UJob &j, const double *in, c64 *fourier,
                                            double *out) {
  return JobPtrs{in + j.in_off, fourier + j.f_off, out + j.out_off};
}

// generic body: run_program over the class programs, 16 points per thread
__device__ __forceinline__ void tile_generic(const UClass &C, const UTile &tile,
                                             const UJob *__restrict__ jobs,
                                             const uint32_t *__restrict__ order,
                                             const double *in, c64 *fourier, double *out,
                                             c64 *smem) {
  constexpr int E = 16;
  const uint32_t log2n = C.log2n;
  const uint32_t n = 1u << log2n;
  const uint32_t T = n / E;
  const uint32_t G = kThreads / T;
  const uint32_t g = threadIdx.x / T, lt = threadIdx.x % T;
  const bool active = g < tile.count;
  c64 *lds = smem + g * lds_words(n);
  JobPtrs P{nullptr, nullptr, nullptr};
  if (active) P = job_ptrs(jobs[order[tile.first + g]], in, fourier, out);
  const c64 *tw = reinterpret_cast<const c64 *>(C.tw);
  const c64 *twist = reinterpret_cast<const c64 *>(C.twist);
  const c64 *untwist = reinterpret_cast<const c64 *>(C.untwist);
  (void)G;
  auto load_x = [&](uint32_t j) -> c64 {
    if (!active) return mk(0.0, 0.0);
    return cmul(twist[j], mk(ld_stream(&P.x[j]), ld_stream(&P.x[j + n])));
  };
  auto store_f = [&](uint32_t pos, c64 v) {
    if (active) st_stream(v, &P.f[pos]);
    lds[lds_pad(pos)] = v;
  };
  run_program<E>(C.fwd, lds, lt, T, tw, load_x, store_f);
  auto load_f = [&](uint32_t pos) -> c64 { return lds[lds_pad(pos)]; };
  auto store_y = [&](uint32_t j, c64 v) {
    if (!active) return;
    const c64 r = cmul(untwist[j], v);
    st_stream(r.re, &P.y[j]);
    st_stream(r.im, &P.y[j + n]);
  };
  run_program<E>(C.inv, lds, lt, T, tw, load_f, store_y);
}

__device__ __forceinline__ void run_tile(const UClass &C, const UTile &tile,
                                         const UJob *__restrict__ jobs,
                                         const uint32_t *__restrict__ order, const double *in,
                                         c64 *fourier, double *out, c64 *smem) {
  tile_generic(C, tile, jobs, order, in, fourier, out, smem);
}

__global__ void __launch_bounds__(kThreads) k_uneven_persistent(
    const UJob *__restrict__ jobs, const uint32_t *__restrict__ order,
    const UTile *__restrict__ tiles, uint32_t ntiles, const UClass *__restrict__ classes,
    uint32_t *counter, const double *in, c64 *fourier, double *out) {
  extern __shared__ c64 smem[];
  __shared__ uint32_t s_tile;
  for (;;) {
    if (threadIdx.x == 0) s_tile = atomicAdd(counter, 1u);
    __syncthreads();
    const uint32_t t = s_tile;
    __syncthreads();
    if (t >= ntiles) break;
    const UTile tile = tiles[t];
    run_tile(classes[tile.cls], tile, jobs, order, in, fourier, out, smem);
  }
}

__global__ void __launch_bounds__(kThreads) k_uneven_tiles(
    const UJob *__restrict__ jobs, const uint32_t *__restrict__ order,
    const UTile *__restrict__ tiles, uint32_t first_tile, const UClass *__restrict__ classes,
    const double *in, c64 *fourier, double *out) {
  extern __shared__ c64 smem[];
  const UTile tile = tiles[first_tile + blockIdx.x];
  run_tile(classes[tile.cls], tile, jobs, order, in, fourier, out, smem);
}

__global__ void k_reset_counter(uint32_t *counter) {
  if (threadIdx.x == 0) counter[0] = 0u;
}

}  // namespace

namespace cfft_ub {

hipError_t prepare() {
  hipError_t e = cfft_launch::ensure_lds(reinterpret_cast<const void *>(k_uneven_persistent),
                                        kLdsBytes);
  if (e != hipSuccess) return e;
  return cfft_launch::ensure_lds(reinterpret_cast<const void *>(k_uneven_tiles), kLdsBytes);
}

hipError_t launch_persistent(const UJob *jobs, const uint32_t *order, const UTile *tiles,
                             uint32_t ntiles, const UClass *classes, uint32_t *counter,
                             uint32_t grid, const double *in, cfft_launch::dc64 *fourier,
                             double *out, hipStream_t s) {
  hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(64), 0, s, counter);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (ntiles == 0) return hipSuccess;
  hipLaunchKernelGGL(k_uneven_persistent, dim3(grid), dim3(kThreads), kLdsBytes, s, jobs, order,
                     tiles, ntiles, classes, counter, in, reinterpret_cast<c64 *>(fourier), out);
  return hipGetLastError();
}

hipError_t launch_tiles(const UJob *jobs, const uint32_t *order, const UTile *tiles,
                        uint32_t first_tile, uint32_t count, const UClass *classes,
                        const double *in, cfft_launch::dc64 *fourier, double *out,
                        hipStream_t s) {
  if (count == 0) return hipSuccess;
  hipLaunchKernelGGL(k_uneven_tiles, dim3(count), dim3(kThreads), kLdsBytes, s, jobs, order,
                     tiles, first_tile, classes, in, reinterpret_cast<c64 *>(fourier), out);
  return hipGetLastError();
}

}  // namespace cfft_ub
