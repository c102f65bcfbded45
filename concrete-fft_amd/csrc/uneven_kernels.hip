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
// Two tile bodies, in two kernels (inlined side by side in one persistent
// loop they need more than 256 VGPRs):
//  - generic: the pass-program executor (any plan, 16 points per thread),
//    a workgroup barrier around every pass;
//  - Dif16 (plans Dif16 with base_n = min(n, 1024), 256 <= n <= 8192): the
//    register-resident Dif16 schedule on one wavefront per 1024-point base
//    transform.  For n > 1024 the plan's single depth pass (radix n / 1024,
//    src/unordered.rs:391-489) runs in registers right after the loads
//    (forward) and right before the stores (inverse), so a tile needs two
//    workgroup barriers; jobs with n <= 1024 are wave-local (no barrier
//    inside a tile).  One twiddle block (the 1024-point base transform's)
//    serves every class and is staged in LDS once per workgroup.
#include <hip/hip_runtime.h>

#include "launch.hpp"
#include "lds_attr.hpp"
#include "r16_exec.hpp"
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

// ---- Dif16 body ----

// NV (1 or 2) jobs of n = 2^LOG2N points (256 <= n <= 1024) on T = n/16
// lanes of one wavefront, transformed together (every twiddle, twist and
// untwist read shared): fold, twist, forward, Fourier store, inverse,
// untwist, store.
// A null job is skipped (its lanes compute on zeros); `x` is the group's
// exchange buffer, `twl` the 1024-point twiddle block in LDS.
template <int LOG2N, int NV>
__device__ __forceinline__ void job2_r16_local(const UClass &C, const UJob *ja, const UJob *jb,
                                               const double *in, c64 *fourier, double *out,
                                               c64 *x, uint32_t l, const c64 *twl) {
  using S = R16<LOG2N>;
  constexpr uint32_t n = S::n, T = S::T;
  JobPtrs A{nullptr, nullptr, nullptr}, B{nullptr, nullptr, nullptr};
  if (ja) A = job_ptrs(*ja, in, fourier, out);
  if (jb) B = job_ptrs(*jb, in, fourier, out);
  const c64 *twist = reinterpret_cast<const c64 *>(C.twist);
  const c64 *untwist = reinterpret_cast<const c64 *>(C.untwist);
  c64 va[16], vb[16];
  // every coefficient load first (their latency overlaps), then the twist
  // four points at a time (its loads hit the cache)
#pragma unroll
  for (uint32_t m = 0; m < 16; ++m) {
    const uint32_t j = l + T * m;
    va[m] = ja ? mk(ld_stream(&A.x[j]), ld_stream(&A.x[j + n])) : mk(0.0, 0.0);
    if (NV == 2) vb[m] = jb ? mk(ld_stream(&B.x[j]), ld_stream(&B.x[j + n])) : mk(0.0, 0.0);
  }
  sched_fence();
#pragma unroll
  for (uint32_t m0 = 0; m0 < 16; m0 += 4) {
#pragma unroll
    for (uint32_t m = m0; m < m0 + 4; ++m) {
      const c64 w = twist[l + T * m];
      va[m] = cmul(w, va[m]);
      if (NV == 2) vb[m] = cmul(w, vb[m]);
    }
    sched_fence();
  }
  r16l_transform<LOG2N, true, NV>(va, vb, x, l, twl);
#pragma unroll
  for (uint32_t s = 0; s < 16; ++s) {
    const uint32_t j = l + T * S::end_m(s);
    if (ja) st_stream(va[s], &A.f[j]);
    if (NV == 2 && jb) st_stream(vb[s], &B.f[j]);
  }
  sched_fence();  // stores issued before the inverse reuses their registers
  c64 ua[16], ub[16];
  r16_end_to_input<LOG2N>(va, ua);
  if (NV == 2) r16_end_to_input<LOG2N>(vb, ub);
  r16l_transform<LOG2N, false, NV>(ua, ub, x, l, twl);
  sched_fence();
  // untwist eight points at a time (all sixteen loads at once beside both
  // results exceed the register budget)
#pragma unroll
  for (uint32_t s0 = 0; s0 < 16; s0 += 8) {
#pragma unroll
    for (uint32_t s = s0; s < s0 + 8; ++s) {
      const uint32_t j = l + T * S::end_m(s);
      const c64 w = untwist[j];
      if (ja) {
        const c64 r = cmul(w, ua[s]);
        st_stream(r.re, &A.y[j]);
        st_stream(r.im, &A.y[j + n]);
      }
      if (NV == 2 && jb) {
        const c64 r = cmul(w, ub[s]);
        st_stream(r.re, &B.y[j]);
        st_stream(r.im, &B.y[j + n]);
      }
    }
    sched_fence();
  }
}

// jobs tile.first .. + tile.count, NV per group of T lanes (group g: jobs
// NV g .. NV g + NV - 1); group g uses the exchange buffer at g * lds_words
// (so wave w's groups share w * 1088)
template <int LOG2N>
__device__ __forceinline__ void group_r16_local(const UClass &C, const UTile &tile,
                                                const UJob *__restrict__ jobs,
                                                const uint32_t *__restrict__ order,
                                                const double *in, c64 *fourier, double *out,
                                                c64 *xbase, const c64 *twl, uint32_t tid) {
  using S = R16<LOG2N>;
  constexpr int NV = cfft_ub::jobs_per_group(LOG2N);
  uint32_t g = tid / S::T;
  const uint32_t l = tid % S::T;
  // a group that is a whole wavefront: its job pointers are wave-uniform
  // (scalar registers instead of 64-bit vector pairs)
  if constexpr (S::T == 64) g = __builtin_amdgcn_readfirstlane(g);
  const uint32_t j0 = NV * g;
  const UJob *ja = j0 < tile.count ? &jobs[order[tile.first + j0]] : nullptr;
  const UJob *jb = NV == 2 && j0 + 1 < tile.count ? &jobs[order[tile.first + j0 + 1]] : nullptr;
  job2_r16_local<LOG2N, NV>(C, ja, jb, in, fourier, out, xbase + g * S::lds_words, l, twl);
}

// One job of n = 2048 points on one wavefront (r16l_2048_transform: the
// plan's radix-2 depth pass in registers around its two 1024-point base
// transforms, run together).  Lane l holds the points l + 64 m and
// 1024 + l + 64 m.  A null job is skipped (its lanes compute on zeros).
__device__ __forceinline__ void job_r16_2048(const UClass &C, const UJob *job, const double *in,
                                             c64 *fourier, double *out, c64 *x, uint32_t l,
                                             const c64 *twl) {
  using S = R16<10>;
  constexpr uint32_t n = 2048;
  JobPtrs P{nullptr, nullptr, nullptr};
  if (job) P = job_ptrs(*job, in, fourier, out);
  const c64 *tw = reinterpret_cast<const c64 *>(C.tw);
  const c64 *twist = reinterpret_cast<const c64 *>(C.twist);
  const c64 *untwist = reinterpret_cast<const c64 *>(C.untwist);
  c64 va[16], vb[16];
#pragma unroll
  for (uint32_t m = 0; m < 16; ++m) {
    const uint32_t j = l + 64 * m;
    va[m] = job ? mk(ld_stream(&P.x[j]), ld_stream(&P.x[j + n])) : mk(0.0, 0.0);
    vb[m] = job ? mk(ld_stream(&P.x[1024 + j]), ld_stream(&P.x[1024 + j + n])) : mk(0.0, 0.0);
  }
  sched_fence();
#pragma unroll
  for (uint32_t m0 = 0; m0 < 16; m0 += 4) {
#pragma unroll
    for (uint32_t m = m0; m < m0 + 4; ++m) {
      const uint32_t j = l + 64 * m;
      va[m] = cmul(twist[j], va[m]);
      vb[m] = cmul(twist[1024 + j], vb[m]);
    }
    sched_fence();
  }
  r16l_2048_transform<true>(va, vb, x, l, twl, tw);
  if (job) {
#pragma unroll
    for (uint32_t s = 0; s < 16; ++s) {
      const uint32_t q = l + 64 * S::end_m(s);
      st_stream(va[s], &P.f[q]);
      st_stream(vb[s], &P.f[1024 + q]);
    }
  }
  sched_fence();  // stores issued before the inverse reuses their registers
  c64 ua[16], ub[16];
  r16_end_to_input<10>(va, ua);
  r16_end_to_input<10>(vb, ub);
  r16l_2048_transform<false>(ua, ub, x, l, twl, tw);
  sched_fence();
#pragma unroll
  for (uint32_t s0 = 0; s0 < 16; s0 += 4) {
#pragma unroll
    for (uint32_t s = s0; s < s0 + 4; ++s) {
      const uint32_t q = l + 64 * S::end_m(s);
      const c64 ra = cmul(untwist[q], ua[s]);
      const c64 rb = cmul(untwist[1024 + q], ub[s]);
      if (job) {
        st_stream(ra.re, &P.y[q]);
        st_stream(ra.im, &P.y[q + n]);
        st_stream(rb.re, &P.y[1024 + q]);
        st_stream(rb.im, &P.y[1024 + q + n]);
      }
    }
    sched_fence();
  }
}

__device__ __forceinline__ void run_group_r16_local(const UClass &C, const UTile &tile,
                                                    const UJob *__restrict__ jobs,
                                                    const uint32_t *__restrict__ order,
                                                    const double *in, c64 *fourier, double *out,
                                                    c64 *xbase, const c64 *twl, uint32_t tid) {
  switch (C.log2n) {
    case 8: group_r16_local<8>(C, tile, jobs, order, in, fourier, out, xbase, twl, tid); break;
    case 9: group_r16_local<9>(C, tile, jobs, order, in, fourier, out, xbase, twl, tid); break;
    case 10: group_r16_local<10>(C, tile, jobs, order, in, fourier, out, xbase, twl, tid); break;
    default: {
      // n = 2048: one job per wave, the wave's exchange buffer at 1088 w
      const uint32_t g = __builtin_amdgcn_readfirstlane(tid >> 6);
      const UJob *job = g < tile.count ? &jobs[order[tile.first + g]] : nullptr;
      job_r16_2048(C, job, in, fourier, out, xbase + g * R16<10>::lds_words, tid & 63u, twl);
    } break;
  }
}

// n = 2^LOG2N = R * 1024 (R = 4, 8): one job on R wavefronts.  Depth
// pass (DEPTH_FWD, radix R over the whole job; butterfly j < 1024 reads the
// points k * 1024 + j, its output k, times T[k j], goes to chunk bitrev(k)),
// then wave c runs the base transform of chunk c, forward and inverse, then
// the inverse depth pass (DEPTH_INV: conj T[k j] before the butterfly).
// Chunk c lives in LDS at c * 1088 points (r16 padding); the base transform
// uses that same region as its exchange buffer.
//
// A thread's depth-pass inputs are 16 points (NBD butterflies of R points):
// point (i, k) at p = k * 1024 + t + TJ * i.  The first kAhead of them can be
// held raw (x_p, x_{p+n}) in a DepthIn, loaded a tile ahead; the rest are
// loaded when the tile starts (holding all 16 across the base transforms
// would spill).
constexpr int kAhead = 8;
struct DepthIn {
  double re[kAhead], im[kAhead];
};

template <int LOG2N>
__device__ __forceinline__ void depth_load(const UTile &tile, const UJob *__restrict__ jobs,
                                           const uint32_t *__restrict__ order, const double *in,
                                           uint32_t tid, DepthIn &d) {
  constexpr uint32_t n = 1u << LOG2N;
  constexpr int R = 1 << (LOG2N - 10);
  constexpr uint32_t TJ = n / 16, NBD = 16 / R;
  const uint32_t g = tid / TJ, t = tid % TJ;
  // a group past the tile's last job reads the tile's first job (no branch
  // around the loads: every load is issued, none is waited on here); its
  // results are never stored
  const double *x = in + jobs[order[tile.first + (g < tile.count ? g : 0u)]].in_off;
#pragma unroll
  for (int q = 0; q < kAhead; ++q) {
    const uint32_t p = (uint32_t)(q % R) * 1024 + t + TJ * (uint32_t)(q / R);
    d.re[q] = ld_stream(&x[p]);
    d.im[q] = ld_stream(&x[p + n]);
  }
  (void)NBD;
}

__device__ __forceinline__ void depth_load_any(uint32_t log2n, const UTile &tile,
                                               const UJob *__restrict__ jobs,
                                               const uint32_t *__restrict__ order,
                                               const double *in, uint32_t tid, DepthIn &d) {
  if (log2n == 12) depth_load<12>(tile, jobs, order, in, tid, d);
  else depth_load<13>(tile, jobs, order, in, tid, d);
}

// `after_barrier()` runs right after the first barrier: `din` is free from
// there on (the persistent kernel loads the next tile's inputs into it,
// overlapping the base transforms and the inverse depth pass).
// AHEAD: the first kAhead points are in `din` (else every point is loaded
// here).
template <int LOG2N, bool AHEAD, class AfterBarrier>
__device__ __forceinline__ void tile_r16_depth(const UClass &C, const UTile &tile,
                                               const UJob *__restrict__ jobs,
                                               const uint32_t *__restrict__ order,
                                               const double *in, c64 *fourier, double *out,
                                               c64 *smem, const c64 *twl, uint32_t tid,
                                               DepthIn &din, AfterBarrier after_barrier) {
  using S = R16<10>;
  constexpr uint32_t n = 1u << LOG2N;
  constexpr int R = 1 << (LOG2N - 10);
  constexpr uint32_t TJ = n / 16;        // threads per job
  constexpr uint32_t NBD = 16 / R;       // depth butterflies per thread
  constexpr uint32_t CH = S::lds_words;  // LDS points per chunk
  const uint32_t g = tid / TJ, t = tid % TJ;
  const bool active = g < tile.count;
  c64 *X = smem + g * (R * CH);
  JobPtrs P{nullptr, nullptr, nullptr};
  if (active) P = job_ptrs(jobs[order[tile.first + g]], in, fourier, out);
  const c64 *tw = reinterpret_cast<const c64 *>(C.tw);
  const c64 *twist = reinterpret_cast<const c64 *>(C.twist);
  const c64 *untwist = reinterpret_cast<const c64 *>(C.untwist);
  {
    c64 v[NBD][R];
#pragma unroll
    for (uint32_t i = 0; i < NBD; ++i) {
      const uint32_t j = t + TJ * i;
#pragma unroll
      for (int k = 0; k < R; ++k) {
        const uint32_t p = (uint32_t)k * 1024 + j;
        const int q = (int)i * R + k;
        if (AHEAD && q < kAhead)
          v[i][k] = cmul(twist[p], mk(din.re[q < kAhead ? q : 0], din.im[q < kAhead ? q : 0]));
        else
          v[i][k] = active ? cmul(twist[p], mk(ld_stream(&P.x[p]), ld_stream(&P.x[p + n])))
                           : mk(0.0, 0.0);
      }
    }
#pragma unroll
    for (uint32_t i = 0; i < NBD; ++i) {
      const uint32_t j = t + TJ * i;
      butterfly<R, true>(v[i]);
#pragma unroll
      for (int k = 1; k < R; ++k) v[i][k] = cmul(tw[(uint32_t)k * j], v[i][k]);
#pragma unroll
      for (int k = 0; k < R; ++k) X[(uint32_t)bitrev_digit<R>(k) * CH + r16_pad(j)] = v[i][k];
    }
  }
  __syncthreads();
  after_barrier();
  {
    const uint32_t c = t >> 6, l = t & 63;
    c64 *xc = X + c * CH;
    c64 v[16];
#pragma unroll
    for (uint32_t m = 0; m < 16; ++m) v[m] = xc[r16_pad(l + 64 * m)];
    r16l_transform<10, true>(v, v, xc, l, twl);
    if (active) {
#pragma unroll
      for (uint32_t s = 0; s < 16; ++s) st_stream(v[s], &P.f[c * 1024 + l + 64 * S::end_m(s)]);
    }
    c64 u[16];
    r16_end_to_input<10>(v, u);
    r16l_transform<10, false>(u, u, xc, l, twl);
    wave_lds_fence();
#pragma unroll
    for (uint32_t s = 0; s < 16; ++s) xc[r16_pad(l + 64 * S::end_m(s))] = u[s];
  }
  __syncthreads();
  // one butterfly at a time: its R points, R - 1 twiddles and R untwist
  // factors are all that is live beside the inputs loaded ahead
#pragma unroll 1
  for (uint32_t i = 0; i < NBD; ++i) {
    const uint32_t j = t + TJ * i;
    c64 v[R];
#pragma unroll
    for (int k = 0; k < R; ++k) v[k] = X[(uint32_t)bitrev_digit<R>(k) * CH + r16_pad(j)];
#pragma unroll
    for (int k = 1; k < R; ++k) {
      c64 w = tw[(uint32_t)k * j];
      w.im = -w.im;
      v[k] = cmul(w, v[k]);
    }
    butterfly<R, false>(v);
    if (active) {
#pragma unroll
      for (int k = 0; k < R; ++k) {
        const uint32_t p = (uint32_t)k * 1024 + j;
        const c64 r = cmul(untwist[p], v[k]);
        st_stream(r.re, &P.y[p]);
        st_stream(r.im, &P.y[p + n]);
      }
    }
  }
}

// a depth tile whose inputs were not loaded ahead
__device__ __forceinline__ void run_tile_r16_depth(const UClass &C, const UTile &tile,
                                                   const UJob *__restrict__ jobs,
                                                   const uint32_t *__restrict__ order,
                                                   const double *in, c64 *fourier, double *out,
                                                   c64 *smem, const c64 *twl, uint32_t tid) {
  DepthIn din;
  auto none = [] {};
  switch (C.log2n) {
    case 12:
      tile_r16_depth<12, false>(C, tile, jobs, order, in, fourier, out, smem, twl, tid, din, none);
      break;
    default:
      tile_r16_depth<13, false>(C, tile, jobs, order, in, fourier, out, smem, twl, tid, din, none);
      break;
  }
}

__device__ __forceinline__ void stage_r16_twiddles(const c64 *__restrict__ src, c64 *twl) {
  for (uint32_t i = threadIdx.x; i < cfft_ub::kTwR16Points; i += kThreads) twl[i] = src[i];
}

// Tiles of the Dif16 classes on persistent workgroups.  With `counter`
// (zeroed before the launch) the tiles come from a device work queue split
// into P = min(kQueueParts, grid) partitions: partition p holds the tiles
// t = p + P i, has its own counter (counter[p * kCounterStride], each on its
// own cache line) and serves the workgroups w with w % P == p (workgroups
// are dealt to the XCDs round-robin, so a partition's fetches stay on one
// XCD).  A workgroup takes its next index when a tile starts and publishes
// it when the tile ends, so the atomic's round trip overlaps the tile.
// Without `counter` the tiles are dealt round-robin (workgroup w: tiles w,
// w + G, ...).  Device-scope atomics on one address are serialised (measured
// ~50 ns each); the partitions divide that cost, the static deal removes it.
//
// The schedule queues the depth classes (n = 4096, 8192) first.  While a
// workgroup's tiles are depth tiles, each tile's inputs are loaded into
// registers during the previous tile, right after its first barrier (the
// next index is published before that barrier), so the loads overlap the
// base transforms.  The remaining tiles run as before; a depth tile met
// there loads its own inputs.
__global__ void __launch_bounds__(kThreads) k_uneven_r16_persistent(
    const UJob *__restrict__ jobs, const uint32_t *__restrict__ order,
    const UTile *__restrict__ tiles, uint32_t ntiles, const UClass *__restrict__ classes,
    const c64 *__restrict__ tw_block, uint32_t *counter, const double *in, c64 *fourier,
    double *out) {
  // every partition needs a workgroup: P = min(kQueueParts, grid)
  const uint32_t P = gridDim.x < cfft_ub::kQueueParts ? gridDim.x : cfft_ub::kQueueParts;
  extern __shared__ c64 smem[];
  __shared__ uint32_t s_next[2];
  c64 *twl = smem + cfft_ub::kLdsPoints;
  const bool queue = counter != nullptr;
  const uint32_t part = blockIdx.x % P;
  uint32_t *ctr = queue ? counter + part * cfft_ub::kCounterStride : nullptr;
  // queue index i of partition `part` -> tile part + P i
  stage_r16_twiddles(tw_block, twl);
  if (queue && threadIdx.x == 0) s_next[0] = part + P * atomicAdd(ctr, 1u);
  __syncthreads();
  uint32_t t = queue ? __builtin_amdgcn_readfirstlane(s_next[0]) : blockIdx.x;
  uint32_t slot = 1;
  // depth tiles, inputs loaded a tile ahead
  if (t < ntiles && classes[tiles[t].cls].log2n >= 12) {
    DepthIn din;
    {
      const UTile tile = tiles[t];
      depth_load_any(classes[tile.cls].log2n, tile, jobs, order, in,
                     launder_u32(threadIdx.x), din);
    }
    bool ahead = true;
    while (ahead) {
      uint32_t next = 0;
      if (queue && threadIdx.x == 0) next = part + P * atomicAdd(ctr, 1u);
      const UTile tile = tiles[t];
      const UClass &C = classes[tile.cls];
      const uint32_t tid = launder_u32(threadIdx.x);
      uint32_t tn = 0;
      if (queue && threadIdx.x == 0) s_next[slot] = next;
      auto load_next = [&] {
        tn = queue ? __builtin_amdgcn_readfirstlane(s_next[slot]) : t + gridDim.x;
        ahead = false;
        if (tn < ntiles) {
          const UTile nt = tiles[tn];
          const uint32_t nl = classes[nt.cls].log2n;
          if (nl >= 12) {
            depth_load_any(nl, nt, jobs, order, in, launder_u32(threadIdx.x), din);
            ahead = true;
          }
        }
      };
      if (C.log2n == 12)
        tile_r16_depth<12, true>(C, tile, jobs, order, in, fourier, out, smem, twl, tid, din,
                                 load_next);
      else
        tile_r16_depth<13, true>(C, tile, jobs, order, in, fourier, out, smem, twl, tid, din,
                                 load_next);
      __syncthreads();
      t = tn;
      slot ^= 1u;
    }
  }
  // queue slots: a slot is written one tile after every thread read it
  for (; t < ntiles; slot ^= 1u) {
    uint32_t next = 0;
    if (queue && threadIdx.x == 0) next = part + P * atomicAdd(ctr, 1u);
    const UTile tile = tiles[t];
    const UClass &C = classes[tile.cls];
    // opaque per tile: address arithmetic of the bodies is not hoisted out
    // of the loop (it would stay live across it and spill)
    const uint32_t tid = launder_u32(threadIdx.x);
    if (C.log2n <= 11)
      run_group_r16_local(C, tile, jobs, order, in, fourier, out, smem, twl, tid);
    else
      run_tile_r16_depth(C, tile, jobs, order, in, fourier, out, smem, twl, tid);
    if (queue && threadIdx.x == 0) s_next[slot] = next;
    __syncthreads();
    t = queue ? __builtin_amdgcn_readfirstlane(s_next[slot]) : t + gridDim.x;
  }
}

// plain form: one workgroup per tile of one class
__global__ void __launch_bounds__(kThreads) k_uneven_r16_tiles(
    const UJob *__restrict__ jobs, const uint32_t *__restrict__ order,
    const UTile *__restrict__ tiles, uint32_t first_tile, const UClass *__restrict__ classes,
    const c64 *__restrict__ tw_block, const double *in, c64 *fourier, double *out) {
  extern __shared__ c64 smem[];
  c64 *twl = smem + cfft_ub::kLdsPoints;
  const UTile tile = tiles[first_tile + blockIdx.x];
  const UClass &C = classes[tile.cls];
  stage_r16_twiddles(tw_block, twl);
  __syncthreads();
  if (C.log2n <= 11)
    run_group_r16_local(C, tile, jobs, order, in, fourier, out, smem, twl, threadIdx.x);
  else
    run_tile_r16_depth(C, tile, jobs, order, in, fourier, out, smem, twl, threadIdx.x);
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

// zeroes `count` queue counters kCounterStride words apart
__global__ void k_reset_counter(uint32_t *counter, uint32_t count) {
  if (threadIdx.x < count) counter[threadIdx.x * cfft_ub::kCounterStride] = 0u;
}

}  // namespace

namespace cfft_ub {

hipError_t prepare() {
  hipError_t e = cfft_launch::ensure_lds(reinterpret_cast<const void *>(k_uneven_persistent),
                                        kLdsBytes);
  if (e == hipSuccess)
    e = cfft_launch::ensure_lds(reinterpret_cast<const void *>(k_uneven_tiles), kLdsBytes);
  if (e == hipSuccess)
    e = cfft_launch::ensure_lds(reinterpret_cast<const void *>(k_uneven_r16_persistent),
                                kR16LdsBytes);
  if (e == hipSuccess)
    e = cfft_launch::ensure_lds(reinterpret_cast<const void *>(k_uneven_r16_tiles), kR16LdsBytes);
  return e;
}

hipError_t launch_persistent(const UJob *jobs, const uint32_t *order, const UTile *tiles,
                             uint32_t ntiles, const UClass *classes, uint32_t *counter,
                             uint32_t grid, const double *in, cfft_launch::dc64 *fourier,
                             double *out, hipStream_t s) {
  hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(64), 0, s, counter, 1u);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (ntiles == 0) return hipSuccess;
  hipLaunchKernelGGL(k_uneven_persistent, dim3(grid), dim3(kThreads), kLdsBytes, s, jobs, order,
                     tiles, ntiles, classes, counter, in, reinterpret_cast<c64 *>(fourier), out);
  return hipGetLastError();
}

hipError_t launch_r16_persistent(const UJob *jobs, const uint32_t *order, const UTile *tiles,
                                 uint32_t ntiles, const UClass *classes,
                                 const cfft_launch::dc64 *tw_block, uint32_t *counter,
                                 uint32_t grid, const double *in, cfft_launch::dc64 *fourier,
                                 double *out, hipStream_t s) {
  if (counter) {
    hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(64), 0, s, counter, kQueueParts);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  if (ntiles == 0) return hipSuccess;
  hipLaunchKernelGGL(k_uneven_r16_persistent, dim3(grid), dim3(kThreads), kR16LdsBytes, s, jobs,
                     order, tiles, ntiles, classes, reinterpret_cast<const c64 *>(tw_block),
                     counter, in, reinterpret_cast<c64 *>(fourier), out);
  return hipGetLastError();
}

hipError_t launch_tiles(const UJob *jobs, const uint32_t *order, const UTile *tiles,
                        uint32_t first_tile, uint32_t count, const UClass *classes,
                        const cfft_launch::dc64 *tw_block, const double *in,
                        cfft_launch::dc64 *fourier, double *out, hipStream_t s) {
  if (count == 0) return hipSuccess;
  c64 *f = reinterpret_cast<c64 *>(fourier);
  if (tw_block)
    hipLaunchKernelGGL(k_uneven_r16_tiles, dim3(count), dim3(kThreads), kR16LdsBytes, s, jobs,
                       order, tiles, first_tile, classes, reinterpret_cast<const c64 *>(tw_block),
                       in, f, out);
  else
    hipLaunchKernelGGL(k_uneven_tiles, dim3(count), dim3(kThreads), kLdsBytes, s, jobs, order,
                       tiles, first_tile, classes, in, f, out);
  return hipGetLastError();
}

}  // namespace cfft_ub
