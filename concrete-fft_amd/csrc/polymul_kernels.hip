// polymul_kernels.hip -- register-resident kernels for Dif16 plans with
// base_n = n (512 <= n <= 4096) or base_n = 1024 (n = 2048, 4096), and Dif8
// plans with n = 1024: plain forward / inverse transforms and the fused
// negacyclic products at N = 512, 1024 and 4096 ... 16384 (N = 2048 has a
// kernel of its own, polymul_n2048_kernels.hip).
#include <hip/hip_runtime.h>

#include "launch.hpp"
#include "lds_attr.hpp"
#include "r16_exec.hpp"
#include "r8_exec.hpp"


using namespace cfft_dev;
using cfft_launch::dc64;

namespace {

// twist zeta_j and its scaled conjugate conj(zeta_j) / n (exact: n is a power
// of two), the latter recomputed instead of loaded
__device__ __forceinline__ c64 untwist_from(c64 z, double inv_n) {
  return mk(z.re * inv_n, (-z.im) * inv_n);
}

template <int LOG2N, bool FWD>
__global__ void __launch_bounds__(R16<LOG2N>::T) k_r16_transform(c64 *__restrict__ data,
                                                                  uint64_t batch, uint64_t stride,
                                                                  const c64 *__restrict__ tw) {
  using S = R16<LOG2N>;
  __shared__ c64 lds[S::lds_words];
  const uint32_t l = threadIdx.x;
  const uint64_t t = blockIdx.x;
  if (t >= batch) return;
  c64 *x = data + t * stride;
  c64 v[16];
#pragma unroll
  for (uint32_t m = 0; m < 16; ++m) v[m] = ld_stream(&x[l + S::T * m]);
  r16_transform<LOG2N, FWD>(v, lds, l, tw);
#pragma unroll
  for (uint32_t s = 0; s < 16; ++s) st_stream(v[s], &x[l + S::T * S::end_m(s)]);
}

constexpr uint32_t kPW = 8;  // waves per workgroup
constexpr uint32_t kTwWords = 16 * 64 + 16 * 4;
// the wave-local kernels below keep the padded exchange and the whole block
constexpr size_t kWaveLds = (kTwWords + kPW * R16<10>::lds_words) * sizeof(c64);

template <int LOG2N>
hipError_t launch_transform(bool fwd, dc64 *data, uint64_t batch, uint64_t stride,
                            const dc64 *tw, hipStream_t s) {
  using S = R16<LOG2N>;
  const uint32_t grid = (uint32_t)batch;
  c64 *d = reinterpret_cast<c64 *>(data);
  const c64 *t = reinterpret_cast<const c64 *>(tw);
  if (fwd)
    hipLaunchKernelGGL((k_r16_transform<LOG2N, true>), dim3(grid), dim3(S::T), 0, s, d, batch,
                       stride, t);
  else
    hipLaunchKernelGGL((k_r16_transform<LOG2N, false>), dim3(grid), dim3(S::T), 0, s, d, batch,
                       stride, t);
  return hipGetLastError();
}

// ---- Dif8, n = 1024: one transform / product per 128-thread workgroup ----

template <bool FWD>
__global__ void __launch_bounds__(R8::T) k_r8_transform(c64 *__restrict__ data, uint64_t batch,
                                                        uint64_t stride,
                                                        const c64 *__restrict__ tw) {
  __shared__ c64 x[R8::lds_words];
  const uint32_t l = threadIdx.x;
  const uint64_t t = blockIdx.x;
  if (t >= batch) return;
  c64 *d = data + t * stride;
  c64 v[8];
#pragma unroll
  for (uint32_t m = 0; m < 8; ++m) v[m] = ld_stream(&d[l + R8::T * m]);
  r8_transform<FWD>(v, x, l, tw);
#pragma unroll
  for (uint32_t m = 0; m < 8; ++m) st_stream(v[m], &d[l + R8::T * m]);
}

__global__ void __launch_bounds__(R8::T)
k_r8_polymul(const uint64_t *__restrict__ A, const int64_t *__restrict__ B,
             uint64_t *__restrict__ C, uint64_t batch, const c64 *__restrict__ tw,
             const c64 *__restrict__ twist) {
  constexpr uint32_t n = R8::n;
  __shared__ c64 x[R8::lds_words];
  const uint32_t l = threadIdx.x;
  const uint64_t poly = blockIdx.x;
  if (poly >= batch) return;
  const uint64_t *a = A + poly * (2 * n);
  const int64_t *b = B + poly * (2 * n);
  uint64_t *c = C + poly * (2 * n);
  // A is loaded and transformed first; B's raw words are fetched before that
  // transform (their latency overlaps it) and twisted after it, which keeps
  // the live state under 128 VGPRs
  c64 fa[8], fb[8];
  int64_t rb[16];
#pragma unroll
  for (uint32_t m = 0; m < 8; ++m) {
    const uint32_t j = l + R8::T * m;
    const c64 z = twist[j];
    fa[m] = cmul(z, mk(torus_to_f64(ld_stream(&a[j])), torus_to_f64(ld_stream(&a[j + n]))));
  }
#pragma unroll
  for (uint32_t m = 0; m < 8; ++m) {
    const uint32_t j = l + R8::T * m;
    rb[2 * m] = ld_stream(&b[j]);
    rb[2 * m + 1] = ld_stream(&b[j + n]);
  }
  sched_fence();
  r8_transform<true>(fa, x, l, tw);
  const c64 *tw1 = launder(twist);
#pragma unroll
  for (uint32_t m = 0; m < 8; ++m)
    fb[m] = cmul(tw1[l + R8::T * m], mk(i64_to_f64(rb[2 * m]), i64_to_f64(rb[2 * m + 1])));
  r8_transform<true>(fb, x, l, tw);
#pragma unroll
  for (uint32_t m = 0; m < 8; ++m) fa[m] = cmul(fa[m], fb[m]);
  r8_transform<false>(fa, x, l, tw);
  const double inv_n = 1.0 / (double)n;
  const c64 *tws = launder(twist);
#pragma unroll
  for (uint32_t m = 0; m < 8; ++m) {
    const uint32_t j = l + R8::T * m;
    const c64 r = cmul(untwist_from(tws[j], inv_n), fa[m]);
    st_stream(from_torus(r.re), &c[j]);
    st_stream(from_torus(r.im), &c[j + n]);
  }
}

// ---- Dif16 plan with n = 2048, base_n = 1024: one transform per wavefront ----
// r16l_2048_transform (the plan's radix-2 depth pass in registers around its
// two 1024-point base transforms, run together), 8 waves per workgroup
// sharing the 1024-point twiddle block staged in LDS.  In place: each wave
// reads its whole transform before writing it.
template <bool FWD>
__global__ void __launch_bounds__(64 * kPW) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_w2048_transform(c64 *__restrict__ data, uint64_t batch, uint64_t stride,
                  const c64 *__restrict__ tw, const c64 *__restrict__ tw_blk) {
  using S = R16<10>;
  extern __shared__ c64 smem[];
  c64 *ltw = smem;
  for (uint32_t i = threadIdx.x; i < kTwWords; i += blockDim.x) ltw[i] = tw_blk[i];
  __syncthreads();
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint64_t t = (uint64_t)blockIdx.x * kPW + wave;
  if (t >= batch) return;
  c64 *x = smem + kTwWords + wave * S::lds_words;
  const uint32_t l = threadIdx.x & 63u;
  c64 *d = data + t * stride;
  c64 va[16], vb[16];
#pragma unroll
  for (uint32_t m = 0; m < 16; ++m) {
    va[m] = ld_stream(&d[l + 64 * m]);
    vb[m] = ld_stream(&d[1024 + l + 64 * m]);
  }
  r16l_2048_transform<FWD>(va, vb, x, l, ltw, tw);
#pragma unroll
  for (uint32_t s = 0; s < 16; ++s) {
    const uint32_t q = l + 64 * S::end_m(s);
    st_stream(va[s], &d[q]);
    st_stream(vb[s], &d[1024 + q]);
  }
}

// ---- Dif16 plan with base_n = 1024, n = R * 1024 (R = 4: n = 4096) ----
// One transform on R waves, 8 / R transforms per 8-wave workgroup.  Forward:
// the plan's one depth pass (DEPTH_FWD, radix R: butterfly j < 1024 reads the
// points k * 1024 + j, its output k, times T[k j], goes to chunk bitrev(k))
// runs in registers right after the loads and writes the chunks to LDS; one
// barrier; wave c runs chunk c's 1024-point base transform in registers and
// stores it.  Inverse: wave c loads and inverse-transforms chunk c into LDS;
// one barrier; the inverse depth pass (conj T[k j] before the butterfly)
// runs right before the stores.  The same operations as the uneven batch's
// depth tiles, so the plan's bits.  In place: every point is read before the
// barrier and written after it.  Chunk c lives in LDS at c * 1088 points (r16
// padding) and doubles as the base transform's exchange buffer.
template <int LOG2N, bool FWD>
__global__ void __launch_bounds__(64 * kPW) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_depth_transform(c64 *__restrict__ data, uint64_t batch, uint64_t stride,
                  const c64 *__restrict__ tw, const c64 *__restrict__ tw_blk) {
  using S = R16<10>;
  constexpr uint32_t n = 1u << LOG2N;
  constexpr int R = 1 << (LOG2N - 10);
  constexpr uint32_t TJ = n / 16;        // threads per transform
  constexpr uint32_t G = 64 * kPW / TJ;  // transforms per workgroup
  constexpr uint32_t NBD = 16 / R;       // depth butterflies per thread
  constexpr uint32_t CH = S::lds_words;
  extern __shared__ c64 smem[];
  c64 *ltw = smem;
  for (uint32_t i = threadIdx.x; i < kTwWords; i += blockDim.x) ltw[i] = tw_blk[i];
  const uint32_t g = threadIdx.x / TJ, t = threadIdx.x % TJ;
  const uint64_t tr = (uint64_t)blockIdx.x * G + g;
  // a group past the batch works on the workgroup's first transform and
  // stores nothing; every wave reaches both barriers
  const bool active = tr < batch;
  c64 *d = data + (active ? tr : (uint64_t)blockIdx.x * G) * stride;
  c64 *X = smem + kTwWords + g * (R * CH);
  const uint32_t c = t >> 6, l = t & 63;
  c64 *xc = X + c * CH;
  if (FWD) {
    __syncthreads();  // twiddle block staged
#pragma unroll 1
    for (uint32_t i = 0; i < NBD; ++i) {
      const uint32_t j = t + TJ * i;
      c64 v[R];
#pragma unroll
      for (int k = 0; k < R; ++k) v[k] = ld_stream(&d[(uint32_t)k * 1024 + j]);
      butterfly<R, true>(v);
#pragma unroll
      for (int k = 1; k < R; ++k) v[k] = cmul(tw[(uint32_t)k * j], v[k]);
#pragma unroll
      for (int k = 0; k < R; ++k) X[(uint32_t)bitrev_digit<R>(k) * CH + r16_pad(j)] = v[k];
    }
    __syncthreads();
    c64 v[16];
#pragma unroll
    for (uint32_t m = 0; m < 16; ++m) v[m] = xc[r16_pad(l + 64 * m)];
    r16l_transform<10, true>(v, v, xc, l, ltw);
    if (active) {
#pragma unroll
      for (uint32_t s = 0; s < 16; ++s) st_stream(v[s], &d[c * 1024 + l + 64 * S::end_m(s)]);
    }
  } else {
    c64 v[16], u[16];
#pragma unroll
    for (uint32_t s = 0; s < 16; ++s) v[s] = ld_stream(&d[c * 1024 + l + 64 * S::end_m(s)]);
    __syncthreads();  // twiddle block staged
    r16_end_to_input<10>(v, u);
    r16l_transform<10, false>(u, u, xc, l, ltw);
    wave_lds_fence();
#pragma unroll
    for (uint32_t s = 0; s < 16; ++s) xc[r16_pad(l + 64 * S::end_m(s))] = u[s];
    __syncthreads();
#pragma unroll 1
    for (uint32_t i = 0; i < NBD; ++i) {
      const uint32_t j = t + TJ * i;
      c64 w[R];
#pragma unroll
      for (int k = 0; k < R; ++k) w[k] = X[(uint32_t)bitrev_digit<R>(k) * CH + r16_pad(j)];
#pragma unroll
      for (int k = 1; k < R; ++k) {
        c64 f = tw[(uint32_t)k * j];
        f.im = -f.im;
        w[k] = cmul(f, w[k]);
      }
      butterfly<R, false>(w);
      if (active) {
#pragma unroll
        for (int k = 0; k < R; ++k) st_stream(w[k], &d[(uint32_t)k * 1024 + j]);
      }
    }
  }
}

// ---- negacyclic product, Dif16 plan with base_n = 1024, n = R * 1024 ----
// (R = 2, 4, 8: N = 4096, 8192, 16384.)  One product on R waves, 8 / R
// products per 8-wave workgroup, each operand through the structure of
// k_depth_transform: fold + twist + the plan's depth pass in registers after
// the loads, chunks to LDS, one 1024-point base transform per wave.  A's
// spectrum stays in registers while B's is formed; the pointwise product,
// the inverse base transform, the inverse depth pass, untwist (with 1/n)
// and torus rounding follow before the stores.  The same operations in the
// same order as the generic fused product (k_polymul_generic), so the same
// bits.

// One operand's raw depth-pass inputs for one thread: point (i, k) is
// q = k * 1024 + t + TJ i, held as the two u64 words at q and q + n.
template <int LOG2N>
struct DepthRaw {
  static constexpr int R = 1 << (LOG2N - 10);
  static constexpr uint32_t NBD = 16 / R;
  uint64_t re[NBD][R], im[NBD][R];
};

template <int LOG2N, class T>
__device__ __forceinline__ void depth_raw_load(DepthRaw<LOG2N> &d, const T *x, uint32_t t) {
  constexpr uint32_t n = 1u << LOG2N, TJ = n / 16;
#pragma unroll
  for (uint32_t i = 0; i < DepthRaw<LOG2N>::NBD; ++i)
#pragma unroll
    for (int k = 0; k < DepthRaw<LOG2N>::R; ++k) {
      const uint32_t q = (uint32_t)k * 1024 + t + TJ * i;
      d.re[i][k] = (uint64_t)ld_stream(&x[q]);
      d.im[i][k] = (uint64_t)ld_stream(&x[q + n]);
    }
}

// Fold + twist + the plan's depth pass (radix R, as k_depth_transform) on
// inputs already in registers (TORUS: A's u64 torus words, else B's i64
// digits), chunks to LDS
template <int LOG2N, bool TORUS>
__device__ __forceinline__ void depth_raw_to_lds(const DepthRaw<LOG2N> &d, c64 *X, uint32_t t,
                                                 const c64 *tw, const c64 *twist) {
  constexpr uint32_t n = 1u << LOG2N, TJ = n / 16, CH = R16<10>::lds_words;
  constexpr int R = DepthRaw<LOG2N>::R;
#pragma unroll
  for (uint32_t i = 0; i < DepthRaw<LOG2N>::NBD; ++i) {
    const uint32_t j = t + TJ * i;
    c64 v[R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
      const uint32_t q = (uint32_t)k * 1024 + j;
      const c64 x = TORUS ? mk(torus_to_f64(d.re[i][k]), torus_to_f64(d.im[i][k]))
                          : mk(i64_to_f64((int64_t)d.re[i][k]), i64_to_f64((int64_t)d.im[i][k]));
      v[k] = cmul(twist[q], x);
    }
    butterfly<R, true>(v);
#pragma unroll
    for (int k = 1; k < R; ++k) v[k] = cmul(tw[(uint32_t)k * j], v[k]);
#pragma unroll
    for (int k = 0; k < R; ++k) X[(uint32_t)bitrev_digit<R>(k) * CH + r16_pad(j)] = v[k];
    sched_fence();  // one butterfly's twist and twiddle loads in flight at a time
  }
}

// Persistent form: each workgroup loops over its share of the batch, G
// products at a time.  PRE_B: B's inputs are loaded while A's base transform
// runs; PRE_A: the next round's A inputs are loaded while the inverse base
// transform runs (so only the first round waits for A).  A group past the
// batch works on the workgroup's first product of that round and stores
// nothing; every wave runs every round and reaches every barrier.
template <int LOG2N, bool PRE_B, bool PRE_A>
__global__ void __launch_bounds__(64 * kPW) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_depth_polymul(const uint64_t *__restrict__ A, const int64_t *__restrict__ B,
                uint64_t *__restrict__ C, uint64_t batch, const c64 *__restrict__ tw,
                const c64 *__restrict__ tw_blk, const c64 *__restrict__ twist,
                const c64 *__restrict__ untwist) {
  using S = R16<10>;
  constexpr uint32_t n = 1u << LOG2N;
  constexpr int R = 1 << (LOG2N - 10);
  constexpr uint32_t TJ = n / 16;        // threads per product
  constexpr uint32_t G = 64 * kPW / TJ;  // products per workgroup and round
  constexpr uint32_t NBD = 16 / R;
  constexpr uint32_t CH = S::lds_words;
  extern __shared__ c64 smem[];
  c64 *ltw = smem;
  for (uint32_t i = threadIdx.x; i < kTwWords; i += blockDim.x) ltw[i] = tw_blk[i];
  const uint64_t step = (uint64_t)gridDim.x * G;
  // round r of this workgroup: products blockIdx.x * G + r * step + g
  auto offset_of = [&](uint64_t first, uint32_t g) -> uint64_t {
    const uint64_t pr = first + g;
    return (pr < batch ? pr : first) * (2ull * n);
  };
  uint64_t first = (uint64_t)blockIdx.x * G;
  DepthRaw<LOG2N> ra;
  if (PRE_A) depth_raw_load<LOG2N>(ra, A + offset_of(first, threadIdx.x / TJ), threadIdx.x % TJ);
  __syncthreads();  // twiddle block staged
  while (first < batch) {
    // recomputed per round (opaque) so the per-lane addresses below are not
    // hoisted out of the loop and kept live across it
    const uint32_t tid = launder_u32(threadIdx.x);
    const uint32_t g = tid / TJ, t = tid % TJ;
    c64 *X = smem + kTwWords + g * (R * CH);
    const uint32_t c = t >> 6, l = t & 63;
    c64 *xc = X + c * CH;
    const bool active = first + g < batch;
    const uint64_t off = offset_of(first, g);
    if (!PRE_A) depth_raw_load<LOG2N>(ra, A + off, t);
    depth_raw_to_lds<LOG2N, true>(ra, X, t, tw, twist);
    __syncthreads();
    c64 fa[16];
    {
      DepthRaw<LOG2N> rb;
      if (PRE_B) {
        depth_raw_load<LOG2N>(rb, B + off, t);
        sched_fence();
      }
#pragma unroll
      for (uint32_t m = 0; m < 16; ++m) fa[m] = xc[r16_pad(l + 64 * m)];
      r16l_transform<10, true>(fa, fa, xc, l, ltw);
      __syncthreads();  // every chunk read before B's depth pass overwrites it
      if (!PRE_B) depth_raw_load<LOG2N>(rb, B + off, t);
      depth_raw_to_lds<LOG2N, false>(rb, X, t, tw, twist);
    }
    __syncthreads();
    {
      c64 fb[16];
#pragma unroll
      for (uint32_t m = 0; m < 16; ++m) fb[m] = xc[r16_pad(l + 64 * m)];
      r16l_transform<10, true>(fb, fb, xc, l, ltw);
#pragma unroll
      for (uint32_t s = 0; s < 16; ++s) fa[s] = cmul(fa[s], fb[s]);
    }
    const uint64_t next = first + step;
    if (PRE_A) {
      // the next round's A (past the batch: this round's, never used)
      depth_raw_load<LOG2N>(ra, A + offset_of(next < batch ? next : first, g), t);
      sched_fence();
    }
    c64 u[16];
    r16_end_to_input<10>(fa, u);
    r16l_transform<10, false>(u, u, xc, l, ltw);
    wave_lds_fence();
#pragma unroll
    for (uint32_t s = 0; s < 16; ++s) xc[r16_pad(l + 64 * S::end_m(s))] = u[s];
    __syncthreads();
    uint64_t *cp = C + off;
#pragma unroll 1
    for (uint32_t i = 0; i < NBD; ++i) {
      const uint32_t j = t + TJ * i;
      c64 w[R];
#pragma unroll
      for (int k = 0; k < R; ++k) w[k] = X[(uint32_t)bitrev_digit<R>(k) * CH + r16_pad(j)];
#pragma unroll
      for (int k = 1; k < R; ++k) {
        c64 f = tw[(uint32_t)k * j];
        f.im = -f.im;
        w[k] = cmul(f, w[k]);
      }
      butterfly<R, false>(w);
      if (active) {
#pragma unroll
        for (int k = 0; k < R; ++k) {
          const uint32_t q = (uint32_t)k * 1024 + j;
          const c64 r = cmul(untwist[q], w[k]);
          st_stream(from_torus(r.re), &cp[q]);
          st_stream(from_torus(r.im), &cp[q + n]);
        }
      }
    }
    __syncthreads();  // every chunk read before the next round's depth pass
    first = next;
  }
}

// ---- Dif16 plan with n = base_n = 1024: NV transforms per wavefront ----
// r16w_transform (twiddle reads shared by the NV transforms), 8 waves per
// workgroup sharing the twiddle block staged in LDS.  In place: each wave
// reads its transforms before writing them.  A wave whose second transform
// would lie past the batch repeats its first one and stores it once.
template <bool FWD, int NV>
__global__ void __launch_bounds__(64 * kPW) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_w1024_transform(c64 *__restrict__ data, uint64_t batch, uint64_t stride,
                  const c64 *__restrict__ tw_r16) {
  using S = R16<10>;
  extern __shared__ c64 smem[];
  c64 *ltw = smem;
  for (uint32_t i = threadIdx.x; i < kTwWords; i += blockDim.x) ltw[i] = tw_r16[i];
  __syncthreads();
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint64_t t0 = ((uint64_t)blockIdx.x * kPW + wave) * NV;
  if (t0 >= batch) return;
  const bool two = NV == 2 && t0 + 1 < batch;
  c64 *x = smem + kTwWords + wave * S::lds_words;
  const uint32_t l = threadIdx.x & 63u;
  c64 *d0 = data + t0 * stride;
  c64 *d1 = two ? d0 + stride : d0;
  c64 va[16], vb[16];
#pragma unroll
  for (uint32_t m = 0; m < 16; ++m) {
    va[m] = ld_stream(&d0[l + 64 * m]);
    if (NV == 2) vb[m] = ld_stream(&d1[l + 64 * m]);
  }
  r16w_transform<FWD, NV>(va, vb, x, l, ltw);
#pragma unroll
  for (uint32_t s = 0; s < 16; ++s) {
    const uint32_t q = l + 64 * S::end_m(s);
    st_stream(va[s], &d0[q]);
    if (NV == 2 && two) st_stream(vb[s], &d1[q]);
  }
}

// ---- Dif16 plan with n = base_n = 512: 2 x NV transforms per wavefront ----
// Each 32-lane half of a wave runs NV transforms (r16l_transform<9>, twiddles
// read from the 1024-point block at column stride 2, twiddle reads shared),
// 8 waves per workgroup sharing that block staged in LDS; each half has its
// own 544-point exchange buffer.  A half whose transform would lie past the
// batch repeats the last one and stores nothing.
template <bool FWD, int NV>
__global__ void __launch_bounds__(64 * kPW) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_w512_transform(c64 *__restrict__ data, uint64_t batch, uint64_t stride,
                 const c64 *__restrict__ tw_blk) {
  using S = R16<9>;
  extern __shared__ c64 smem[];
  c64 *ltw = smem;
  for (uint32_t i = threadIdx.x; i < kTwWords; i += blockDim.x) ltw[i] = tw_blk[i];
  __syncthreads();
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint64_t first = ((uint64_t)blockIdx.x * kPW + wave) * (2 * NV);
  if (first >= batch) return;
  const uint32_t half = (threadIdx.x >> 5) & 1u, l = threadIdx.x & 31u;
  const uint64_t t0 = first + half * NV;
  const bool st0 = t0 < batch, st1 = NV == 2 && t0 + 1 < batch;
  c64 *d0 = data + (st0 ? t0 : batch - 1) * stride;
  c64 *d1 = st1 ? d0 + stride : d0;
  c64 *x = smem + kTwWords + (wave * 2 + half) * S::lds_words;
  c64 va[16], vb[16];
#pragma unroll
  for (uint32_t m = 0; m < 16; ++m) {
    va[m] = ld_stream(&d0[l + 32 * m]);
    if (NV == 2) vb[m] = ld_stream(&d1[l + 32 * m]);
  }
  r16l_transform<9, FWD, NV>(va, vb, x, l, ltw);
#pragma unroll
  for (uint32_t s = 0; s < 16; ++s) {
    const uint32_t q = l + 32 * S::end_m(s);
    if (st0) st_stream(va[s], &d0[q]);
    if (st1) st_stream(vb[s], &d1[q]);
  }
}

// ---- negacyclic product, N = 512 / 1024 on a Dif16 plan with n = base_n ----
// (n = 2^LOG2N = 256 or 512.)  One product per group of T = n / 16 lanes
// (a quarter- or half-wave), the same operations as the n = 1024 product
// kernel: fold + twist at the loads, both forward transforms together
// (r16l_transform<LOG2N>, twiddle reads shared), pointwise product, inverse,
// untwist and torus rounding at the stores.  8 waves per workgroup share the
// 1024-point twiddle block in LDS (read at column stride 1024 / n).
// Persistent: at most one workgroup per CU, the waves step over the batch
// 64 / T products at a time, and B of a group's next product is loaded
// during the inverse.  A group past the batch repeats the last product and
// stores nothing.
template <int LOG2N>
__global__ void __launch_bounds__(64 * kPW) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_wsub_polymul(const uint64_t *__restrict__ A, const int64_t *__restrict__ B,
               uint64_t *__restrict__ C, uint64_t batch, const c64 *__restrict__ tw_blk,
               const c64 *__restrict__ twist) {
  using S = R16<LOG2N>;
  constexpr uint32_t n = S::n;
  constexpr uint32_t PW = 64 / S::T;  // products per wave
  static_assert(PW * S::lds_words <= R16<10>::lds_words, "per-wave exchange space");
  extern __shared__ c64 smem[];
  c64 *ltw = smem;
  for (uint32_t i = threadIdx.x; i < kTwWords; i += blockDim.x) ltw[i] = tw_blk[i];
  __syncthreads();
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint64_t step = (uint64_t)gridDim.x * kPW * PW;
  uint64_t first = ((uint64_t)blockIdx.x * kPW + wave) * PW;
  uint64_t pb[32];  // raw words of B for the group's current product
  if (first < batch) {
    const uint32_t g0 = (threadIdx.x & 63u) / S::T, l0 = threadIdx.x % S::T;
    const uint64_t p0 = first + g0 < batch ? first + g0 : batch - 1;
    const int64_t *b = B + p0 * (2 * n);
#pragma unroll
    for (uint32_t m = 0; m < 16; ++m) {
      pb[2 * m] = (uint64_t)ld_stream(&b[l0 + S::T * m]);
      pb[2 * m + 1] = (uint64_t)ld_stream(&b[l0 + S::T * m + n]);
    }
  }
  for (; first < batch; first += step) {
    // recomputed per product so per-lane addresses are not kept live across
    // the loop
    const uint32_t lane = launder_u32(threadIdx.x) & 63u;
    const uint32_t grp = lane / S::T, l = lane % S::T;
    const uint64_t p = first + grp;
    const bool act = p < batch;
    const uint64_t pc = act ? p : batch - 1;
    const uint64_t *a = A + pc * (2 * n);
    uint64_t *c = C + pc * (2 * n);
    c64 *x = smem + kTwWords + wave * R16<10>::lds_words + grp * S::lds_words;
    const double inv_n = 1.0 / (double)n;
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
        const c64 z = twist[l + S::T * m];
        fa[m] = cmul(z, mk(torus_to_f64(ra[2 * m]), torus_to_f64(ra[2 * m + 1])));
        v[m] = cmul(z, mk(i64_to_f64((int64_t)pb[2 * m]), i64_to_f64((int64_t)pb[2 * m + 1])));
      }
    }
    r16l_transform<LOG2N, true, 2>(fa, v, x, l, ltw);
    c64 z[16];
#pragma unroll
    for (uint32_t s = 0; s < 16; ++s) fa[s] = cmul(fa[s], v[s]);
    r16_end_to_input<LOG2N>(fa, z);
    sched_fence();
    {
      // B of the group's next product (past the batch: the L2-resident twist
      // table, 2 n words, never used)
      const uint64_t nx = p + step;
      const int64_t *bn =
          nx < batch ? B + nx * (2 * n) : reinterpret_cast<const int64_t *>(twist);
#pragma unroll
      for (uint32_t m = 0; m < 16; ++m) {
        pb[2 * m] = (uint64_t)ld_stream(&bn[l + S::T * m]);
        pb[2 * m + 1] = (uint64_t)ld_stream(&bn[l + S::T * m + n]);
      }
    }
    sched_fence();
    r16l_transform<LOG2N, false, 1>(z, z, x, l, ltw);
    const c64 *tws = launder(twist);
#pragma unroll
    for (uint32_t s = 0; s < 16; ++s) {
      const uint32_t j = l + S::T * S::end_m(s);
      const c64 r = cmul(untwist_from(tws[j], inv_n), z[s]);
      if (act) {
        st_stream(from_torus(r.re), &c[j]);
        st_stream(from_torus(r.im), &c[j + n]);
      }
    }
  }
}

// ---- Dif16 plan with n = 1024 and base_n = 512: one transform per half-wave ----
// r16l_depth2_transform<9> (the radix-2 depth pass in registers around the
// two 512-point base transforms, run together), two transforms per wave,
// 8 waves per workgroup sharing the 1024-point block in LDS; each half has
// its own 544-point exchange buffer.  A half past the batch repeats the last
// transform and stores nothing.
template <bool FWD>
__global__ void __launch_bounds__(64 * kPW) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_wd2_512_transform(c64 *__restrict__ data, uint64_t batch, uint64_t stride,
                    const c64 *__restrict__ tw, const c64 *__restrict__ tw_blk) {
  using S = R16<9>;
  extern __shared__ c64 smem[];
  c64 *ltw = smem;
  for (uint32_t i = threadIdx.x; i < kTwWords; i += blockDim.x) ltw[i] = tw_blk[i];
  __syncthreads();
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint64_t first = ((uint64_t)blockIdx.x * kPW + wave) * 2;
  if (first >= batch) return;
  const uint32_t half = (threadIdx.x >> 5) & 1u, l = threadIdx.x & 31u;
  const uint64_t t = first + half;
  const bool act = t < batch;
  c64 *d = data + (act ? t : batch - 1) * stride;
  c64 *x = smem + kTwWords + (wave * 2 + half) * S::lds_words;
  c64 va[16], vb[16];
#pragma unroll
  for (uint32_t m = 0; m < 16; ++m) {
    va[m] = ld_stream(&d[l + 32 * m]);
    vb[m] = ld_stream(&d[512 + l + 32 * m]);
  }
  r16l_depth2_transform<9, FWD>(va, vb, x, l, ltw, tw);
  // store addresses formed again here rather than kept live across the
  // transform
  const uint32_t ls = launder_u32(threadIdx.x) & 31u;
  c64 *ds = data + (act ? t : batch - 1) * stride;
#pragma unroll
  for (uint32_t s = 0; s < 16; ++s) {
    const uint32_t q = ls + 32 * S::end_m(s);
    if (act) {
      st_stream(va[s], &ds[q]);
      st_stream(vb[s], &ds[512 + q]);
    }
  }
}

// ---- negacyclic product, N = 2048 on a Dif16 plan with base_n = 512 ----
// Persistent: one product per wavefront at a time, 8 waves per workgroup, at
// most one workgroup per CU, products strided over the waves; B of a wave's
// next product is loaded during its inverse.  The plan's radix-2 depth pass
// (src/unordered.rs:391-489) pairs the 1024-point transform's points j and
// 512 + j.  Lane 32 h + l (h = 0, 1; l < 32) holds chunk h -- points
// 512 h + l + 32 m in its slots m -- so each depth butterfly takes one value
// from each half-wave: the two halves swap their values through the wave's
// exchange buffer (slot m of lane i at 64 m + i), and each then forms its own
// output of the butterfly: chunk 0 (own value a) b + a, chunk 1 (own value b)
// a + (-b) -- the same IEEE results as a + b and a - b, since addition
// commutes and subtraction is addition of the negated operand.  Each
// half then runs its chunk's 512-point base transforms (r16l_transform<9>, A
// and B together, in its own 544-point half of the buffer).  Otherwise the
// operations of the generic fused product: fold + twist at the loads,
// pointwise product, inverse (base transforms, then the depth pass with
// conj(T[q]) on chunk 1's point before the butterfly), untwist with the
// exact 1/n and torus rounding at the stores.

// the depth butterfly's partner swap for slots m0 .. m0 + 7 of one operand
__device__ __forceinline__ void half_swap_write(const c64 *u, c64 *xw, uint32_t lane,
                                                uint32_t m0, uint32_t base) {
#pragma unroll
  for (uint32_t m = 0; m < 8; ++m) xw[base + 64 * m + lane] = u[m0 + m];
}

__global__ void __launch_bounds__(64 * kPW) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_wd2_512_polymul(const uint64_t *__restrict__ A, const int64_t *__restrict__ B,
                  uint64_t *__restrict__ C, uint64_t batch, const c64 *__restrict__ tw,
                  const c64 *__restrict__ tw_blk, const c64 *__restrict__ twist) {
  using S = R16<9>;
  constexpr uint32_t n = 1024;
  static_assert(2 * S::lds_words == R16<10>::lds_words, "two half-wave buffers per wave");
  extern __shared__ c64 smem[];
  c64 *ltw = smem;
  for (uint32_t i = threadIdx.x; i < kTwWords; i += blockDim.x) ltw[i] = tw_blk[i];
  __syncthreads();
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  c64 *xw = smem + kTwWords + wave * R16<10>::lds_words;  // the wave's 1088 points
  const uint64_t step = (uint64_t)gridDim.x * kPW;
  uint64_t p = (uint64_t)blockIdx.x * kPW + wave;
  uint64_t pb[32];  // raw words of B for the wave's current product
  if (p < batch) {
    const uint32_t j0 = 512 * ((threadIdx.x >> 5) & 1u) + (threadIdx.x & 31u);
    const int64_t *b = B + p * (2 * n);
#pragma unroll
    for (uint32_t m = 0; m < 16; ++m) {
      pb[2 * m] = (uint64_t)ld_stream(&b[j0 + 32 * m]);
      pb[2 * m + 1] = (uint64_t)ld_stream(&b[j0 + 32 * m + n]);
    }
  }
  for (; p < batch; p += step) {
    // recomputed per product so per-lane addresses are not kept live across
    // the loop
    const uint32_t lane = launder_u32(threadIdx.x) & 63u, h = lane >> 5, l = lane & 31u;
    const uint32_t pl = lane ^ 32u;  // the other half's lane with the same l
    const uint64_t *a = A + p * (2 * n);
    uint64_t *c = C + p * (2 * n);
    c64 *x = xw + h * S::lds_words;  // this half's 544 points
    const double inv_n = 1.0 / (double)n;
    c64 fa[16], v[16];
    {
      uint64_t ra[32];
#pragma unroll
      for (uint32_t m = 0; m < 16; ++m) {
        const uint32_t j = 512 * h + l + 32 * m;
        ra[2 * m] = ld_stream(&a[j]);
        ra[2 * m + 1] = ld_stream(&a[j + n]);
      }
      sched_fence();
#pragma unroll
      for (uint32_t m = 0; m < 16; ++m) {
        const c64 z = twist[512 * h + l + 32 * m];
        fa[m] = cmul(z, mk(torus_to_f64(ra[2 * m]), torus_to_f64(ra[2 * m + 1])));
        v[m] = cmul(z, mk(i64_to_f64((int64_t)pb[2 * m]), i64_to_f64((int64_t)pb[2 * m + 1])));
      }
    }
    // DEPTH_FWD radix 2: chunk 0 gets a + b, chunk 1 (a - b) T[j], j = l + 32 m;
    // eight slots of A and B per swap (1024 points of the buffer)
    {
      const c64 *twd = launder(tw);
#pragma unroll
      for (uint32_t m0 = 0; m0 < 16; m0 += 8) {
        wave_lds_fence();
        half_swap_write(fa, xw, lane, m0, 0);
        half_swap_write(v, xw, lane, m0, 512);
        wave_lds_fence();
#pragma unroll
        for (uint32_t m1 = 0; m1 < 8; m1 += 4) {
          c64 w[4];
#pragma unroll
          for (uint32_t e = 0; e < 4; ++e) w[e] = twd[l + 32 * (m0 + m1 + e)];
#pragma unroll
          for (uint32_t e = 0; e < 4; ++e) {
            const uint32_t m = m0 + m1 + e;
            const c64 oa = xw[64 * (m1 + e) + pl], ob = xw[512 + 64 * (m1 + e) + pl];
            const c64 sa = h ? mk(-fa[m].re, -fa[m].im) : fa[m];
            const c64 sb = h ? mk(-v[m].re, -v[m].im) : v[m];
            const c64 ta = cadd(oa, sa), tb = cadd(ob, sb);
            fa[m] = h ? cmul(w[e], ta) : ta;
            v[m] = h ? cmul(w[e], tb) : tb;
          }
          sched_fence();
        }
      }
      wave_lds_fence();
    }
    r16l_transform<9, true, 2>(fa, v, x, l, ltw);
    c64 z[16];
#pragma unroll
    for (uint32_t s = 0; s < 16; ++s) fa[s] = cmul(fa[s], v[s]);
    r16_end_to_input<9>(fa, z);
    sched_fence();
    {
      // B of the wave's next product (past its last product: the L2-resident
      // twist table, 2048 words, never used)
      const uint64_t nx = p + step;
      const int64_t *bn = nx < batch ? B + nx * (2 * n) : reinterpret_cast<const int64_t *>(twist);
#pragma unroll
      for (uint32_t m = 0; m < 16; ++m) {
        const uint32_t j = 512 * h + l + 32 * m;
        pb[2 * m] = (uint64_t)ld_stream(&bn[j]);
        pb[2 * m + 1] = (uint64_t)ld_stream(&bn[j + n]);
      }
    }
    sched_fence();
    r16l_transform<9, false, 1>(z, z, x, l, ltw);
    // DEPTH_INV radix 2: u1 = conj(T[q]) chunk 1's point; chunk 0 gets u0 + u1,
    // chunk 1 u0 - u1, q = l + 32 end_m(s)
    {
      const c64 *twd = launder(tw);
#pragma unroll
      for (uint32_t s0 = 0; s0 < 16; s0 += 4) {
        c64 w[4];
#pragma unroll
        for (uint32_t e = 0; e < 4; ++e) {
          w[e] = twd[l + 32 * S::end_m(s0 + e)];
          w[e].im = -w[e].im;
        }
#pragma unroll
        for (uint32_t e = 0; e < 4; ++e) z[s0 + e] = h ? cmul(w[e], z[s0 + e]) : z[s0 + e];
      }
      wave_lds_fence();
#pragma unroll
      for (uint32_t s = 0; s < 16; ++s) xw[64 * s + lane] = z[s];
      wave_lds_fence();
#pragma unroll
      for (uint32_t s = 0; s < 16; ++s) {
        const c64 o = xw[64 * s + pl];
        z[s] = cadd(o, h ? mk(-z[s].re, -z[s].im) : z[s]);
      }
    }
    const c64 *tws = launder(twist);
#pragma unroll
    for (uint32_t s = 0; s < 16; ++s) {
      const uint32_t j = 512 * h + l + 32 * S::end_m(s);
      const c64 r = cmul(untwist_from(tws[j], inv_n), z[s]);
      st_stream(from_torus(r.re), &c[j]);
      st_stream(from_torus(r.im), &c[j + n]);
    }
  }
}

}  // namespace

namespace cfft_launch {

hipError_t wd2_512_transform(bool fwd, dc64 *data, uint64_t batch, uint64_t stride,
                             const dc64 *tw, const dc64 *tw_blk, hipStream_t s) {
  if (batch == 0) return hipSuccess;
  const void *k = fwd ? reinterpret_cast<const void *>(k_wd2_512_transform<true>)
                      : reinterpret_cast<const void *>(k_wd2_512_transform<false>);
  hipError_t e = cfft_launch::ensure_lds(k, kWaveLds);
  if (e != hipSuccess) return e;
  c64 *d = reinterpret_cast<c64 *>(data);
  const c64 *t = reinterpret_cast<const c64 *>(tw);
  const c64 *tb = reinterpret_cast<const c64 *>(tw_blk);
  const uint64_t per_wg = (uint64_t)kPW * 2;
  const dim3 grid((uint32_t)((batch + per_wg - 1) / per_wg));
  if (fwd)
    hipLaunchKernelGGL(k_wd2_512_transform<true>, grid, dim3(64 * kPW), kWaveLds, s, d, batch,
                       stride, t, tb);
  else
    hipLaunchKernelGGL(k_wd2_512_transform<false>, grid, dim3(64 * kPW), kWaveLds, s, d, batch,
                       stride, t, tb);
  return hipGetLastError();
}

hipError_t wd2_512_polymul(const uint64_t *a, const int64_t *b, uint64_t *c, uint64_t batch,
                           const dc64 *tw, const dc64 *tw_blk, const dc64 *twist,
                           hipStream_t s) {
  if (batch == 0) return hipSuccess;
  hipError_t e =
      cfft_launch::ensure_lds(reinterpret_cast<const void *>(k_wd2_512_polymul), kWaveLds);
  if (e != hipSuccess) return e;
  int cus = 0;
  e = cu_count(&cus);
  if (e != hipSuccess) return e;
  const uint64_t want = (batch + kPW - 1) / kPW;
  const dim3 grid((uint32_t)(want < (uint64_t)cus ? want : (uint64_t)cus));
  hipLaunchKernelGGL(k_wd2_512_polymul, grid, dim3(64 * kPW), kWaveLds, s, a, b, c, batch,
                     reinterpret_cast<const c64 *>(tw), reinterpret_cast<const c64 *>(tw_blk),
                     reinterpret_cast<const c64 *>(twist));
  return hipGetLastError();
}

hipError_t wsub_polymul(uint32_t log2n, const uint64_t *a, const int64_t *b, uint64_t *c,
                        uint64_t batch, const dc64 *tw_blk, const dc64 *twist, hipStream_t s) {
  if (batch == 0) return hipSuccess;
  if (log2n != 8 && log2n != 9) return hipErrorInvalidValue;
  const void *k = log2n == 8 ? reinterpret_cast<const void *>(k_wsub_polymul<8>)
                             : reinterpret_cast<const void *>(k_wsub_polymul<9>);
  hipError_t e = cfft_launch::ensure_lds(k, kWaveLds);
  if (e != hipSuccess) return e;
  int cus = 0;
  e = cu_count(&cus);
  if (e != hipSuccess) return e;
  const uint64_t per_wg = (uint64_t)kPW * (log2n == 8 ? 4 : 2);
  const uint64_t want = (batch + per_wg - 1) / per_wg;
  const dim3 grid((uint32_t)(want < (uint64_t)cus ? want : (uint64_t)cus)), block(64 * kPW);
  const c64 *tb = reinterpret_cast<const c64 *>(tw_blk), *z = reinterpret_cast<const c64 *>(twist);
  if (log2n == 8)
    hipLaunchKernelGGL(k_wsub_polymul<8>, grid, block, kWaveLds, s, a, b, c, batch, tb, z);
  else
    hipLaunchKernelGGL(k_wsub_polymul<9>, grid, block, kWaveLds, s, a, b, c, batch, tb, z);
  return hipGetLastError();
}

hipError_t w512_transform(bool fwd, dc64 *data, uint64_t batch, uint64_t stride,
                          const dc64 *tw_blk, hipStream_t s) {
  if (batch == 0) return hipSuccess;
  constexpr int NV = 2;
  const void *k = fwd ? reinterpret_cast<const void *>(k_w512_transform<true, NV>)
                      : reinterpret_cast<const void *>(k_w512_transform<false, NV>);
  hipError_t e = cfft_launch::ensure_lds(k, kWaveLds);
  if (e != hipSuccess) return e;
  c64 *d = reinterpret_cast<c64 *>(data);
  const c64 *t = reinterpret_cast<const c64 *>(tw_blk);
  const uint64_t per_wg = (uint64_t)kPW * 2 * NV;
  const dim3 grid((uint32_t)((batch + per_wg - 1) / per_wg));
  if (fwd)
    hipLaunchKernelGGL((k_w512_transform<true, NV>), grid, dim3(64 * kPW), kWaveLds, s, d,
                       batch, stride, t);
  else
    hipLaunchKernelGGL((k_w512_transform<false, NV>), grid, dim3(64 * kPW), kWaveLds, s, d,
                       batch, stride, t);
  return hipGetLastError();
}

hipError_t w1024_transform(bool fwd, dc64 *data, uint64_t batch, uint64_t stride,
                           const dc64 *tw_r16, hipStream_t s) {
  if (batch == 0) return hipSuccess;
  constexpr int NV = 2;
  const void *k = fwd ? reinterpret_cast<const void *>(k_w1024_transform<true, NV>)
                      : reinterpret_cast<const void *>(k_w1024_transform<false, NV>);
  hipError_t e = cfft_launch::ensure_lds(k, kWaveLds);
  if (e != hipSuccess) return e;
  c64 *d = reinterpret_cast<c64 *>(data);
  const c64 *t = reinterpret_cast<const c64 *>(tw_r16);
  const uint64_t per_wg = (uint64_t)kPW * NV;
  const dim3 grid((uint32_t)((batch + per_wg - 1) / per_wg));
  if (fwd)
    hipLaunchKernelGGL((k_w1024_transform<true, NV>), grid, dim3(64 * kPW), kWaveLds, s, d,
                       batch, stride, t);
  else
    hipLaunchKernelGGL((k_w1024_transform<false, NV>), grid, dim3(64 * kPW), kWaveLds, s, d,
                       batch, stride, t);
  return hipGetLastError();
}

hipError_t w2048_transform(bool fwd, dc64 *data, uint64_t batch, uint64_t stride, const dc64 *tw,
                           const dc64 *tw_blk, hipStream_t s) {
  if (batch == 0) return hipSuccess;
  const void *k = fwd ? reinterpret_cast<const void *>(k_w2048_transform<true>)
                      : reinterpret_cast<const void *>(k_w2048_transform<false>);
  hipError_t e = cfft_launch::ensure_lds(k, kWaveLds);
  if (e != hipSuccess) return e;
  c64 *d = reinterpret_cast<c64 *>(data);
  const c64 *t = reinterpret_cast<const c64 *>(tw);
  const c64 *tb = reinterpret_cast<const c64 *>(tw_blk);
  const dim3 grid((uint32_t)((batch + kPW - 1) / kPW));
  if (fwd)
    hipLaunchKernelGGL(k_w2048_transform<true>, grid, dim3(64 * kPW), kWaveLds, s, d, batch,
                       stride, t, tb);
  else
    hipLaunchKernelGGL(k_w2048_transform<false>, grid, dim3(64 * kPW), kWaveLds, s, d, batch,
                       stride, t, tb);
  return hipGetLastError();
}

hipError_t depth_transform(bool fwd, uint32_t log2n, dc64 *data, uint64_t batch, uint64_t stride,
                           const dc64 *tw, const dc64 *tw_blk, hipStream_t s) {
  if (batch == 0) return hipSuccess;
  if (log2n != 12) return hipErrorInvalidValue;
  constexpr size_t lds = (kTwWords + 8 * R16<10>::lds_words) * sizeof(c64);
  const void *k = fwd ? reinterpret_cast<const void *>(k_depth_transform<12, true>)
                      : reinterpret_cast<const void *>(k_depth_transform<12, false>);
  hipError_t e = cfft_launch::ensure_lds(k, lds);
  if (e != hipSuccess) return e;
  c64 *d = reinterpret_cast<c64 *>(data);
  const c64 *t = reinterpret_cast<const c64 *>(tw);
  const c64 *tb = reinterpret_cast<const c64 *>(tw_blk);
  const dim3 grid((uint32_t)((batch + 1) / 2)), block(64 * kPW);
  if (fwd)
    hipLaunchKernelGGL((k_depth_transform<12, true>), grid, block, lds, s, d, batch, stride, t, tb);
  else
    hipLaunchKernelGGL((k_depth_transform<12, false>), grid, block, lds, s, d, batch, stride, t, tb);
  return hipGetLastError();
}

hipError_t depth_polymul(uint32_t log2n, const uint64_t *a, const int64_t *b, uint64_t *c,
                         uint64_t batch, const dc64 *tw, const dc64 *tw_blk, const dc64 *twist,
                         const dc64 *untwist, hipStream_t s) {
  if (batch == 0) return hipSuccess;
  if (log2n < 11 || log2n > 13) return hipErrorInvalidValue;
  constexpr size_t lds = (kTwWords + 8 * R16<10>::lds_words) * sizeof(c64);
  // both operands loaded a step ahead at N = 4096 and 8192; at N = 16384
  // (8 points per depth butterfly) B's early load would spill, so only the
  // next round's A is (measured: 38.0 / 18.1 / 7.0 M products/s against
  // 31.6 / 14.1 / 6.1 M with every load at its depth pass)
  const void *k = log2n == 11   ? reinterpret_cast<const void *>(k_depth_polymul<11, true, true>)
                  : log2n == 12 ? reinterpret_cast<const void *>(k_depth_polymul<12, true, true>)
                                : reinterpret_cast<const void *>(k_depth_polymul<13, false, true>);
  hipError_t e = cfft_launch::ensure_lds(k, lds);
  if (e != hipSuccess) return e;
  int cus = 0;
  e = cfft_launch::cu_count(&cus);
  if (e != hipSuccess) return e;
  const c64 *t = reinterpret_cast<const c64 *>(tw), *tb = reinterpret_cast<const c64 *>(tw_blk);
  const c64 *z = reinterpret_cast<const c64 *>(twist), *u = reinterpret_cast<const c64 *>(untwist);
  const uint64_t per_wg = 8u >> (log2n - 10);  // 8 / R products per round
  const uint64_t want = (batch + per_wg - 1) / per_wg;
  const dim3 grid((uint32_t)(want < (uint64_t)cus ? want : (uint64_t)cus)), block(64 * kPW);
  if (log2n == 11)
    hipLaunchKernelGGL((k_depth_polymul<11, true, true>), grid, block, lds, s, a, b, c, batch, t,
                       tb, z, u);
  else if (log2n == 12)
    hipLaunchKernelGGL((k_depth_polymul<12, true, true>), grid, block, lds, s, a, b, c, batch, t,
                       tb, z, u);
  else
    hipLaunchKernelGGL((k_depth_polymul<13, false, true>), grid, block, lds, s, a, b, c, batch, t,
                       tb, z, u);
  return hipGetLastError();
}

bool r16_supported(uint32_t log2n) { return log2n >= 9 && log2n <= 12; }
bool r8_supported(uint32_t log2n) { return log2n == 10; }

hipError_t r8_transform(bool fwd, dc64 *data, uint64_t batch, uint64_t stride, const dc64 *tw,
                        hipStream_t s) {
  if (batch == 0) return hipSuccess;
  c64 *d = reinterpret_cast<c64 *>(data);
  const c64 *t = reinterpret_cast<const c64 *>(tw);
  if (fwd)
    hipLaunchKernelGGL(k_r8_transform<true>, dim3((uint32_t)batch), dim3(R8::T), 0, s, d, batch,
                       stride, t);
  else
    hipLaunchKernelGGL(k_r8_transform<false>, dim3((uint32_t)batch), dim3(R8::T), 0, s, d, batch,
                       stride, t);
  return hipGetLastError();
}

hipError_t r8_polymul(const uint64_t *a, const int64_t *b, uint64_t *c, uint64_t batch,
                      const dc64 *tw, const dc64 *twist, hipStream_t s) {
  if (batch == 0) return hipSuccess;
  hipLaunchKernelGGL(k_r8_polymul, dim3((uint32_t)batch), dim3(R8::T), 0, s, a, b, c, batch,
                     reinterpret_cast<const c64 *>(tw), reinterpret_cast<const c64 *>(twist));
  return hipGetLastError();
}

hipError_t r16_transform(uint32_t log2n, bool fwd, dc64 *data, uint64_t batch, uint64_t stride,
                         const dc64 *tw, hipStream_t s) {
  if (batch == 0) return hipSuccess;
  switch (log2n) {
    case 9: return launch_transform<9>(fwd, data, batch, stride, tw, s);
    case 10: return w1024_transform(fwd, data, batch, stride, tw, s);
    case 11: return launch_transform<11>(fwd, data, batch, stride, tw, s);
    case 12: return launch_transform<12>(fwd, data, batch, stride, tw, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace cfft_launch
