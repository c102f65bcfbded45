// r16_exec.hpp -- register-resident schedule of the reference's Dif16 plan
// for transforms of n = 2^LOG2N points, 9 <= LOG2N <= 12 (base_n == n).
//
// The reference's Dif16 recursion (src/dif16.rs:900-920) runs, for these n,
// two radix-16 Stockham cores (strides 1 and 16) and one end butterfly of
// radix 2^(LOG2N-8) (src/dif2.rs / dif4.rs / dif8.rs / dif16.rs *_end).
// With T = n/16 threads per transform:
//   core s=1   thread l reads points l + T*k        (k < 16), writes 16 l + k
//   core s=16  thread l reads points l + T*k,  writes q + 256 p + 16 k
//              (p = l / 16, q = l % 16)
//   end        thread l reads and writes points l + T*(i + NB*k)
//              (i < NB = 16 / REND butterflies, k < REND)
// so every thread's inputs and outputs of a whole transform are the points
// l + T*m, m < 16 ("lane-linear"): loads, stores, the twist, the pointwise
// product and the hand-over from a forward to an inverse transform happen in
// registers, and only the two interior exchanges go through LDS.  Each
// butterfly and twiddle product is the reference's operation DAG, so results
// are bit-identical to the generic schedule.
#pragma once

#include "device_math.hpp"

namespace cfft_dev {

template <bool FWD>
__device__ __forceinline__ c64 tw_at_impl(const c64 *__restrict__ tw, uint32_t j) {
  c64 w = tw[j];
  if (!FWD) w.im = -w.im;
  return w;
}

__device__ __forceinline__ uint32_t r16_pad(uint32_t i) { return i + (i >> 4); }

template <int LOG2N>
struct R16 {
  static constexpr uint32_t n = 1u << LOG2N;
  static constexpr uint32_t T = n / 16;
  static constexpr uint32_t LOG2REND = LOG2N - 8;
  static constexpr uint32_t REND = 1u << LOG2REND;
  static constexpr uint32_t NB = 16 / REND;
  static constexpr uint32_t lds_words = n + n / 16;
  // point index (minus l, over T) held by register slot s after the end pass:
  // slot s = i * REND + k holds m = i + NB * k
  static __device__ __forceinline__ constexpr uint32_t end_m(uint32_t s) {
    return (s / REND) + NB * (s % REND);
  }
};

// Hides a pointer's value from the optimiser so loads through it are not
// merged across transforms (the three transforms of a product read the same
// twiddles; keeping them live between transforms costs 120 VGPRs).
// (an opaque zero offset keeps the pointer's global address space, unlike
// passing the pointer itself through asm)
template <class P>
__device__ __forceinline__ const P *launder(const P *p) {
  uint32_t zero = 0;
  asm volatile("" : "+s"(zero));
  return p + zero;
}

__device__ __forceinline__ uint32_t launder_u32(uint32_t v) {
  asm volatile("" : "+v"(v));
  return v;
}

// Scheduling fence: keeps the machine scheduler from hoisting every load of a
// transform to its top, which would need far more than 256 VGPRs.
__device__ __forceinline__ void sched_fence() { __builtin_amdgcn_sched_barrier(0); }

// v[k] *= w_k for k = 1..15, twiddles fetched four at a time
template <bool FWD, class Idx>
__device__ __forceinline__ void twiddle16(c64 *v, const c64 *__restrict__ tw, Idx idx) {
#pragma unroll
  for (int g = 0; g < 4; ++g) {
    c64 w[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int k = 4 * g + e;
      if (k > 0) w[e] = tw_at_impl<FWD>(tw, idx(k));
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int k = 4 * g + e;
      if (k > 0) v[k] = cmul(w[e], v[k]);
    }
    sched_fence();
  }
}

template <bool FWD>
__device__ __forceinline__ c64 tw_at(const c64 *__restrict__ tw, uint32_t j) {
  c64 w = tw[j];
  if (!FWD) w.im = -w.im;
  return w;
}

// One transform, data in registers in the lane-linear layout:
// in:  v[m] = point l + T*m
// out: v[s] = point l + T*end_m(s)
// `lds` holds lds_words points; the caller's barrier discipline: this
// function begins with writes to LDS, so any earlier reads of LDS by other
// threads must be complete (it ends with a barrier after its last read).
// `tw` is the plan's r16 twiddle block: rows k = 0..15 of T entries holding
// T[k*l] for lane l (pass 1), followed by rows k = 0..15 of T/16 entries
// holding T[16*k*p] (pass 2).  Same values as the full table, laid out so
// every twiddle load of a wave is contiguous.
template <int LOG2N, bool FWD>
__device__ __forceinline__ void r16_transform(c64 *v, c64 *__restrict__ lds, uint32_t l,
                                              const c64 *__restrict__ tw) {
  using S = R16<LOG2N>;
  tw = launder(tw);
  // ---- core, stride 1: butterfly of points l + T*k, twiddle k*l ----
  butterfly<16, FWD>(v);
  sched_fence();
  twiddle16<FWD>(v, tw, [&](int k) { return (uint32_t)k * S::T + l; });
#pragma unroll
  for (int k = 0; k < 16; ++k) lds[r16_pad(16 * l + (uint32_t)k)] = v[k];
  __syncthreads();
  // ---- core, stride 16 ----
#pragma unroll
  for (int k = 0; k < 16; ++k) v[k] = lds[r16_pad(l + S::T * (uint32_t)k)];
  const uint32_t p = l >> 4, q = l & 15;
  butterfly<16, FWD>(v);
  sched_fence();
  twiddle16<FWD>(v, tw + 16 * S::T, [&](int k) { return (uint32_t)k * (S::T / 16) + p; });
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 16; ++k) lds[r16_pad(q + 256 * p + 16 * (uint32_t)k)] = v[k];
  __syncthreads();
  // ---- end butterflies of radix REND at stride n / REND ----
#pragma unroll
  for (uint32_t i = 0; i < S::NB; ++i) {
    c64 *u = v + i * S::REND;
#pragma unroll
    for (uint32_t k = 0; k < S::REND; ++k) u[k] = lds[r16_pad(l + S::T * (i + S::NB * k))];
    butterfly<S::REND, FWD>(u);
  }
  __syncthreads();
}

// Move register slots from the end-pass layout to the input layout.
template <int LOG2N>
__device__ __forceinline__ void r16_end_to_input(const c64 *src, c64 *dst) {
#pragma unroll
  for (uint32_t s = 0; s < 16; ++s) dst[R16<LOG2N>::end_m(s)] = src[s];
}

}  // namespace cfft_dev

namespace cfft_dev {

// ---- wave-local variant: one transform per wavefront (T == 64) ----
//
// LDS instructions of one wavefront execute in issue order, so an exchange
// among the lanes of a single wave needs no s_barrier; the fence only stops
// the compiler from moving LDS accesses across it.
__device__ __forceinline__ void wave_lds_fence() {
  __asm__ __volatile__("" ::: "memory");
  __builtin_amdgcn_wave_barrier();
}

// gfx950 cross-lane swaps of one f64 held in two VGPRs per lane:
// permlane32: lanes 32-63 of `a` <-> lanes 0-31 of `b`;
// permlane16: odd 16-lane rows of `a` <-> even rows of `b`.
__device__ __forceinline__ void swap_rows32(double &a, double &b) {
  const uint64_t ua = __builtin_bit_cast(uint64_t, a), ub = __builtin_bit_cast(uint64_t, b);
  const auto lo = __builtin_amdgcn_permlane32_swap((uint32_t)ua, (uint32_t)ub, false, false);
  const auto hi =
      __builtin_amdgcn_permlane32_swap((uint32_t)(ua >> 32), (uint32_t)(ub >> 32), false, false);
  a = __builtin_bit_cast(double, ((uint64_t)hi[0] << 32) | lo[0]);
  b = __builtin_bit_cast(double, ((uint64_t)hi[1] << 32) | lo[1]);
}
__device__ __forceinline__ void swap_rows16(double &a, double &b) {
  const uint64_t ua = __builtin_bit_cast(uint64_t, a), ub = __builtin_bit_cast(uint64_t, b);
  const auto lo = __builtin_amdgcn_permlane16_swap((uint32_t)ua, (uint32_t)ub, false, false);
  const auto hi =
      __builtin_amdgcn_permlane16_swap((uint32_t)(ua >> 32), (uint32_t)(ub >> 32), false, false);
  a = __builtin_bit_cast(double, ((uint64_t)hi[0] << 32) | lo[0]);
  b = __builtin_bit_cast(double, ((uint64_t)hi[1] << 32) | lo[1]);
}

// The second exchange of a wave-local Dif16 transform (n = 512 or 1024,
// T = n / 16 lanes per transform, lane l = q + 16 p with q < 16; p is the
// lane's 16-lane row within its transform) in registers.  Through LDS it
// writes slot k of lane (q, p) at q + 256 p + 16 k and reads slot
// i REND + k of lane (q, p) from l + T (i + NB k); matching the two indices,
//   n = 1024 (REND = 4): slot 4 i + k of row p <- slot 4 i + p of row k,
//   n = 512  (REND = 2): slot 2 i + k of row p <- slot 2 i + p of row k,
// with q unchanged: a transpose of the row index against the low register
// bits, which the row swaps perform without LDS.
template <int LOG2N>
__device__ __forceinline__ void r16_exchange2_rows(c64 *v) {
  // (n = 256 has no second exchange: its callers return before it)
  if constexpr (LOG2N == 10) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      c64 *g = v + 4 * i;
      swap_rows32(g[0].re, g[2].re);
      swap_rows32(g[0].im, g[2].im);
      swap_rows32(g[1].re, g[3].re);
      swap_rows32(g[1].im, g[3].im);
      swap_rows16(g[0].re, g[1].re);
      swap_rows16(g[0].im, g[1].im);
      swap_rows16(g[2].re, g[3].re);
      swap_rows16(g[2].im, g[3].im);
    }
  } else if constexpr (LOG2N == 9) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      swap_rows16(v[2 * i].re, v[2 * i + 1].re);
      swap_rows16(v[2 * i].im, v[2 * i + 1].im);
    }
  }
}

template <bool FWD>
__device__ __forceinline__ void twiddle16_x2(c64 *va, c64 *vb, const c64 *tw_row0,
                                             uint32_t row_stride, uint32_t col) {
#pragma unroll
  for (int g = 0; g < 4; ++g) {
    c64 w[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int k = 4 * g + e;
      if (k > 0) w[e] = tw_at_impl<FWD>(tw_row0, (uint32_t)k * row_stride + col);
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int k = 4 * g + e;
      if (k > 0) {
        va[k] = cmul(w[e], va[k]);
        vb[k] = cmul(w[e], vb[k]);
      }
    }
    sched_fence();
  }
}

template <bool FWD>
__device__ __forceinline__ void twiddle16_x1(c64 *v, const c64 *tw_row0, uint32_t row_stride,
                                             uint32_t col) {
#pragma unroll
  for (int g = 0; g < 4; ++g) {
    c64 w[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int k = 4 * g + e;
      if (k > 0) w[e] = tw_at_impl<FWD>(tw_row0, (uint32_t)k * row_stride + col);
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int k = 4 * g + e;
      if (k > 0) v[k] = cmul(w[e], v[k]);
    }
    sched_fence();
  }
}

// n = 1024 points per wave.  `tw` points at the r16 twiddle block in LDS;
// `x` is this wave's exchange buffer (R16<10>::lds_words points).
// Lane-linear layout as r16_transform.  NV = 1 or 2 transforms sharing
// every twiddle read (va, vb).
template <bool FWD, int NV>
__device__ __forceinline__ void r16w_transform(c64 *va, c64 *vb, c64 *x, uint32_t l,
                                               const c64 *tw) {
  using S = R16<10>;
  // core, stride 1
  butterfly<16, FWD>(va);
  if (NV == 2) butterfly<16, FWD>(vb);
  sched_fence();
  if (NV == 2) twiddle16_x2<FWD>(va, vb, tw, S::T, l);
  else twiddle16_x1<FWD>(va, tw, S::T, l);
  // exchange 1
  const uint32_t p = l >> 4;
#pragma unroll
  for (int t = 0; t < NV; ++t) {
    c64 *v = t == 0 ? va : vb;
    wave_lds_fence();
#pragma unroll
    for (int k = 0; k < 16; ++k) x[r16_pad(16 * l + (uint32_t)k)] = v[k];
    wave_lds_fence();
#pragma unroll
    for (int k = 0; k < 16; ++k) v[k] = x[r16_pad(l + S::T * (uint32_t)k)];
  }
  wave_lds_fence();  // last use of x
  // core, stride 16
  butterfly<16, FWD>(va);
  if (NV == 2) butterfly<16, FWD>(vb);
  sched_fence();
  if (NV == 2) twiddle16_x2<FWD>(va, vb, tw + 16 * S::T, S::T / 16, p);
  else twiddle16_x1<FWD>(va, tw + 16 * S::T, S::T / 16, p);
  // exchange 2, in registers
  r16_exchange2_rows<10>(va);
  if (NV == 2) r16_exchange2_rows<10>(vb);
  // end butterflies (radix 4)
#pragma unroll
  for (uint32_t i = 0; i < S::NB; ++i) {
    butterfly<S::REND, FWD>(va + i * S::REND);
    if (NV == 2) butterfly<S::REND, FWD>(vb + i * S::REND);
  }
}

// NV (1 or 2) transforms of n = 2^LOG2N points (LOG2N = 8, 9 or 10) on
// T = n/16 lanes of one wavefront (64 / T lane groups per wave, each with its
// own transforms), sharing every twiddle read.  `tw` is the r16 twiddle
// block of the 1024-point transform (in LDS): the twiddles of a smaller
// transform are its entries at column stride 1024 / n (T_n[j] =
// T_1024[j * 1024 / n], the same sincospi argument).  `x` is the group's
// exchange buffer (R16<LOG2N>::lds_words points; the NV transforms exchange
// through it in turn).  Same layouts and operation order as r16_transform;
// no s_barrier.  At n = 256 the reference's Dif16 is one core at stride 1
// and a radix-16 end at stride 16 (points l + 16 k, in place), so there is
// one exchange and end_m(s) = s.
template <int LOG2N, bool FWD, int NV = 1>
__device__ __forceinline__ void r16l_transform(c64 *va, c64 *vb, c64 *x, uint32_t l,
                                               const c64 *tw) {
  using S = R16<LOG2N>;
  static_assert(S::T <= 64 && LOG2N >= 8, "one transform must fit a wavefront");
  constexpr uint32_t cs = 1u << (10 - LOG2N);  // column stride in the 1024 block
  butterfly<16, FWD>(va);
  if (NV == 2) butterfly<16, FWD>(vb);
  sched_fence();
  if (NV == 2) twiddle16_x2<FWD>(va, vb, tw, 64, l * cs);
  else twiddle16_x1<FWD>(va, tw, 64, l * cs);
  if constexpr (LOG2N == 8) {
#pragma unroll
    for (int t = 0; t < NV; ++t) {
      c64 *v = t == 0 ? va : vb;
      wave_lds_fence();
#pragma unroll
      for (int k = 0; k < 16; ++k) x[r16_pad(16 * l + (uint32_t)k)] = v[k];
      wave_lds_fence();
#pragma unroll
      for (int k = 0; k < 16; ++k) v[k] = x[r16_pad(l + 16 * (uint32_t)k)];
    }
    wave_lds_fence();
    butterfly<16, FWD>(va);
    if (NV == 2) butterfly<16, FWD>(vb);
    return;
  }
  const uint32_t p = l >> 4;
#pragma unroll
  for (int t = 0; t < NV; ++t) {
    c64 *v = t == 0 ? va : vb;
    wave_lds_fence();
#pragma unroll
    for (int k = 0; k < 16; ++k) x[r16_pad(16 * l + (uint32_t)k)] = v[k];
    wave_lds_fence();
#pragma unroll
    for (int k = 0; k < 16; ++k) v[k] = x[r16_pad(l + S::T * (uint32_t)k)];
  }
  wave_lds_fence();  // last use of x
  butterfly<16, FWD>(va);
  if (NV == 2) butterfly<16, FWD>(vb);
  sched_fence();
  if (NV == 2) twiddle16_x2<FWD>(va, vb, tw + 16 * 64, 4, p * cs);
  else twiddle16_x1<FWD>(va, tw + 16 * 64, 4, p * cs);
  // exchange 2, in registers
  r16_exchange2_rows<LOG2N>(va);
  if (NV == 2) r16_exchange2_rows<LOG2N>(vb);
#pragma unroll
  for (uint32_t i = 0; i < S::NB; ++i) {
    if constexpr (S::REND > 1) {
      butterfly<S::REND, FWD>(va + i * S::REND);
      if (NV == 2) butterfly<S::REND, FWD>(vb + i * S::REND);
    }
  }
}

// A Dif16 plan with n = 2 b and base_n = b (b = 2^LOG2B = 512 or 1024) on
// the T = b / 16 lanes of one transform group (a half-wave at b = 512, a
// whole wave at b = 1024).  The plan is one radix-2 depth pass
// (src/unordered.rs:391-489) around two b-point base transforms; the depth
// butterfly pairs the points j and b + j, which lane l holds in the same slot
// (va[m] = point l + T m, vb[m] = point b + l + T m), so the depth pass runs
// in registers and the two base transforms run together (NV = 2, twiddle
// reads shared).
//   forward: in  va[m], vb[m] as above; out va[s] = point l + T end_m(s),
//            vb[s] = point b + l + T end_m(s)   (plan order)
//   inverse: in  va[m] = point l + T m, vb[m] = point b + l + T m (plan
//            order); out va[s] = point q, vb[s] = point b + q with
//            q = l + T end_m(s)
// `tw` is the plan's full table T[j] (the depth twiddles T[j], j < b, in
// global memory), `twl` the 1024-point r16 block in LDS (read at column
// stride 1024 / b), `x` the group's exchange buffer (R16<LOG2B>::lds_words
// points).
template <int LOG2B, bool FWD>
__device__ __forceinline__ void r16l_depth2_transform(c64 *va, c64 *vb, c64 *x, uint32_t l,
                                                      const c64 *twl,
                                                      const c64 *__restrict__ tw) {
  using S = R16<LOG2B>;
  // a forward and an inverse of one job read the same depth twiddles; merged
  // loads would stay live across a whole transform
  tw = launder(tw);
  if constexpr (FWD) {
    // DEPTH_FWD radix 2: (a + b, (a - b) T[j]) to chunks 0 and 1
#pragma unroll
    for (uint32_t m0 = 0; m0 < 16; m0 += 4) {
      c64 w[4];
#pragma unroll
      for (uint32_t e = 0; e < 4; ++e) w[e] = tw[l + S::T * (m0 + e)];
#pragma unroll
      for (uint32_t e = 0; e < 4; ++e) {
        c64 t2[2] = {va[m0 + e], vb[m0 + e]};
        butterfly<2, true>(t2);
        va[m0 + e] = t2[0];
        vb[m0 + e] = cmul(w[e], t2[1]);
      }
      sched_fence();
    }
    r16l_transform<LOG2B, true, 2>(va, vb, x, l, twl);
  } else {
    r16l_transform<LOG2B, false, 2>(va, vb, x, l, twl);
    // DEPTH_INV radix 2: conj(T[q]) on chunk 1's point before the butterfly
#pragma unroll
    for (uint32_t s0 = 0; s0 < 16; s0 += 4) {
      c64 w[4];
#pragma unroll
      for (uint32_t e = 0; e < 4; ++e) {
        w[e] = tw[l + S::T * S::end_m(s0 + e)];
        w[e].im = -w[e].im;
      }
#pragma unroll
      for (uint32_t e = 0; e < 4; ++e) {
        c64 t2[2] = {va[s0 + e], cmul(w[e], vb[s0 + e])};
        butterfly<2, false>(t2);
        va[s0 + e] = t2[0];
        vb[s0 + e] = t2[1];
      }
      sched_fence();
    }
  }
}

// n = 2048, base_n = 1024 on one wavefront (T = 64)
template <bool FWD>
__device__ __forceinline__ void r16l_2048_transform(c64 *va, c64 *vb, c64 *x, uint32_t l,
                                                    const c64 *twl, const c64 *__restrict__ tw) {
  r16l_depth2_transform<10, FWD>(va, vb, x, l, twl, tw);
}

}  // namespace cfft_dev

namespace cfft_dev {

// ---- wave-local 1024-point transform with an unpadded exchange ----
//
// Same operations and layouts as r16w_transform; two differences free LDS
// for the callers that keep more tables there:
//  * the twiddle rows for k = 0 (T[0] = 1, never multiplied) are not stored:
//    tw1[(k-1) 64 + l] = T[k l] and tw2[(k-1) 4 + p] = T[16 k p], k = 1..15;
//  * the exchange buffer holds exactly 1024 points: point i is stored at
//    i ^ ((i >> 4) & 15) (the low four index bits XORed with the next four),
//    which spreads both the row writes (i = 16 l + k) and the column reads
//    (i = l + 64 k) of 16 consecutive lanes over 16 distinct 16-byte bank
//    groups, as a one-point pad per 16 would.
template <bool FWD>
__device__ __forceinline__ void twiddle_rows_x2(c64 *va, c64 *vb, int nv, const c64 *rows,
                                                uint32_t row_stride, uint32_t col) {
#pragma unroll
  for (int g = 0; g < 4; ++g) {
    c64 w[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int k = 4 * g + e;
      if (k > 0) w[e] = tw_at_impl<FWD>(rows, (uint32_t)(k - 1) * row_stride + col);
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int k = 4 * g + e;
      if (k > 0) {
        va[k] = cmul(w[e], va[k]);
        if (nv == 2) vb[k] = cmul(w[e], vb[k]);
      }
    }
    sched_fence();
  }
}

template <bool FWD, int NV>
__device__ __forceinline__ void r16x_transform(c64 *va, c64 *vb, c64 *x, uint32_t l,
                                               const c64 *tw1, const c64 *tw2) {
  using S = R16<10>;
  butterfly<16, FWD>(va);
  if (NV == 2) butterfly<16, FWD>(vb);
  sched_fence();
  twiddle_rows_x2<FWD>(va, vb, NV, tw1, 64, l);
  // exchange 1 (write 16 l + k, read l + 64 k; swizzled as above)
  const uint32_t p = l >> 4, q = l & 15;
#pragma unroll
  for (int t = 0; t < NV; ++t) {
    c64 *v = t == 0 ? va : vb;
    wave_lds_fence();
#pragma unroll
    for (int k = 0; k < 16; ++k) x[16 * l + ((uint32_t)k ^ q)] = v[k];
    wave_lds_fence();
#pragma unroll
    for (int k = 0; k < 16; ++k)
      v[k] = x[64 * (uint32_t)k + 16 * p + (q ^ (p + 4 * ((uint32_t)k & 3)))];
  }
  wave_lds_fence();  // last use of x
  butterfly<16, FWD>(va);
  if (NV == 2) butterfly<16, FWD>(vb);
  sched_fence();
  twiddle_rows_x2<FWD>(va, vb, NV, tw2, 4, p);
  r16_exchange2_rows<10>(va);
  if (NV == 2) r16_exchange2_rows<10>(vb);
#pragma unroll
  for (uint32_t i = 0; i < S::NB; ++i) {
    butterfly<S::REND, FWD>(va + i * S::REND);
    if (NV == 2) butterfly<S::REND, FWD>(vb + i * S::REND);
  }
}

}  // namespace cfft_dev
