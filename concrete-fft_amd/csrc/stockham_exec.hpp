// stockham_exec.hpp -- runs a pass program (program.hpp) over one transform
// held in LDS.
//
// Each thread owns E points of the transform (E = 16, or n when n < 16) and
// processes E/R butterflies per radix-R pass.  A pass first gathers all of
// the thread's inputs into registers, waits for the whole workgroup, then
// computes and scatters its outputs, so a single LDS buffer serves as both
// Stockham buffers of the reference (x and y).  The first pass gathers
// through a caller-supplied `load` functor (global memory, fused
// conversions) and the last pass scatters through `store`; intermediate
// passes stay in LDS.  LDS addresses carry one 16-byte pad slot per 16
// points to spread power-of-two strides over the banks.
#pragma once

#include "device_math.hpp"
#include "program.hpp"

namespace cfft_dev {

using cfft_prog::Pass;
using cfft_prog::Program;

__device__ __forceinline__ uint32_t lds_pad(uint32_t i) { return i + (i >> 4); }
__host__ __device__ constexpr uint32_t lds_words(uint32_t n) { return n + (n >> 4); }

template <int R>
struct RadixLog {
  static constexpr uint32_t value = R == 2 ? 1 : R == 4 ? 2 : R == 8 ? 3 : 4;
};

// Input/output positions and twiddle exponents of butterfly `b` of pass P.
template <int R>
__device__ __forceinline__ void pass_positions(const Pass &P, uint32_t b, uint32_t *in,
                                               uint32_t *out, uint32_t *twi) {
  constexpr uint32_t lr = RadixLog<R>::value;
  const uint32_t log2m = P.log2len - lr;
  const uint32_t m = 1u << log2m;
  const uint32_t c = b >> log2m, bb = b & (m - 1);
  const uint32_t base = c << P.log2len;
  const uint32_t log2s = P.s;       // stores log2(s)
  const uint32_t log2tw = P.tw_mul; // stores log2(n / len)
  switch (P.kind) {
    case cfft_prog::DEPTH_FWD:
#pragma unroll
      for (int k = 0; k < R; ++k) {
        in[k] = base + ((uint32_t)k << log2m) + bb;
        out[k] = base + ((uint32_t)bitrev_digit<R>(k) << log2m) + bb;
        twi[k] = ((uint32_t)k * bb) << log2tw;
      }
      break;
    case cfft_prog::DEPTH_INV:
#pragma unroll
      for (int k = 0; k < R; ++k) {
        in[k] = base + ((uint32_t)bitrev_digit<R>(k) << log2m) + bb;
        out[k] = base + ((uint32_t)k << log2m) + bb;
        twi[k] = ((uint32_t)k * bb) << log2tw;
      }
      break;
    case cfft_prog::DIF_CORE: {
      const uint32_t p = bb >> log2s, q = bb & ((1u << log2s) - 1);
#pragma unroll
      for (int k = 0; k < R; ++k) {
        in[k] = base + ((uint32_t)k << log2m) + bb;
        out[k] = base + q + (((p << lr) + (uint32_t)k) << log2s);
        twi[k] = ((uint32_t)k * p) << (log2s + log2tw);
      }
    } break;
    case cfft_prog::DIT_CORE: {
      const uint32_t p = bb >> log2s, q = bb & ((1u << log2s) - 1);
#pragma unroll
      for (int k = 0; k < R; ++k) {
        in[k] = base + (p << (lr + log2s)) + ((uint32_t)k << log2s) + q;
        out[k] = base + ((uint32_t)k << log2m) + bb;
        twi[k] = ((uint32_t)k * p) << (log2s + log2tw);
      }
    } break;
    default:  // END: s = len / R
#pragma unroll
      for (int k = 0; k < R; ++k) {
        in[k] = base + bb + ((uint32_t)k << log2m);
        out[k] = in[k];
        twi[k] = 0;
      }
      break;
  }
}

template <int R>
__device__ __forceinline__ void apply_butterfly(const Pass &P, c64 *v, const uint32_t *twi,
                                                const c64 *__restrict__ tw) {
  const bool fwd = P.fwd != 0;
  const bool pre = P.kind == cfft_prog::DIT_CORE || P.kind == cfft_prog::DEPTH_INV;
  const bool post = P.kind == cfft_prog::DIF_CORE || P.kind == cfft_prog::DEPTH_FWD;
  if (pre) {
#pragma unroll
    for (int k = 1; k < R; ++k) {
      c64 w = tw[twi[k]];
      if (!fwd) w.im = -w.im;
      v[k] = cmul(w, v[k]);
    }
  }
  if (fwd) butterfly<R, true>(v);
  else butterfly<R, false>(v);
  if (post) {
#pragma unroll
    for (int k = 1; k < R; ++k) {
      c64 w = tw[twi[k]];
      if (!fwd) w.im = -w.im;
      v[k] = cmul(w, v[k]);
    }
  }
}

template <int E, int R, class Load, class Store>
__device__ __forceinline__ void run_pass(const Pass &P, bool first, bool last, c64 *lds,
                                         uint32_t lt, uint32_t T, const c64 *__restrict__ tw,
                                         Load &load, Store &store) {
  if constexpr (R > E) {
    return;
  } else {
    constexpr int NB = E / R;
    c64 v[NB][R];
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      uint32_t in[R], out[R], twi[R];
      pass_positions<R>(P, lt + T * (uint32_t)i, in, out, twi);
#pragma unroll
      for (int k = 0; k < R; ++k) v[i][k] = first ? load(in[k]) : lds[lds_pad(in[k])];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      uint32_t in[R], out[R], twi[R];
      pass_positions<R>(P, lt + T * (uint32_t)i, in, out, twi);
      apply_butterfly<R>(P, v[i], twi, tw);
#pragma unroll
      for (int k = 0; k < R; ++k) {
        if (last) store(out[k], v[i][k]);
        else lds[lds_pad(out[k])] = v[i][k];
      }
    }
    __syncthreads();
  }
}

// values the compiler cannot see through, so what is formed from them stays
// where it is used (an opaque zero offset keeps a pointer's address space)
__device__ __forceinline__ uint32_t pass_opaque(uint32_t v) {
  asm volatile("" : "+v"(v));
  return v;
}
__device__ __forceinline__ const c64 *pass_opaque(const c64 *p) {
  uint32_t zero = 0;
  asm volatile("" : "+s"(zero));
  return p + zero;
}

template <int E, class Load, class Store>
__device__ __forceinline__ void run_program(const Program &prog, c64 *lds, uint32_t lt, uint32_t T,
                                            const c64 *__restrict__ tw, Load &load, Store &store) {
  for (uint32_t pi = 0; pi < prog.npass; ++pi) {
    const Pass P = prog.p[pi];
    const bool first = pi == 0, last = pi + 1 == prog.npass;
    switch (P.radix) {
      case 2: run_pass<E, 2>(P, first, last, lds, lt, T, tw, load, store); break;
      case 4: run_pass<E, 4>(P, first, last, lds, lt, T, tw, load, store); break;
      case 8: run_pass<E, 8>(P, first, last, lds, lt, T, tw, load, store); break;
      default: run_pass<E, 16>(P, first, last, lds, lt, T, tw, load, store); break;
    }
  }
}

// A program fixed at compile time (FP::prog, a cfft_prog::fixed_program):
// the passes are unrolled and every pass field is a constant, so positions
// and twiddle indices fold and no pass descriptor is read at run time.
template <class FP, int PI, int E, class Load, class Store>
__device__ __forceinline__ void run_fixed_program(c64 *lds, uint32_t lt, uint32_t T,
                                                  const c64 *__restrict__ tw, Load &load,
                                                  Store &store) {
  if constexpr (PI < (int)FP::prog.npass) {
    constexpr Pass P = FP::prog.p[PI];
    // opaque per pass: a pass's addresses and twiddle loads are formed in it,
    // not hoisted into earlier passes (where they would hold registers)
    run_pass<E, P.radix>(P, PI == 0, PI + 1 == (int)FP::prog.npass, lds, pass_opaque(lt), T,
                         pass_opaque(tw), load, store);
    run_fixed_program<FP, PI + 1, E>(lds, lt, T, tw, load, store);
  }
}

}  // namespace cfft_dev
