// program.hpp -- a plan compiled into a flat list of radix passes.
//
// The reference composes its transforms recursively (src/unordered.rs:
// 391-489 fwd_depth / inv_depth over src/dif*.rs, src/dit*.rs Stockham
// recursions).  On the device the same arithmetic is scheduled as a flat
// program of passes over one transform held in LDS; every pass is described
// by where each butterfly reads its R inputs, which twiddle index it uses and
// where it writes its R outputs.  `b` is the butterfly index in [0, len/R)
// within a chunk of `len` points; twiddles are entries of the single table
// T[j] = sincospi64(-2j/n) of the top-level size n (conjugated for the
// inverse), addressed as T[(k * p * s) * tw_mul].
//
//   DEPTH_FWD  (src/unordered.rs:222-277, fwd_process_x{2,4,8})
//              len = level size n_l, m = n_l/R, j = b:
//              in_k = k*m + j, out = bitrev(k)*m + j, twiddle after (k*j)
//   DEPTH_INV  (src/unordered.rs:233-293, inv_process_x{2,4,8})
//              in_k = bitrev(k)*m + j, out_k = k*m + j, twiddle before
//   DIF_CORE   (stockham_core_generic of src/dif{2,4,8,16}.rs), len = base_n
//              p = b / s, q = b % s: in_k = k*len/R + b,
//              out_k = q + s*(R*p + k), twiddle after (k*p*s)
//   DIT_CORE   (stockham_core_generic of src/dit{2,4,8,16}.rs)
//              in_k = R*s*p + k*s + q, out_k = k*len/R + b, twiddle before
//   END        (stockham_{dif,dit}{2,4,8,16}_end): s = len/R,
//              in_k = out_k = b + k*s, no twiddle
#pragma once

#include <stdint.h>

namespace cfft_prog {

enum PassKind : uint8_t { DEPTH_FWD = 0, DEPTH_INV = 1, DIF_CORE = 2, DIT_CORE = 3, END = 4 };

struct Pass {
  uint8_t kind;
  uint8_t radix;   // 2, 4, 8 or 16
  uint8_t fwd;     // butterfly direction (mul_j sign) and twiddle conjugation
  uint8_t log2len; // log2 of the chunk length the pass works on
  uint32_t s;      // Stockham stride (DIF_CORE / DIT_CORE)
  uint32_t tw_mul; // twiddle index multiplier n / len
  uint32_t pad;
};

constexpr int kMaxPasses = 24;

struct Program {
  uint32_t npass;
  uint32_t log2n;  // log2 of the transform (chunk) the program runs over
  Pass p[kMaxPasses];
};

}  // namespace cfft_prog
