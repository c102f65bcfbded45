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

// ---- building a plan's pass lists ----
// A plan's whole-transform programs: the depth levels above base_n (radix 2
// when n = 2 base_n, 4 when n = 4 base_n, else 8: src/unordered.rs:361-367,
// :407-413; forward top level first, inverse in reverse) around the base
// recursion of src/dif{2,4,8,16}.rs / src/dit{2,4,8,16}.rs as a pass list
// (core passes while more than log2(R) levels remain, then the end butterfly
// of radix 2^L; DIT runs the end first and the cores in reverse).  constexpr,
// so the plans are built with it at run time and kernels specialised to one
// (algo, n, base_n) use it at compile time (every pass field a constant);
// their launchers still compare the two before such a launch.  algo: 0 Dif2,
// 1 Dit2, 2 Dif4, ..., 7 Dit16.
constexpr uint32_t c_log2(uint32_t n) {
  uint32_t l = 0;
  while ((1u << l) < n) ++l;
  return l;
}

constexpr Pass c_pass(uint8_t kind, int radix, bool fwd, uint32_t log2len, uint32_t log2s,
                      uint32_t log2tw) {
  Pass p{};
  p.kind = kind;
  p.radix = (uint8_t)radix;
  p.fwd = fwd ? 1 : 0;
  p.log2len = (uint8_t)log2len;
  p.s = log2s;
  p.tw_mul = log2tw;
  return p;
}

constexpr Program fixed_program(int algo, uint32_t log2n, uint32_t log2bn, bool fwd) {
  Program P{};
  P.log2n = log2n;
  // a list longer than kMaxPasses is cut there (npass = kMaxPasses + 1
  // marks it); plans up to n = 2^30 with base_n <= 1024 take at most 20
  bool over = false;
  auto add = [&](Pass q) {
    if (P.npass < (uint32_t)kMaxPasses) P.p[P.npass++] = q;
    else over = true;
  };
  const int R = 2 << (algo / 2);
  const uint32_t lr = c_log2((uint32_t)R);
  const bool dit = (algo & 1) != 0;
  int radix[16] = {};
  uint32_t lvl[16] = {};
  int nlev = 0;
  for (uint32_t nl = log2n; nl > log2bn; ++nlev) {
    if (nlev == 16) {  // beyond any supported size
      over = true;
      break;
    }
    radix[nlev] = nl == log2bn + 1 ? 2 : nl == log2bn + 2 ? 4 : 8;
    lvl[nlev] = nl;
    nl -= c_log2((uint32_t)radix[nlev]);
  }
  if (fwd)
    for (int i = 0; i < nlev; ++i)
      add(c_pass(DEPTH_FWD, radix[i], true, lvl[i], 0, log2n - lvl[i]));
  const uint32_t tw = log2n - log2bn;
  if (log2bn > 0) {
    uint32_t cores[16] = {};
    int nc = 0;
    uint32_t L = log2bn, st = 0;
    while (L > lr && nc < 16) {
      cores[nc++] = st;
      st += lr;
      L -= lr;
    }
    if (L > lr) over = true;  // beyond any supported size
    const int rend = 1 << L;
    if (!dit) {
      for (int i = 0; i < nc; ++i) add(c_pass(DIF_CORE, R, fwd, log2bn, cores[i], tw));
      add(c_pass(END, rend, fwd, log2bn, 0, tw));
    } else {
      add(c_pass(END, rend, fwd, log2bn, 0, tw));
      for (int i = nc - 1; i >= 0; --i)
        add(c_pass(DIT_CORE, R, fwd, log2bn, cores[i], tw));
    }
  }
  if (!fwd)
    for (int i = nlev - 1; i >= 0; --i)
      add(c_pass(DEPTH_INV, radix[i], false, lvl[i], 0, log2n - lvl[i]));
  if (over) P.npass = kMaxPasses + 1;
  return P;
}

// same passes (the fields the executor reads)
inline bool same_passes(const Program &a, const Program &b) {
  if (a.npass != b.npass) return false;
  for (uint32_t i = 0; i < a.npass; ++i) {
    const Pass &x = a.p[i], &y = b.p[i];
    if (x.kind != y.kind || x.radix != y.radix || x.fwd != y.fwd || x.log2len != y.log2len ||
        x.s != y.s || x.tw_mul != y.tw_mul)
      return false;
  }
  return true;
}

}  // namespace cfft_prog
