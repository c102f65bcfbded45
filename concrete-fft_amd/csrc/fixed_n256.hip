// fixed_n256.hip -- compile-time pass programs (fixed_exec.hpp) for n = 256:
// every algo on base 256 (the autotuner's base for n = 256) except Dif16
// (register-resident).
#include "fixed_exec.hpp"

namespace cfft_launch {

bool fixed_polymul_n256(const cfft_prog::Program &fwd, const cfft_prog::Program &inv,
                       const uint64_t *a, const int64_t *b, uint64_t *c, uint64_t batch,
                       const dc64 *tw, const dc64 *twist, const dc64 *untwist, hipStream_t s,
                       hipError_t *e) {
  const c64 *t = reinterpret_cast<const c64 *>(tw);
  const c64 *tws = reinterpret_cast<const c64 *>(twist);
  const c64 *utw = reinterpret_cast<const c64 *>(untwist);
  return fixed_polymul_try<0, 8, 8>(fwd, inv, a, b, c, batch, t, tws, utw, s, e) ||
         fixed_polymul_try<1, 8, 8>(fwd, inv, a, b, c, batch, t, tws, utw, s, e) ||
         fixed_polymul_try<2, 8, 8>(fwd, inv, a, b, c, batch, t, tws, utw, s, e) ||
         fixed_polymul_try<3, 8, 8>(fwd, inv, a, b, c, batch, t, tws, utw, s, e) ||
         fixed_polymul_try<4, 8, 8>(fwd, inv, a, b, c, batch, t, tws, utw, s, e) ||
         fixed_polymul_try<5, 8, 8>(fwd, inv, a, b, c, batch, t, tws, utw, s, e) ||
         fixed_polymul_try<7, 8, 8>(fwd, inv, a, b, c, batch, t, tws, utw, s, e);
}

bool fixed_fft_n256(const cfft_prog::Program &prog, dc64 *data, uint64_t ntrans, uint64_t stride,
                   const dc64 *tw, hipStream_t s, hipError_t *e) {
  c64 *d = reinterpret_cast<c64 *>(data);
  const c64 *t = reinterpret_cast<const c64 *>(tw);
  return fixed_fft_try<0, 8, 8>(prog, d, ntrans, stride, t, s, e) ||
         fixed_fft_try<1, 8, 8>(prog, d, ntrans, stride, t, s, e) ||
         fixed_fft_try<2, 8, 8>(prog, d, ntrans, stride, t, s, e) ||
         fixed_fft_try<3, 8, 8>(prog, d, ntrans, stride, t, s, e) ||
         fixed_fft_try<4, 8, 8>(prog, d, ntrans, stride, t, s, e) ||
         fixed_fft_try<5, 8, 8>(prog, d, ntrans, stride, t, s, e) ||
         fixed_fft_try<7, 8, 8>(prog, d, ntrans, stride, t, s, e);
}

}  // namespace cfft_launch
