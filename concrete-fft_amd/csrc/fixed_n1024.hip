// fixed_n1024.hip -- compile-time pass programs (fixed_exec.hpp) for n = 1024:
// every algo on bases 1024 and 512 except those with a register-resident kernel
// (Dif16 / 1024, Dit16 / 1024, Dif8 / 1024, Dif16 / 512).
#include "fixed_exec.hpp"

namespace cfft_launch {

bool fixed_polymul_n1024(const cfft_prog::Program &fwd, const cfft_prog::Program &inv,
                       const uint64_t *a, const int64_t *b, uint64_t *c, uint64_t batch,
                       const dc64 *tw, const dc64 *twist, const dc64 *untwist, hipStream_t s,
                       hipError_t *e) {
  const c64 *t = reinterpret_cast<const c64 *>(tw);
  const c64 *tws = reinterpret_cast<const c64 *>(twist);
  const c64 *utw = reinterpret_cast<const c64 *>(untwist);
  return fixed_polymul_try<0, 10, 10>(fwd, inv, a, b, c, batch, t, tws, utw, s, e) ||
         fixed_polymul_try<1, 10, 10>(fwd, inv, a, b, c, batch, t, tws, utw, s, e) ||
         fixed_polymul_try<2, 10, 10>(fwd, inv, a, b, c, batch, t, tws, utw, s, e) ||
         fixed_polymul_try<3, 10, 10>(fwd, inv, a, b, c, batch, t, tws, utw, s, e) ||
         fixed_polymul_try<5, 10, 10>(fwd, inv, a, b, c, batch, t, tws, utw, s, e) ||
         fixed_polymul_try<0, 10, 9>(fwd, inv, a, b, c, batch, t, tws, utw, s, e) ||
         fixed_polymul_try<1, 10, 9>(fwd, inv, a, b, c, batch, t, tws, utw, s, e) ||
         fixed_polymul_try<2, 10, 9>(fwd, inv, a, b, c, batch, t, tws, utw, s, e) ||
         fixed_polymul_try<3, 10, 9>(fwd, inv, a, b, c, batch, t, tws, utw, s, e) ||
         fixed_polymul_try<4, 10, 9>(fwd, inv, a, b, c, batch, t, tws, utw, s, e) ||
         fixed_polymul_try<5, 10, 9>(fwd, inv, a, b, c, batch, t, tws, utw, s, e) ||
         fixed_polymul_try<7, 10, 9>(fwd, inv, a, b, c, batch, t, tws, utw, s, e);
}

bool fixed_fft_n1024(const cfft_prog::Program &prog, dc64 *data, uint64_t ntrans, uint64_t stride,
                   const dc64 *tw, hipStream_t s, hipError_t *e) {
  c64 *d = reinterpret_cast<c64 *>(data);
  const c64 *t = reinterpret_cast<const c64 *>(tw);
  return fixed_fft_try<0, 10, 10>(prog, d, ntrans, stride, t, s, e) ||
         fixed_fft_try<1, 10, 10>(prog, d, ntrans, stride, t, s, e) ||
         fixed_fft_try<2, 10, 10>(prog, d, ntrans, stride, t, s, e) ||
         fixed_fft_try<3, 10, 10>(prog, d, ntrans, stride, t, s, e) ||
         fixed_fft_try<5, 10, 10>(prog, d, ntrans, stride, t, s, e) ||
         fixed_fft_try<0, 10, 9>(prog, d, ntrans, stride, t, s, e) ||
         fixed_fft_try<1, 10, 9>(prog, d, ntrans, stride, t, s, e) ||
         fixed_fft_try<2, 10, 9>(prog, d, ntrans, stride, t, s, e) ||
         fixed_fft_try<3, 10, 9>(prog, d, ntrans, stride, t, s, e) ||
         fixed_fft_try<4, 10, 9>(prog, d, ntrans, stride, t, s, e) ||
         fixed_fft_try<5, 10, 9>(prog, d, ntrans, stride, t, s, e) ||
         fixed_fft_try<7, 10, 9>(prog, d, ntrans, stride, t, s, e);
}

}  // namespace cfft_launch
