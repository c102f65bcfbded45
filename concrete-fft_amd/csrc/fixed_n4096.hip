// fixed_n4096.hip -- compile-time pass programs (fixed_exec.hpp) for n = 4096:
// every algo on bases 1024 and 512 (and, for the transforms, the ordered
// plans' base_n = n) except Dif16 / 1024 (register-resident).
#include "fixed_exec.hpp"

namespace cfft_launch {

bool fixed_polymul_n4096(const cfft_prog::Program &fwd, const cfft_prog::Program &inv,
                       const uint64_t *a, const int64_t *b, uint64_t *c, uint64_t batch,
                       const dc64 *tw, const dc64 *twist, const dc64 *untwist, hipStream_t s,
                       hipError_t *e) {
  const c64 *t = reinterpret_cast<const c64 *>(tw);
  const c64 *tws = reinterpret_cast<const c64 *>(twist);
  const c64 *utw = reinterpret_cast<const c64 *>(untwist);
  return fixed_polymul_try<0, 12, 10>(fwd, inv, a, b, c, batch, t, tws, utw, s, e) ||
         fixed_polymul_try<1, 12, 10>(fwd, inv, a, b, c, batch, t, tws, utw, s, e) ||
         fixed_polymul_try<2, 12, 10>(fwd, inv, a, b, c, batch, t, tws, utw, s, e) ||
         fixed_polymul_try<3, 12, 10>(fwd, inv, a, b, c, batch, t, tws, utw, s, e) ||
         fixed_polymul_try<4, 12, 10>(fwd, inv, a, b, c, batch, t, tws, utw, s, e) ||
         fixed_polymul_try<5, 12, 10>(fwd, inv, a, b, c, batch, t, tws, utw, s, e) ||
         fixed_polymul_try<7, 12, 10>(fwd, inv, a, b, c, batch, t, tws, utw, s, e) ||
         fixed_polymul_try<0, 12, 9>(fwd, inv, a, b, c, batch, t, tws, utw, s, e) ||
         fixed_polymul_try<1, 12, 9>(fwd, inv, a, b, c, batch, t, tws, utw, s, e) ||
         fixed_polymul_try<2, 12, 9>(fwd, inv, a, b, c, batch, t, tws, utw, s, e) ||
         fixed_polymul_try<3, 12, 9>(fwd, inv, a, b, c, batch, t, tws, utw, s, e) ||
         fixed_polymul_try<4, 12, 9>(fwd, inv, a, b, c, batch, t, tws, utw, s, e) ||
         fixed_polymul_try<5, 12, 9>(fwd, inv, a, b, c, batch, t, tws, utw, s, e) ||
         fixed_polymul_try<6, 12, 9>(fwd, inv, a, b, c, batch, t, tws, utw, s, e) ||
         fixed_polymul_try<7, 12, 9>(fwd, inv, a, b, c, batch, t, tws, utw, s, e);
}

bool fixed_fft_n4096(const cfft_prog::Program &prog, dc64 *data, uint64_t ntrans, uint64_t stride,
                   const dc64 *tw, hipStream_t s, hipError_t *e) {
  c64 *d = reinterpret_cast<c64 *>(data);
  const c64 *t = reinterpret_cast<const c64 *>(tw);
  return fixed_fft_try<0, 12, 10>(prog, d, ntrans, stride, t, s, e) ||
         fixed_fft_try<1, 12, 10>(prog, d, ntrans, stride, t, s, e) ||
         fixed_fft_try<2, 12, 10>(prog, d, ntrans, stride, t, s, e) ||
         fixed_fft_try<3, 12, 10>(prog, d, ntrans, stride, t, s, e) ||
         fixed_fft_try<4, 12, 10>(prog, d, ntrans, stride, t, s, e) ||
         fixed_fft_try<5, 12, 10>(prog, d, ntrans, stride, t, s, e) ||
         fixed_fft_try<7, 12, 10>(prog, d, ntrans, stride, t, s, e) ||
         fixed_fft_try<0, 12, 9>(prog, d, ntrans, stride, t, s, e) ||
         fixed_fft_try<1, 12, 9>(prog, d, ntrans, stride, t, s, e) ||
         fixed_fft_try<2, 12, 9>(prog, d, ntrans, stride, t, s, e) ||
         fixed_fft_try<3, 12, 9>(prog, d, ntrans, stride, t, s, e) ||
         fixed_fft_try<4, 12, 9>(prog, d, ntrans, stride, t, s, e) ||
         fixed_fft_try<5, 12, 9>(prog, d, ntrans, stride, t, s, e) ||
         fixed_fft_try<6, 12, 9>(prog, d, ntrans, stride, t, s, e) ||
         fixed_fft_try<7, 12, 9>(prog, d, ntrans, stride, t, s, e) ||
         fixed_fft_try<0, 12, 12>(prog, d, ntrans, stride, t, s, e) ||
         fixed_fft_try<1, 12, 12>(prog, d, ntrans, stride, t, s, e) ||
         fixed_fft_try<2, 12, 12>(prog, d, ntrans, stride, t, s, e) ||
         fixed_fft_try<3, 12, 12>(prog, d, ntrans, stride, t, s, e) ||
         fixed_fft_try<4, 12, 12>(prog, d, ntrans, stride, t, s, e) ||
         fixed_fft_try<5, 12, 12>(prog, d, ntrans, stride, t, s, e) ||
         fixed_fft_try<6, 12, 12>(prog, d, ntrans, stride, t, s, e) ||
         fixed_fft_try<7, 12, 12>(prog, d, ntrans, stride, t, s, e);
}

}  // namespace cfft_launch
