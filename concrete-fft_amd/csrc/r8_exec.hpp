// r8_exec.hpp -- register-resident schedule of the reference's Dif8 plan for
// n = 1024 points (base_n == n), one transform per 128-thread workgroup.
//
// The reference's Dif8 recursion for n = 1024 (src/dif8.rs:393-455) runs
// three radix-8 Stockham cores at strides 1, 8 and 64 and a radix-2 end at
// stride 512 (stockham_dif2_end).  With T = 128 threads and butterfly index
// b = l, each core reads the points l + 128 k (k < 8) and the end pairs the
// points l + 128 i and l + 128 (i + 4): every thread's inputs and outputs of a
// whole transform are the points l + 128 m, m < 8 ("lane-linear", slot m),
// and three LDS exchanges separate the four passes.  Butterflies and twiddle
// products are the reference's operation DAG (twiddle after the butterfly,
// index k * p * s into T[j] = sincospi64(-2 j / n), conjugated for the
// inverse), so results are bit-identical to the generic schedule.
//
// Per thread: 8 points of a transform (32 VGPRs), so a product's two operands
// fit in well under 128 VGPRs and four waves run per SIMD; the 17 KiB LDS
// exchange buffer is shared by the workgroup's two waves.
#pragma once

#include "device_math.hpp"
#include "r16_exec.hpp"

namespace cfft_dev {

struct R8 {
  static constexpr uint32_t n = 1024;
  static constexpr uint32_t T = 128;
  static constexpr uint32_t lds_words = n + n / 16;  // r16_pad padding
  // twiddle block: [k][l] T[k l] (8 x 128), [k][p] T[8 k p] (8 x 16),
  // [k][p] T[64 k p] (8 x 2)
  static constexpr uint32_t tw_words = 8 * 128 + 8 * 16 + 8 * 2;
};

template <bool FWD>
__device__ __forceinline__ void twiddle8(c64 *v, const c64 *__restrict__ tw_row0,
                                         uint32_t row_stride, uint32_t col) {
  c64 w[7];
#pragma unroll
  for (int k = 1; k < 8; ++k) w[k - 1] = tw_at_impl<FWD>(tw_row0, (uint32_t)k * row_stride + col);
#pragma unroll
  for (int k = 1; k < 8; ++k) v[k] = cmul(w[k - 1], v[k]);
}

// One transform; v[m] = point l + 128 m on entry and on exit.  `x` holds
// R8::lds_words points shared by the workgroup; begins with LDS writes and
// ends after its last LDS read and a barrier.
template <bool FWD>
__device__ __forceinline__ void r8_transform(c64 *v, c64 *__restrict__ x, uint32_t l,
                                             const c64 *__restrict__ tw) {
  tw = launder(tw);
  // core, stride 1: outputs 8 l + k, twiddle k l
  butterfly<8, FWD>(v);
  twiddle8<FWD>(v, tw, 128, l);
#pragma unroll
  for (int k = 0; k < 8; ++k) x[r16_pad(8 * l + (uint32_t)k)] = v[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 8; ++k) v[k] = x[r16_pad(l + 128 * (uint32_t)k)];
  __syncthreads();
  // core, stride 8: p = l / 8, q = l % 8, outputs q + 8 (8 p + k), twiddle 8 k p
  {
    const uint32_t p = l >> 3, q = l & 7;
    butterfly<8, FWD>(v);
    twiddle8<FWD>(v, tw + 8 * 128, 16, p);
#pragma unroll
    for (int k = 0; k < 8; ++k) x[r16_pad(q + 64 * p + 8 * (uint32_t)k)] = v[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 8; ++k) v[k] = x[r16_pad(l + 128 * (uint32_t)k)];
  __syncthreads();
  // core, stride 64: p = l / 64, q = l % 64, outputs q + 64 (8 p + k), twiddle 64 k p
  {
    const uint32_t p = l >> 6, q = l & 63;
    butterfly<8, FWD>(v);
    twiddle8<FWD>(v, tw + 8 * 128 + 8 * 16, 2, p);
#pragma unroll
    for (int k = 0; k < 8; ++k) x[r16_pad(q + 512 * p + 64 * (uint32_t)k)] = v[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 8; ++k) v[k] = x[r16_pad(l + 128 * (uint32_t)k)];
  __syncthreads();
  // end, radix 2 at stride 512: points l + 128 i and l + 128 (i + 4)
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    c64 t[2] = {v[i], v[i + 4]};
    butterfly<2, FWD>(t);
    v[i] = t[0];
    v[i + 4] = t[1];
  }
}

}  // namespace cfft_dev
