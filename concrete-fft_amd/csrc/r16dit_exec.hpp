// r16dit_exec.hpp -- register-resident schedule of the reference's Dit16
// plan for n = 1024 (base_n = n) on one wavefront.
//
// The Dit16 recursion (src/dit16.rs) runs, for n = 1024, the radix-4 end
// butterfly first, then the radix-16 DIT cores at stride 16 and stride 1,
// each core multiplying its inputs by the twiddles before the butterfly
// (the DIT_CORE / END passes of the pass program, with their positions and
// twiddle exponents).  With 64 lanes, lane l = 16 p + q:
//   end        lane l, group i < 4: points l + 64 i + 256 k (k < 4), in place,
//              held in slots i + 4 k of the lane-linear layout (slot m =
//              point l + 64 m);
//   core s=16  lane l reads points 256 p + 16 k + q, twiddles T[16 k p],
//              writes 64 k + l;
//   core s=1   lane l reads points 16 l + k, twiddles T[k l], writes 64 k + l,
// so the input and the output are lane-linear (slot m = point l + 64 m).
// Between the end and the first core the slots are renamed (slot i + 4 k ->
// 4 i + k) and the 16-lane rows transposed against the low slot bits in
// registers (the same row swaps as the Dif16 schedule's second exchange);
// between the two cores the points go through the wave's LDS buffer (1024
// points, XOR-swizzled as the Dif16 exchange).  Every butterfly and twiddle
// product is the pass program's, so results are bit-identical to it.
#pragma once

#include "r16_exec.hpp"

namespace cfft_dev {

// NV = 1 or 2 transforms sharing every twiddle read (va, vb).  tw1 / tw2:
// the twiddle rows k = 1..15 (tw1[(k-1) 64 + l] = T[k l], tw2[(k-1) 4 + p] =
// T[16 k p]); x: the wave's 1024-point exchange buffer.
template <bool FWD, int NV>
__device__ __forceinline__ void r16dit_transform(c64 *va, c64 *vb, c64 *x, uint32_t l,
                                                 const c64 *tw1, const c64 *tw2) {
  const uint32_t p = l >> 4, q = l & 15;
  // end: radix 4 over slots i, i + 4, i + 8, i + 12; then slot i + 4 k -> 4 i + k
  c64 ua[16], ub[16];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    c64 t[4] = {va[i], va[i + 4], va[i + 8], va[i + 12]};
    butterfly<4, FWD>(t);
#pragma unroll
    for (int k = 0; k < 4; ++k) ua[4 * i + k] = t[k];
    if (NV == 2) {
      c64 s[4] = {vb[i], vb[i + 4], vb[i + 8], vb[i + 12]};
      butterfly<4, FWD>(s);
#pragma unroll
      for (int k = 0; k < 4; ++k) ub[4 * i + k] = s[k];
    }
  }
  // rows against the low slot bits: slot k now holds point 256 p + 16 k + q
  r16_exchange2_rows<10>(ua);
  if (NV == 2) r16_exchange2_rows<10>(ub);
  sched_fence();
  // core, stride 16
  twiddle_rows_x2<FWD>(ua, ub, NV, tw2, 4, p);
  butterfly<16, FWD>(ua);
  if (NV == 2) butterfly<16, FWD>(ub);
  // exchange: write 64 k + l, read 16 l + k (swizzled)
#pragma unroll
  for (int t = 0; t < NV; ++t) {
    c64 *u = t == 0 ? ua : ub;
    wave_lds_fence();
#pragma unroll
    for (int k = 0; k < 16; ++k) x[64 * (uint32_t)k + 16 * p + (q ^ (p + 4 * ((uint32_t)k & 3)))] = u[k];
    wave_lds_fence();
#pragma unroll
    for (int k = 0; k < 16; ++k) u[k] = x[16 * l + ((uint32_t)k ^ q)];
  }
  wave_lds_fence();  // last use of x
  // core, stride 1
  twiddle_rows_x2<FWD>(ua, ub, NV, tw1, 64, l);
  butterfly<16, FWD>(ua);
  if (NV == 2) butterfly<16, FWD>(ub);
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    va[k] = ua[k];
    if (NV == 2) vb[k] = ub[k];
  }
}

}  // namespace cfft_dev
