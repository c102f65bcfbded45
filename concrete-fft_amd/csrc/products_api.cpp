// products_api.cpp -- C ABI of the fused products built on an unordered plan:
// the exact limb-split u64 product.
#include <hip/hip_runtime.h>

#include "cfft.h"
#include "launch.hpp"
#include "plan.hpp"

using cfft_internal::hip_error;
using cfft_internal::set_error;

namespace cfft_launch {
hipError_t exact_polymul(const cfft_prog::Program &fwd, const cfft_prog::Program &inv,
                         const uint64_t *a, const uint64_t *b, uint64_t *c, double *margin,
                         uint64_t batch, const dc64 *tw, const dc64 *twist, const dc64 *untwist,
                         dc64 *F, hipStream_t s);
}

extern "C" {

int cfft_exact_polymul_u64(cfft_plan plan, const uint64_t *a, const uint64_t *b, uint64_t *c,
                           double *margin, size_t batch, cfft_stream_t stream) {
  if (!plan) return set_error(CFFT_ERROR_NULL_POINTER, "plan is null");
  if (plan->kind != cfft_plan_s::UNORDERED)
    return set_error(CFFT_ERROR_WRONG_PLAN_KIND, "the exact product runs on an unordered plan");
  if (batch == 0) return CFFT_SUCCESS;
  if (!a || !b || !c) return set_error(CFFT_ERROR_NULL_POINTER, "null buffer");
  // |D_k| <= 4 N 2^30 must stay well inside the f64 mantissa
  if (plan->n < 2 || plan->n > 2048 || plan->fwd_full.npass == 0)
    return set_error(CFFT_ERROR_SIZE, "exact product supports 4 <= N <= 4096 (fft_size <= 2048)");
  hipStream_t s = (hipStream_t)stream;
  const size_t bytes = batch * 8 * plan->n * sizeof(cfft_launch::dc64);
  void *F = nullptr;
  hipError_t e = hipMallocAsync(&F, bytes, s);
  if (e != hipSuccess) return hip_error(e, "hipMallocAsync (exact product scratch)");
  e = cfft_launch::exact_polymul(plan->fwd_full, plan->inv_full, a, b, c, margin, batch,
                                 plan->d_tw, plan->d_twist, plan->d_untwist,
                                 reinterpret_cast<cfft_launch::dc64 *>(F), s);
  hipError_t e2 = hipFreeAsync(F, s);
  if (e != hipSuccess) return hip_error(e, "exact product");
  return hip_error(e2, "hipFreeAsync");
}

}  // extern "C"
