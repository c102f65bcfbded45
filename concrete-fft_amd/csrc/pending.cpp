// pending.cpp -- entry points whose device implementation is not built yet.
// Each returns CFFT_ERROR_INVALID_ARGUMENT with an explicit message.
#include <hip/hip_runtime.h>

#include "cfft.h"
#include "host_tables.hpp"
#include "plan.hpp"

using cfft_internal::set_error;

extern "C" {

int cfft_ggsw_to_fourier(cfft_plan, const uint64_t *, int, int, void *, cfft_stream_t) {
  return set_error(CFFT_ERROR_INVALID_ARGUMENT, "external product: not built in this library");
}
int cfft_external_product(cfft_plan, const void *, int, int, int, const uint64_t *, uint64_t *,
                          size_t, cfft_stream_t) {
  return set_error(CFFT_ERROR_INVALID_ARGUMENT, "external product: not built in this library");
}
}  // extern "C"
