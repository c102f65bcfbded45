// pending.cpp -- entry points whose device implementation is not built yet.
// Each returns CFFT_ERROR_INVALID_ARGUMENT with an explicit message.
#include <hip/hip_runtime.h>

#include "cfft.h"
#include "host_tables.hpp"
#include "plan.hpp"

using cfft_internal::set_error;

extern "C" {

int cfft_exact_polymul_u64(cfft_plan, const uint64_t *, const uint64_t *, uint64_t *, double *,
                           size_t, cfft_stream_t) {
  return set_error(CFFT_ERROR_INVALID_ARGUMENT, "exact product: not built in this library");
}
int cfft_ggsw_to_fourier(cfft_plan, const uint64_t *, int, int, void *, cfft_stream_t) {
  return set_error(CFFT_ERROR_INVALID_ARGUMENT, "external product: not built in this library");
}
int cfft_external_product(cfft_plan, const void *, int, int, int, const uint64_t *, uint64_t *,
                          size_t, cfft_stream_t) {
  return set_error(CFFT_ERROR_INVALID_ARGUMENT, "external product: not built in this library");
}
int cfft128_plan_create(size_t, cfft_plan *) {
  return set_error(CFFT_ERROR_INVALID_ARGUMENT, "fft128: not built in this library");
}
int cfft128_fwd(cfft_plan, double *, double *, double *, double *, size_t, cfft_stream_t) {
  return set_error(CFFT_ERROR_INVALID_ARGUMENT, "fft128: not built in this library");
}
int cfft128_inv(cfft_plan, double *, double *, double *, double *, size_t, cfft_stream_t) {
  return set_error(CFFT_ERROR_INVALID_ARGUMENT, "fft128: not built in this library");
}
int cfft128_fwd_host(cfft_plan, double *, double *, double *, double *, size_t) {
  return set_error(CFFT_ERROR_INVALID_ARGUMENT, "fft128: not built in this library");
}
int cfft128_inv_host(cfft_plan, double *, double *, double *, double *, size_t) {
  return set_error(CFFT_ERROR_INVALID_ARGUMENT, "fft128: not built in this library");
}
int cfft128_negacyclic_fwd_fft(size_t, double *, double *, double *, double *, const double *,
                               const double *, const double *, const double *, size_t, int,
                               cfft_stream_t) {
  return set_error(CFFT_ERROR_INVALID_ARGUMENT, "fft128: not built in this library");
}
int cfft128_negacyclic_inv_fft(size_t, double *, double *, double *, double *, const double *,
                               const double *, const double *, const double *, size_t, int,
                               cfft_stream_t) {
  return set_error(CFFT_ERROR_INVALID_ARGUMENT, "fft128: not built in this library");
}
int cfft128_init_negacyclic_twiddles(size_t n, double *re0, double *re1, double *im0,
                                     double *im1) {
  if (!re0 || !re1 || !im0 || !im1) return set_error(CFFT_ERROR_NULL_POINTER, "null argument");
  cfft_host::negacyclic_twiddles128(n, re0, re1, im0, im1);
  return CFFT_SUCCESS;
}

}  // extern "C"
