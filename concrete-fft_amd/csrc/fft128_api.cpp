// fft128_api.cpp -- C ABI of the 128-bit negacyclic FFT (fft128::Plan,
// src/fft128/mod.rs:1830-1961, and its #[doc(hidden)] free functions).
#include <hip/hip_runtime.h>

#include <cstring>
#include <new>
#include <vector>

#include "cfft.h"
#include "host_stage.hpp"
#include "host_tables.hpp"
#include "plan.hpp"
#include "scratch.hpp"

using cfft_internal::hip_error;
using cfft_internal::set_error;

namespace cfft_launch {
hipError_t fft128_run(bool fwd, int mulv, uint32_t L, double *re0, double *re1, double *im0,
                      double *im1, const double *t_re0, const double *t_re1,
                      const double *t_im0, const double *t_im1, uint64_t batch, hipStream_t s);
hipError_t exact128_polymul(uint32_t L, const uint64_t *a, const uint64_t *b, uint64_t *c,
                            double *margin, uint64_t batch, const double *t, double *S,
                            hipStream_t s);
}

namespace {

bool is_pow2(size_t n) { return n != 0 && (n & (n - 1)) == 0; }
uint32_t ilog2(size_t n) {
  uint32_t l = 0;
  while (((size_t)1 << l) < n) ++l;
  return l;
}

int run128(size_t n, bool fwd, int mulv, double *re0, double *re1, double *im0, double *im1,
           const double *t0, const double *t1, const double *t2, const double *t3, size_t batch,
           hipStream_t s) {
  if (batch == 0) return CFFT_SUCCESS;
  if (!re0 || !re1 || !im0 || !im1 || !t0 || !t1 || !t2 || !t3)
    return set_error(CFFT_ERROR_NULL_POINTER, "null buffer");
  if (!is_pow2(n)) return set_error(CFFT_ERROR_NOT_POWER_OF_TWO, "assertion failed: n.is_power_of_two()");
  const uint32_t L = ilog2(n);
  if (L < 3 || L > 20) return set_error(CFFT_ERROR_SIZE, "fft128 device transforms support 8 <= n <= 2^20");
  return hip_error(cfft_launch::fft128_run(fwd, mulv, L, re0, re1, im0, im1, t0, t1, t2, t3,
                                           batch, s),
                   "fft128");
}

int run128_plan(cfft_plan p, bool fwd, double *re0, double *re1, double *im0, double *im1,
                size_t batch, hipStream_t s) {
  if (!p) return set_error(CFFT_ERROR_NULL_POINTER, "plan is null");
  if (p->kind != cfft_plan_s::FFT128)
    return set_error(CFFT_ERROR_WRONG_PLAN_KIND, "not an fft128 plan");
  const size_t n = p->n;
  const double *t = p->d_tw128;
  return run128(n, fwd, 0, re0, re1, im0, im1, t, t + n, t + 2 * n, t + 3 * n, batch, s);
}

// fft128 Plan::fwd / Plan::inv on host arrays (src/fft128/mod.rs:1905-1960),
// staged through a pinned slot (host_stage.hpp): a chunk of nb transforms
// sits in the slot as [re0 | re1 | im0 | im1], nb * n doubles each
int run128_host(cfft_plan p, bool fwd, double *re0, double *re1, double *im0, double *im1,
                size_t batch) {
  if (!p) return set_error(CFFT_ERROR_NULL_POINTER, "plan is null");
  if (p->kind != cfft_plan_s::FFT128)
    return set_error(CFFT_ERROR_WRONG_PLAN_KIND, "not an fft128 plan");
  if (batch == 0) return CFFT_SUCCESS;
  if (!re0 || !re1 || !im0 || !im1) return set_error(CFFT_ERROR_NULL_POINTER, "null buffer");
  const size_t n = p->n;
  double *h[4] = {re0, re1, im0, im1};
  return cfft_internal::stage_host(
      batch, 4 * n * sizeof(double),
      [&](void *pin, size_t b0, size_t nb) {
        double *d = static_cast<double *>(pin);
        for (int i = 0; i < 4; ++i)
          std::memcpy(d + i * n * nb, h[i] + n * b0, n * nb * sizeof(double));
      },
      [&](void *dv, size_t nb, hipStream_t s) {
        double *d = static_cast<double *>(dv);
        return run128_plan(p, fwd, d, d + n * nb, d + 2 * n * nb, d + 3 * n * nb, nb, s);
      },
      [&](const void *pin, size_t b0, size_t nb) {
        const double *d = static_cast<const double *>(pin);
        for (int i = 0; i < 4; ++i)
          std::memcpy(h[i] + n * b0, d + i * n * nb, n * nb * sizeof(double));
      });
}

}  // namespace

extern "C" {

int cfft128_plan_create(size_t n, cfft_plan *out) {
  if (!out) return set_error(CFFT_ERROR_NULL_POINTER, "out is null");
  *out = nullptr;
  if (!is_pow2(n)) return set_error(CFFT_ERROR_NOT_POWER_OF_TWO, "assertion failed: n.is_power_of_two()");
  if (n < 32) return set_error(CFFT_ERROR_SIZE, "assertion failed: n >= 32");
  if (ilog2(n) > 20) return set_error(CFFT_ERROR_SIZE, "fft128 device plans support n <= 2^20");
  auto *p = new (std::nothrow) cfft_plan_s();
  if (!p) return set_error(CFFT_ERROR_OUT_OF_MEMORY, "plan allocation failed");
  p->kind = cfft_plan_s::FFT128;
  p->n = n;
  p->base_n = n;
  (void)hipGetDevice(&p->device);
  std::vector<double> t(4 * n);
  cfft_host::negacyclic_twiddles128(n, t.data(), t.data() + n, t.data() + 2 * n, t.data() + 3 * n);
  hipError_t e = hipMalloc(&p->d_tw128, 4 * n * sizeof(double));
  if (e == hipSuccess) e = hipMemcpy(p->d_tw128, t.data(), 4 * n * sizeof(double), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    delete p;
    return hip_error(e, "fft128 plan");
  }
  *out = p;
  return CFFT_SUCCESS;
}

int cfft128_fwd(cfft_plan plan, double *re0, double *re1, double *im0, double *im1, size_t batch,
                cfft_stream_t stream) {
  return run128_plan(plan, true, re0, re1, im0, im1, batch, (hipStream_t)stream);
}
int cfft128_inv(cfft_plan plan, double *re0, double *re1, double *im0, double *im1, size_t batch,
                cfft_stream_t stream) {
  return run128_plan(plan, false, re0, re1, im0, im1, batch, (hipStream_t)stream);
}
int cfft128_fwd_host(cfft_plan plan, double *re0, double *re1, double *im0, double *im1,
                     size_t batch) {
  return run128_host(plan, true, re0, re1, im0, im1, batch);
}
int cfft128_inv_host(cfft_plan plan, double *re0, double *re1, double *im0, double *im1,
                     size_t batch) {
  return run128_host(plan, false, re0, re1, im0, im1, batch);
}
int cfft128_negacyclic_fwd_fft(size_t n, double *re0, double *re1, double *im0, double *im1,
                               const double *tw_re0, const double *tw_re1, const double *tw_im0,
                               const double *tw_im1, size_t batch, int mul_variant,
                               cfft_stream_t stream) {
  if (mul_variant != 0 && mul_variant != 1)
    return set_error(CFFT_ERROR_INVALID_ARGUMENT, "mul_variant must be 0 or 1");
  return run128(n, true, mul_variant, re0, re1, im0, im1, tw_re0, tw_re1, tw_im0, tw_im1, batch,
                (hipStream_t)stream);
}
int cfft128_negacyclic_inv_fft(size_t n, double *re0, double *re1, double *im0, double *im1,
                               const double *tw_re0, const double *tw_re1, const double *tw_im0,
                               const double *tw_im1, size_t batch, int mul_variant,
                               cfft_stream_t stream) {
  if (mul_variant != 0 && mul_variant != 1)
    return set_error(CFFT_ERROR_INVALID_ARGUMENT, "mul_variant must be 0 or 1");
  return run128(n, false, mul_variant, re0, re1, im0, im1, tw_re0, tw_re1, tw_im0, tw_im1, batch,
                (hipStream_t)stream);
}
int cfft128_init_negacyclic_twiddles(size_t n, double *re0, double *re1, double *im0,
                                     double *im1) {
  if (!re0 || !re1 || !im0 || !im1) return set_error(CFFT_ERROR_NULL_POINTER, "null argument");
  if (!is_pow2(n)) return set_error(CFFT_ERROR_NOT_POWER_OF_TWO, "assertion failed: n.is_power_of_two()");
  cfft_host::negacyclic_twiddles128(n, re0, re1, im0, im1);
  return CFFT_SUCCESS;
}

int cfft128_exact_polymul_u64(cfft_plan plan, const uint64_t *a, const uint64_t *b, uint64_t *c,
                              double *margin, size_t batch, cfft_stream_t stream) {
  if (!plan) return set_error(CFFT_ERROR_NULL_POINTER, "plan is null");
  if (plan->kind != cfft_plan_s::FFT128)
    return set_error(CFFT_ERROR_WRONG_PLAN_KIND, "not an fft128 plan");
  if (batch == 0) return CFFT_SUCCESS;
  if (!a || !b || !c) return set_error(CFFT_ERROR_NULL_POINTER, "null buffer");
  hipStream_t s = (hipStream_t)stream;
  const size_t n = plan->n;
  cfft_internal::ScratchLease lease;
  hipError_t e = cfft_internal::scratch_acquire(16 * batch * n * sizeof(double), s, &lease);
  if (e != hipSuccess) return hip_error(e, "fft128 product scratch");
  double *S = static_cast<double *>(lease.ptr);
  e = cfft_launch::exact128_polymul(ilog2(n), a, b, c, margin, batch, plan->d_tw128, S, s);
  hipError_t e2 = cfft_internal::scratch_release(lease, s);
  if (e != hipSuccess) return hip_error(e, "fft128 exact product");
  return hip_error(e2, "fft128 product scratch release");
}

}  // extern "C"
