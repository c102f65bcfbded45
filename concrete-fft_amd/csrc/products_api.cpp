// products_api.cpp -- C ABI of the fused products built on an unordered plan:
// the exact limb-split u64 product and the TFHE external product on a
// resident Fourier-domain GGSW key.
#include <hip/hip_runtime.h>

#include <string>

#include "cfft.h"
#include "launch.hpp"
#include "plan.hpp"
#include "scratch.hpp"

using cfft_internal::hip_error;
using cfft_internal::set_error;

namespace cfft_launch {
hipError_t exact_polymul(const cfft_prog::Program &fwd, const cfft_prog::Program &inv,
                         const uint64_t *a, const uint64_t *b, uint64_t *c, double *margin,
                         uint64_t batch, const dc64 *tw, const dc64 *twist, const dc64 *untwist,
                         dc64 *F, hipStream_t s);
hipError_t exact_polymul_w2(const uint64_t *a, const uint64_t *b, uint64_t *c, double *margin,
                            uint64_t batch, const dc64 *tw, const dc64 *tw_blk,
                            const dc64 *twist, const dc64 *untwist, dc64 *F, hipStream_t s);
hipError_t exact_polymul_r16(const uint64_t *a, const uint64_t *b, uint64_t *c, double *margin,
                             uint64_t batch, const dc64 *tw_r16, const dc64 *twist,
                             hipStream_t s);
hipError_t torus_to_fourier(const cfft_prog::Program &fwd, const uint64_t *key, dc64 *out,
                            uint64_t npolys, const dc64 *tw, const dc64 *twist, hipStream_t s);
hipError_t external_product_split(const cfft_prog::Program &fwd, const cfft_prog::Program &inv,
                                  const dc64 *keyf, const uint64_t *glwe, uint64_t *out,
                                  uint64_t batch, uint32_t k1, uint32_t levels,
                                  uint32_t base_log, const dc64 *tw, const dc64 *twist,
                                  const dc64 *untwist, dc64 *scratch, hipStream_t s);
hipError_t external_product_split_w(uint32_t log2n, const dc64 *keyf, const uint64_t *glwe,
                                    const uint64_t *glwe_tab, uint64_t *out,
                                    const uint64_t *out_tab, uint64_t batch, uint32_t k1,
                                    uint32_t levels, uint32_t base_log, const dc64 *tw,
                                    const dc64 *tw_blk, const dc64 *twist, const dc64 *untwist,
                                    dc64 *scratch, hipStream_t s);
}

namespace {
// The external product runs split (rows and columns over workgroups through a
// Fourier scratch) at every batch size: measured faster than one workgroup
// per GLWE from batch 1 to 4096 at N = 1024, 2048 and 4096.  Large batches
// run in passes over one scratch of at most kSplitMaxRows spectra and
// kSplitMaxBytes, a whole number of GLWEs per pass.
constexpr uint64_t kSplitMaxRows = 16384;
constexpr uint64_t kSplitMaxBytes = (uint64_t)256 << 20;
constexpr int kMaxK1 = 64;

uint64_t glwe_per_pass(const cfft_plan_s *plan, int k1, int levels) {
  const uint64_t per_glwe = (uint64_t)k1 * (uint64_t)levels;
  uint64_t rows = kSplitMaxBytes / (plan->n * sizeof(cfft_launch::dc64));
  if (rows > kSplitMaxRows) rows = kSplitMaxRows;
  const uint64_t g = rows / per_glwe;
  return g > 0 ? g : 1;
}

int check_extprod_plan(cfft_plan plan) {
  if (!plan) return set_error(CFFT_ERROR_NULL_POINTER, "plan is null");
  if (plan->kind != cfft_plan_s::UNORDERED)
    return set_error(CFFT_ERROR_WRONG_PLAN_KIND, "the external product runs on an unordered plan");
  if (plan->n < 2 || plan->fwd_full.npass == 0)
    return set_error(CFFT_ERROR_SIZE, "external product supports 4 <= N <= 2 * 8192");
  return CFFT_SUCCESS;
}
uint32_t ilog2(size_t n) {
  uint32_t l = 0;
  while (((size_t)1 << l) < n) ++l;
  return l;
}
}  // namespace

namespace cfft_internal {

size_t external_product_scratch_bytes(const cfft_plan_s *plan, int k1, int levels, size_t batch) {
  const uint64_t per_pass = glwe_per_pass(plan, k1, levels);
  const uint64_t m = (uint64_t)batch < per_pass ? (uint64_t)batch : per_pass;
  return m * (uint64_t)k1 * (uint64_t)levels * plan->n * sizeof(cfft_launch::dc64);
}

// the twiddle block of the wave-local split kernels, or null when the plan
// runs the generic split form
const cfft_launch::dc64 *split_w_block(const cfft_plan_s *plan) {
  return plan->n == 1024 && plan->d_tw_r16        ? plan->d_tw_r16
         : plan->n == 512 && plan->d_tw_base1024  ? plan->d_tw_base1024
         : plan->n == 2048 && plan->d_tw_base1024 ? plan->d_tw_base1024
                                                  : nullptr;
}

bool external_product_tables_supported(const cfft_plan_s *plan) {
  return plan && plan->kind == cfft_plan_s::UNORDERED && split_w_block(plan) != nullptr;
}

int external_product_with_scratch(cfft_plan plan, const void *key_fourier, int k1, int levels,
                                  int base_log, const uint64_t *glwe, uint64_t *out,
                                  size_t batch, hipStream_t stream, void *scratch,
                                  size_t scratch_bytes, bool no_alloc, const uint64_t *glwe_tab,
                                  const uint64_t *out_tab) {
  int st = check_extprod_plan(plan);
  if (st != CFFT_SUCCESS) return st;
  if (k1 < 1 || k1 > kMaxK1 || levels < 1)
    return set_error(CFFT_ERROR_INVALID_ARGUMENT, "k1 must be in [1, 64] and levels >= 1");
  if (base_log < 1 || base_log > 32 || levels * base_log > 64)
    return set_error(CFFT_ERROR_INVALID_ARGUMENT,
                     "gadget base_log must be in [1, 32] with levels * base_log <= 64");
  if (batch == 0) return CFFT_SUCCESS;
  if (!key_fourier || !(glwe || glwe_tab) || !(out || out_tab))
    return set_error(CFFT_ERROR_NULL_POINTER, "null buffer");
  if ((glwe_tab || out_tab) && !external_product_tables_supported(plan))
    return set_error(CFFT_ERROR_WRONG_PLAN_KIND,
                     "address tables need the wave-local split kernels (Dif16, N = 1024 ... 4096)");
  // Rows (digit polynomials) and columns spread over workgroups through a
  // Fourier scratch: measured faster than one workgroup per GLWE from batch 1
  // to 4096 at N = 1024, 2048 and 4096.
  hipError_t e = hipSuccess;
  uint64_t per_pass = glwe_per_pass(plan, k1, levels);
  size_t bytes = external_product_scratch_bytes(plan, k1, levels, batch);
  if (no_alloc) {
    // passes sized to the caller's scratch; nothing is allocated
    const uint64_t per_glwe = (uint64_t)k1 * (uint64_t)levels * plan->n * sizeof(cfft_launch::dc64);
    if (!scratch || scratch_bytes < per_glwe)
      return set_error(CFFT_ERROR_SIZE,
                       "external product: one GLWE's spectra (" + std::to_string(per_glwe) +
                           " bytes) do not fit the " + std::to_string(scratch_bytes) +
                           "-byte scratch");
    if (per_pass > scratch_bytes / per_glwe) per_pass = scratch_bytes / per_glwe;
    bytes = 0;
  }
  hipStream_t s = stream;
  // the caller's scratch when it holds one pass's spectra, else the cache
  cfft_internal::ScratchLease lease;
  if (!no_alloc && (!scratch || scratch_bytes < bytes)) {
    e = cfft_internal::scratch_acquire(bytes, s, &lease);
    if (e != hipSuccess) return hip_error(e, "external product scratch");
  }
  cfft_launch::dc64 *S = reinterpret_cast<cfft_launch::dc64 *>(lease.ptr ? lease.ptr : scratch);
  const size_t glwe_words = (size_t)k1 * 2 * plan->n;  // one GLWE: k1 polynomials of N words
  for (uint64_t off = 0; off < batch && e == hipSuccess; off += per_pass) {
    const uint64_t m = batch - off < per_pass ? batch - off : per_pass;
    const uint64_t *g = glwe ? glwe + off * glwe_words : nullptr;
    uint64_t *o = out ? out + off * glwe_words : nullptr;
    const uint64_t *gt = glwe_tab ? glwe_tab + off : nullptr;
    const uint64_t *ot = out_tab ? out_tab + off : nullptr;
    // Dif16 plans with n = 512, 1024 and 2048 (base 1024): one wave per row
    // and per column (wave-local register-resident transforms; two per wave
    // at n = 512).  Measured at N = 4096 against two waves per row transform
    // in the generic split form: 45 vs 72 us for 16 GLWEs, 155 vs 402 us for
    // 1024, 0.86 vs 1.52 ms for 4096.
    const cfft_launch::dc64 *blk = split_w_block(plan);
    if (blk)
      e = cfft_launch::external_product_split_w(
          ilog2(plan->n), reinterpret_cast<const cfft_launch::dc64 *>(key_fourier), g, gt, o, ot,
          m, (uint32_t)k1, (uint32_t)levels, (uint32_t)base_log, plan->d_tw, blk, plan->d_twist,
          plan->d_untwist, S, s);
    else
      e = cfft_launch::external_product_split(
          plan->fwd_full, plan->inv_full, reinterpret_cast<const cfft_launch::dc64 *>(key_fourier),
          g, o, m, (uint32_t)k1, (uint32_t)levels, (uint32_t)base_log, plan->d_tw, plan->d_twist,
          plan->d_untwist, S, s);
  }
  hipError_t e2 = cfft_internal::scratch_release(lease, s);
  if (e != hipSuccess) return hip_error(e, "external product (split)");
  return hip_error(e2, "external product scratch release");
}

}  // namespace cfft_internal

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
  // Dif16 with base_n == n == 1024: one fused kernel, no scratch
  if (plan->n == 1024 && plan->d_tw_r16)
    return hip_error(cfft_launch::exact_polymul_r16(a, b, c, margin, batch, plan->d_tw_r16,
                                                    plan->d_twist, s),
                     "exact product");
  // the wave-local form keeps the 4 weight spectra of a product, the generic
  // form its 8 limb spectra
  const bool w2 = plan->n == 2048 && plan->d_tw_base1024;  // Dif16, base_n == 1024
  const size_t bytes = batch * (w2 ? 4 : 8) * plan->n * sizeof(cfft_launch::dc64);
  cfft_internal::ScratchLease lease;
  hipError_t e = cfft_internal::scratch_acquire(bytes, s, &lease);
  if (e != hipSuccess) return hip_error(e, "exact product scratch");
  void *F = lease.ptr;
  if (w2)
    e = cfft_launch::exact_polymul_w2(a, b, c, margin, batch, plan->d_tw, plan->d_tw_base1024,
                                      plan->d_twist, plan->d_untwist,
                                      reinterpret_cast<cfft_launch::dc64 *>(F), s);
  else
    e = cfft_launch::exact_polymul(plan->fwd_full, plan->inv_full, a, b, c, margin, batch,
                                   plan->d_tw, plan->d_twist, plan->d_untwist,
                                   reinterpret_cast<cfft_launch::dc64 *>(F), s);
  hipError_t e2 = cfft_internal::scratch_release(lease, s);
  if (e != hipSuccess) return hip_error(e, "exact product");
  return hip_error(e2, "exact product scratch release");
}

int cfft_ggsw_to_fourier(cfft_plan plan, const uint64_t *key, int k1, int levels, void *key_fourier,
                         cfft_stream_t stream) {
  int st = check_extprod_plan(plan);
  if (st != CFFT_SUCCESS) return st;
  if (k1 < 1 || levels < 1) return set_error(CFFT_ERROR_INVALID_ARGUMENT, "k1 and levels must be >= 1");
  if (!key || !key_fourier) return set_error(CFFT_ERROR_NULL_POINTER, "null buffer");
  const uint64_t npolys = (uint64_t)k1 * (uint64_t)levels * (uint64_t)k1;
  return hip_error(cfft_launch::torus_to_fourier(plan->fwd_full, key,
                                                 reinterpret_cast<cfft_launch::dc64 *>(key_fourier),
                                                 npolys, plan->d_tw, plan->d_twist,
                                                 (hipStream_t)stream),
                   "ggsw to fourier");
}

int cfft_external_product(cfft_plan plan, const void *key_fourier, int k1, int levels,
                          int base_log, const uint64_t *glwe, uint64_t *out, size_t batch,
                          cfft_stream_t stream) {
  return cfft_internal::external_product_with_scratch(plan, key_fourier, k1, levels, base_log,
                                                      glwe, out, batch, (hipStream_t)stream,
                                                      nullptr, 0, false);
}

}  // extern "C"
