// noalloc.hip -- cfft_external_product_noalloc: the external product over a
// fixed per-device scratch, for callers that cannot allocate device memory
// (code under HIP graph capture, tight serving loops, harnesses that forbid
// allocation).  Same kernels, operation order and bits as
// cfft_external_product; only the Fourier scratch differs.
#include <hip/hip_runtime.h>

#include <mutex>

#include "cfft.h"
#include "plan.hpp"

namespace {

constexpr size_t kScratchBytes = (size_t)32 << 20;
constexpr int kMaxDevices = 64;

// One copy per device, held until the process ends.  In host code `scratch`
// only names the symbol: its device address is looked up per device below.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
void *g_scratch[kMaxDevices] = {};

// the current device's copy of `scratch`, looked up on the first call on
// that device with the device current
int device_scratch(void **out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return cfft_internal::hip_error(e, "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices)
    return cfft_internal::set_error(CFFT_ERROR_INVALID_ARGUMENT, "device index out of range");
  std::lock_guard<std::mutex> g(g_mu);
  if (!g_scratch[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return cfft_internal::hip_error(e, "hipGetSymbolAddress(scratch)");
    if (!p) return cfft_internal::set_error(CFFT_ERROR_DEVICE, "no device address for the scratch");
    g_scratch[dev] = p;
  }
  *out = g_scratch[dev];
  return CFFT_SUCCESS;
}

}  // namespace

extern "C" {

size_t cfft_noalloc_scratch_bytes(void) { return kScratchBytes; }

int cfft_external_product_noalloc(cfft_plan plan, const void *key_fourier, int k1, int levels,
                                  int base_log, const uint64_t *glwe, uint64_t *out, size_t batch,
                                  cfft_stream_t stream) {
  void *s = nullptr;
  const int rc = device_scratch(&s);
  if (rc != CFFT_SUCCESS) return rc;
  // the split kernels write every spectrum of a pass before reading it, so
  // the scratch needs no clearing between calls
  return cfft_internal::external_product_with_scratch(plan, key_fourier, k1, levels, base_log,
                                                      glwe, out, batch, (hipStream_t)stream, s,
                                                      kScratchBytes, true);
}

}  // extern "C"
