// host_stage.hpp -- staging for the entry points that take host buffers
// (cfft_fwd_host, cfft_inv_host, cfft_fwd_monomial_host, cfft128_*_host).
//
// A call borrows a slot: a private non-blocking stream and a pinned host
// buffer that the device can address directly (hipHostMalloc, mapped).
// Slots are kept per device and reused, so a call allocates nothing once a
// slot of its size exists.  The caller's data is copied into the pinned
// buffer with memcpy, then either
//   * small calls (at most kZeroCopyBytes): the kernels run on the pinned
//     buffer itself, reading and writing host memory over the bus -- one
//     launch and one stream synchronisation, no DMA copy; or
//   * larger calls: the pinned buffer is copied to a device buffer from the
//     stream-ordered scratch cache, the kernels run there, and the result is
//     copied back; batches above kStageBytes run in chunks.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <string>

#include "cfft.h"
#include "scratch.hpp"

namespace cfft_internal {

constexpr size_t kZeroCopyBytes = (size_t)64 << 10;
constexpr size_t kStageBytes = (size_t)64 << 20;

struct HostSlot {
  int device = -1;
  hipStream_t stream = nullptr;
  void *pinned = nullptr;  // host address
  void *mapped = nullptr;  // the same memory's device address
  size_t bytes = 0;
  bool busy = false;
};

// a slot of the current device with at least `bytes` of pinned memory
hipError_t host_slot_acquire(size_t bytes, HostSlot **out);
void host_slot_release(HostSlot *slot);

int hip_error(hipError_t e, const char *where);
int set_error(int status, const std::string &msg);

// Runs a host-buffer call of `batch` items of `per` bytes each, staged as
// above.  copy_in(pinned, b0, nb) fills the pinned buffer with items
// [b0, b0 + nb) (a no-op when the kernels only write), run(dev, nb, stream)
// issues the kernels on a device buffer of nb items, copy_out(pinned, b0, nb)
// takes the results.  Returns the first failure; the slot's stream is
// synchronised before the slot is given back, whatever happened.
template <class CopyIn, class Run, class CopyOut>
int stage_host(size_t batch, size_t per, CopyIn copy_in, Run run, CopyOut copy_out) {
  if (batch == 0) return 0;
  if (per == 0 || batch > ((size_t)-1 >> 1) / per)
    return set_error(CFFT_ERROR_SIZE_OVERFLOW, "host buffer size overflows");
  const size_t total = batch * per;
  const bool zero_copy = total <= kZeroCopyBytes;
  const size_t chunk = zero_copy ? batch : std::min(batch, std::max<size_t>(1, kStageBytes / per));
  HostSlot *slot = nullptr;
  hipError_t e = host_slot_acquire(chunk * per, &slot);
  if (e != hipSuccess) return hip_error(e, "host staging");
  hipStream_t s = slot->stream;
  ScratchLease lease;
  int rc = 0;
  if (!zero_copy) rc = hip_error(scratch_acquire(chunk * per, s, &lease), "device scratch");
  for (size_t b0 = 0; b0 < batch && rc == 0; b0 += chunk) {
    const size_t nb = std::min(chunk, batch - b0);
    copy_in(slot->pinned, b0, nb);
    if (zero_copy) {
      rc = run(slot->mapped, nb, s);
    } else {
      rc = hip_error(hipMemcpyAsync(lease.ptr, slot->pinned, nb * per, hipMemcpyHostToDevice, s),
                     "hipMemcpyAsync");
      if (rc == 0) rc = run(lease.ptr, nb, s);
      if (rc == 0)
        rc = hip_error(hipMemcpyAsync(slot->pinned, lease.ptr, nb * per, hipMemcpyDeviceToHost, s),
                       "hipMemcpyAsync");
    }
    const hipError_t es = hipStreamSynchronize(s);
    if (rc == 0) rc = hip_error(es, "hipStreamSynchronize");
    if (rc == 0) copy_out(slot->pinned, b0, nb);
  }
  if (lease.ptr) {
    const int rr = hip_error(scratch_release(lease, s), "device scratch");
    if (rc == 0) rc = rr;
  }
  (void)hipStreamSynchronize(s);
  host_slot_release(slot);
  return rc;
}

}  // namespace cfft_internal
