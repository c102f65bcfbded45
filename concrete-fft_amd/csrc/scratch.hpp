// scratch.hpp -- stream-ordered device scratch for the product entry points.
//
// A per-device cache of device buffers replaces a hipMallocAsync /
// hipFreeAsync pair per call (measured 17-28 us of device time per pair on
// MI355X, a third of a single external product).  An entry is handed to a
// call on stream s when its last use was on s (stream order covers it) or
// its completion event has fired; otherwise another entry is created (a
// power of two of at least 32 MiB; up to 8 per device and 2 GiB in all), or,
// as a last resort, s waits for the entry's event.  Requests above the cache's limits fall back
// to hipMallocAsync / hipFreeAsync, and so do calls on a capturing stream
// (a graph allocation node).  The cache keeps its buffers until the process
// ends.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace cfft_internal {

struct ScratchLease {
  void *ptr = nullptr;
  int device = -1;
  int slot = -1;  // -1: stream-ordered allocation to free with hipFreeAsync
};

// A device buffer of at least `bytes`, usable in stream order on `s`.
hipError_t scratch_acquire(size_t bytes, hipStream_t s, ScratchLease *out);
// Work using the lease has been issued on `s`; the buffer returns to the
// cache once that work completes.
hipError_t scratch_release(const ScratchLease &lease, hipStream_t s);

}  // namespace cfft_internal
