// host_stage.cpp -- per-device pool of host staging slots (host_stage.hpp).
#include "host_stage.hpp"

#include <map>
#include <memory>
#include <mutex>
#include <vector>

namespace cfft_internal {
namespace {

std::mutex g_mu;
// slots live until the process ends (their streams and pinned buffers are
// reused by later calls)
std::map<int, std::vector<std::unique_ptr<HostSlot>>> g_slots;

// pinned sizes are powers of two from 64 KiB, so a few sizes serve all calls
size_t slot_size(size_t bytes) {
  size_t b = (size_t)64 << 10;
  while (b < bytes) b <<= 1;
  return b;
}

hipError_t grow(HostSlot *s, size_t bytes) {
  const size_t want = slot_size(bytes);
  if (s->pinned) {
    // the slot's last call has synchronised its stream
    hipError_t e = hipHostFree(s->pinned);
    s->pinned = s->mapped = nullptr;
    s->bytes = 0;
    if (e != hipSuccess) return e;
  }
  hipError_t e = hipHostMalloc(&s->pinned, want, hipHostMallocMapped);
  if (e != hipSuccess) {
    s->pinned = nullptr;
    return e;
  }
  e = hipHostGetDevicePointer(&s->mapped, s->pinned, 0);
  if (e != hipSuccess) {
    (void)hipHostFree(s->pinned);
    s->pinned = s->mapped = nullptr;
    return e;
  }
  s->bytes = want;
  return hipSuccess;
}

}  // namespace

hipError_t host_slot_acquire(size_t bytes, HostSlot **out) {
  *out = nullptr;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  HostSlot *slot = nullptr;
  {
    std::lock_guard<std::mutex> g(g_mu);
    std::vector<std::unique_ptr<HostSlot>> &v = g_slots[dev];
    // an idle slot that fits, else the largest idle one, else a new one
    for (auto &s : v)
      if (!s->busy && s->bytes >= bytes && (!slot || s->bytes < slot->bytes)) slot = s.get();
    if (!slot)
      for (auto &s : v)
        if (!s->busy && (!slot || s->bytes > slot->bytes)) slot = s.get();
    if (!slot) {
      v.push_back(std::unique_ptr<HostSlot>(new HostSlot));
      slot = v.back().get();
      slot->device = dev;
    }
    slot->busy = true;
  }
  if (!slot->stream) e = hipStreamCreateWithFlags(&slot->stream, hipStreamNonBlocking);
  if (e == hipSuccess && slot->bytes < bytes) e = grow(slot, bytes);
  if (e != hipSuccess) {
    host_slot_release(slot);
    return e;
  }
  *out = slot;
  return hipSuccess;
}

void host_slot_release(HostSlot *slot) {
  if (!slot) return;
  std::lock_guard<std::mutex> g(g_mu);
  slot->busy = false;
}

}  // namespace cfft_internal
