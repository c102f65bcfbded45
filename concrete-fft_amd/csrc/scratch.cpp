// scratch.cpp -- per-device, stream-ordered scratch cache (see scratch.hpp).
#include "scratch.hpp"

#include <map>
#include <mutex>
#include <vector>

namespace cfft_internal {
namespace {

constexpr size_t kMaxEntries = 8;
constexpr size_t kMaxBytes = (size_t)2 << 30;
// 32 MiB holds the split-form external product of up to 1024 spectra of 2048
// points (serving-size batches) and the small exact products; larger calls
// take larger power-of-two entries
constexpr size_t kMinEntry = (size_t)32 << 20;

// entry sizes are powers of two (at least kMinEntry), so the few entries made
// early serve every later request of the small-batch paths: a synchronous
// hipMalloc in the middle of a serving run stalls it for milliseconds
size_t entry_size(size_t bytes) {
  size_t b = kMinEntry;
  while (b < bytes) b <<= 1;
  return b;
}

struct Entry {
  void *ptr = nullptr;
  size_t bytes = 0;
  hipEvent_t done = nullptr;  // recorded after the last use's work
  hipStream_t last = nullptr;
  bool used = false;          // `done` has been recorded at least once
  bool leased = false;        // between acquire and release on the host
};

struct DevicePool {
  std::vector<Entry> entries;
  size_t bytes = 0;
};

std::mutex g_mu;
std::map<int, DevicePool> g_pools;

// an entry this stream can take without waiting on the device: never used,
// last used on this stream (stream order), or its last use has completed
bool idle_for(Entry &e, hipStream_t s) {
  if (!e.used || e.last == s) return true;
  return hipEventQuery(e.done) == hipSuccess;
}

}  // namespace

hipError_t scratch_acquire(size_t bytes, hipStream_t s, ScratchLease *out) {
  *out = ScratchLease{};
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (bytes == 0) bytes = 1;
  // under stream capture: a graph allocation node, as before
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  e = hipStreamIsCapturing(s, &cap);
  if (e != hipSuccess) return e;
  if (cap != hipStreamCaptureStatusNone) {
    out->device = dev;
    return hipMallocAsync(&out->ptr, bytes, s);
  }
  std::lock_guard<std::mutex> g(g_mu);
  DevicePool &pool = g_pools[dev];
  // 1. the smallest idle entry that fits
  int best = -1;
  for (size_t i = 0; i < pool.entries.size(); ++i) {
    Entry &en = pool.entries[i];
    if (en.leased || en.bytes < bytes || !idle_for(en, s)) continue;
    if (best < 0 || en.bytes < pool.entries[(size_t)best].bytes) best = (int)i;
  }
  // 2. a new entry
  const size_t eb = entry_size(bytes);
  if (best < 0 && pool.entries.size() < kMaxEntries && pool.bytes + eb <= kMaxBytes) {
    Entry en;
    en.bytes = eb;
    e = hipMalloc(&en.ptr, eb);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&en.done, hipEventDisableTiming);
    if (e != hipSuccess) {
      if (en.ptr) (void)hipFree(en.ptr);
      (void)hipGetLastError();  // fall through to the stream-ordered path
    } else {
      pool.entries.push_back(en);
      pool.bytes += eb;
      best = (int)pool.entries.size() - 1;
    }
  }
  // 3. a busy entry that fits: this stream waits for its last use
  if (best < 0) {
    for (size_t i = 0; i < pool.entries.size(); ++i) {
      Entry &en = pool.entries[i];
      if (!en.leased && en.bytes >= bytes) {
        e = hipStreamWaitEvent(s, en.done, 0);
        if (e != hipSuccess) return e;
        best = (int)i;
        break;
      }
    }
  }
  if (best < 0) {  // above the cache's limits
    e = hipMallocAsync(&out->ptr, bytes, s);
    out->device = dev;
    return e;
  }
  Entry &en = pool.entries[(size_t)best];
  // always ordered after the entry's last use: free when it has completed or
  // was on this stream, and correct even if a stream handle was reused
  if (en.used) {
    e = hipStreamWaitEvent(s, en.done, 0);
    if (e != hipSuccess) return e;
  }
  en.leased = true;
  out->ptr = en.ptr;
  out->device = dev;
  out->slot = best;
  return hipSuccess;
}

hipError_t scratch_release(const ScratchLease &lease, hipStream_t s) {
  if (!lease.ptr) return hipSuccess;
  if (lease.slot < 0) return hipFreeAsync(lease.ptr, s);
  std::lock_guard<std::mutex> g(g_mu);
  Entry &en = g_pools[lease.device].entries[(size_t)lease.slot];
  hipError_t e = hipEventRecord(en.done, s);
  if (e != hipSuccess) return e;  // completion unknown: the entry stays leased, never reused
  en.used = true;
  en.last = s;
  en.leased = false;
  return hipSuccess;
}

}  // namespace cfft_internal
