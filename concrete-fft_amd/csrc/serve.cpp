// serve.cpp -- request server: continuous batching of negacyclic products and
// external products on one device, with admission by free device memory.
//
// Threads: callers submit requests (any thread); one dispatcher thread owns
// every device resource.  The dispatcher repeatedly
//   1. retires batches whose completion event has fired: copies results from
//      pinned staging to the callers' buffers, releases their device and
//      staging ranges, marks their tickets done;
//   2. while a launch slot (stream) is free, forms a batch from the waiting
//      requests -- the group (kind, N, key) of the oldest request, in arrival
//      order, as many as fit the batch limit and the free device / staging
//      memory -- and launches it: one host-to-device copy, one batched kernel,
//      one device-to-host copy, one event (device-buffer requests: the
//      callers' device buffers are named by a table of addresses in mapped
//      pinned staging; the wave-local external-product kernels read and
//      write them through it, other batches are gathered into the batch's
//      block by a kernel and scattered back by another; nothing crosses the
//      bus but the table);
//   3. sleeps until a submission arrives, or polls while batches are in flight.
// So requests that arrive while others run on the device are admitted as soon
// as a slot and memory free up, and batches on different streams overlap.
// A request that does not fit the free memory waits (admission); one larger
// than the whole budget is refused at submission.
#include <hip/hip_runtime.h>

#include <array>
#include <atomic>
#include <chrono>
#include <condition_variable>

#include <sys/prctl.h>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "cfft.h"
#include "launch.hpp"
#include "plan.hpp"

using cfft_internal::hip_error;
using cfft_internal::set_error;

namespace {

using Clock = std::chrono::steady_clock;

constexpr size_t kAlign = 256;
// an idle dispatcher polls for new submissions this long before it blocks on
// the condition variable (a blocked thread's wake-up adds tens of us)
constexpr int kIdleSpinUs = 500;
constexpr int kMaxKeys = 256;
constexpr int kMaxLog2N = 14;
// rows of one gather / scatter launch (two input rows per product request)
constexpr size_t kMaxDeviceBatch = 32767;
// host-buffer external-product batches up to this size are read and written
// by the wave-local kernels in the mapped pinned staging itself (no DMA
// copies: each input is read once per gadget level over the bus, which
// larger batches pay more for than two copies)
constexpr size_t kHostDirectMax = 16;

size_t round_up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

// first-fit range allocator over [0, cap)
class RangeAlloc {
 public:
  void reset(size_t cap) {
    cap_ = cap;
    free_.clear();
    if (cap) free_[0] = cap;
    used_ = peak_ = 0;
  }
  bool alloc(size_t len, size_t *off) {
    len = round_up(len);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < len) continue;
      *off = it->first;
      const size_t rest = it->second - len, next = it->first + len;
      free_.erase(it);
      if (rest) free_[next] = rest;
      used_ += len;
      if (used_ > peak_) peak_ = used_;
      return true;
    }
    return false;
  }
  void release(size_t off, size_t len) {
    len = round_up(len);
    used_ -= len;
    auto it = free_.emplace(off, len).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
  size_t capacity() const { return cap_; }
  size_t peak() const { return peak_; }

 private:
  std::map<size_t, size_t> free_;
  size_t cap_ = 0, used_ = 0, peak_ = 0;
};

struct Key {
  size_t N = 0;
  int k1 = 0, levels = 0, base_log = 0;
  cfft_plan plan = nullptr;
  void *d_fourier = nullptr;
};

struct Req {
  uint64_t ticket;
  int kind;
  bool dev;      // in0, in1, out are device addresses
  int key;       // external product
  uint32_t log2N;
  const void *in0, *in1;
  void *out;
  size_t in_bytes, out_bytes;
  size_t host_bytes;  // pinned staging: the data, or the address table
  Clock::time_point t_submit;
};

struct Batch {
  int kind, key;
  bool dev;
  uint32_t log2N;
  std::vector<Req> reqs;
  size_t dev_off = 0, dev_len = 0, host_off = 0, host_len = 0, in_bytes = 0;
  int slot = -1;
};

struct TicketState {
  bool done = false;
  bool claimed = false;   // a thread is waiting on it (at most one may)
  bool released = false;  // the caller gave it up: forgotten when it completes
  int status = CFFT_SUCCESS;
  Clock::time_point t_done;
};

uint32_t ilog2(size_t n) {
  uint32_t l = 0;
  while (((size_t)1 << l) < n) ++l;
  return l;
}

}  // namespace

struct cfft_server_s {
  cfft_server_config cfg{};
  int device = 0;
  // device and pinned staging pools, owned by the dispatcher; h_pool_dev is
  // the device's address of the (mapped) pinned pool
  char *d_pool = nullptr, *h_pool = nullptr, *h_pool_dev = nullptr;
  RangeAlloc dalloc, halloc;
  std::vector<hipStream_t> streams;
  std::vector<hipEvent_t> events;
  std::vector<int> slot_busy;
  std::vector<Batch> inflight;
  // resources published to the dispatcher (written once, then read-only)
  std::array<Key, kMaxKeys> keys{};
  std::atomic<int> nkeys{0};
  std::array<std::atomic<cfft_plan>, kMaxLog2N + 1> plans{};
  std::mutex plan_mu;  // creation of plans and keys
  // submission queue and tickets
  std::mutex mu;
  std::condition_variable cv_submit, cv_done;
  std::deque<Req> pending;
  std::unordered_map<uint64_t, TicketState> tickets;
  uint64_t next_ticket = 1;
  bool stop = false;
  std::thread worker;
  // dispatcher-local
  std::deque<Req> waiting;
  // statistics
  std::atomic<uint64_t> n_batches{0}, n_done{0}, n_admission_waits{0}, max_inflight{0};
  std::atomic<uint64_t> n_submitted{0};  // bumped after each push to `pending`

  ~cfft_server_s();
  void run();
  void retire(Batch &b, hipError_t err);
  bool launch_one();
  hipError_t launch(Batch &b);
  size_t dev_scratch(const Batch &b, size_t m) const;
};

namespace {

size_t req_in_bytes(int kind, const Key *k, uint32_t log2N) {
  const size_t N = (size_t)1 << log2N;
  if (kind == CFFT_REQUEST_EXTERNAL_PRODUCT) return (size_t)k->k1 * N * 8;
  return 2 * N * 8;
}
size_t req_out_bytes(int kind, const Key *k, uint32_t log2N) {
  const size_t N = (size_t)1 << log2N;
  if (kind == CFFT_REQUEST_EXTERNAL_PRODUCT) return (size_t)k->k1 * N * 8;
  return N * 8;
}

// device bytes beyond inputs and outputs: the split external product's
// Fourier scratch for `m` GLWEs of one key (one pass's spectra)
size_t key_scratch(const Key &k, size_t m) {
  return round_up(cfft_internal::external_product_scratch_bytes(k.plan, k.k1, k.levels, m));
}

int get_plan(cfft_server_s *s, uint32_t log2N, cfft_plan *out) {
  cfft_plan p = s->plans[log2N].load(std::memory_order_acquire);
  if (!p) {
    std::lock_guard<std::mutex> g(s->plan_mu);
    p = s->plans[log2N].load(std::memory_order_acquire);
    if (!p) {
      const size_t n = ((size_t)1 << log2N) / 2;
      int rc = cfft_unordered_plan_create(n, CFFT_METHOD_USER_PROVIDED, CFFT_DIF16,
                                          n < 1024 ? n : 1024, 0.0, &p);
      if (rc) return rc;
      s->plans[log2N].store(p, std::memory_order_release);
    }
  }
  *out = p;
  return CFFT_SUCCESS;
}

}  // namespace

cfft_server_s::~cfft_server_s() {
  {
    std::lock_guard<std::mutex> g(mu);
    stop = true;
  }
  cv_submit.notify_all();
  if (worker.joinable()) worker.join();
  for (hipStream_t st : streams) (void)hipStreamSynchronize(st);
  for (hipEvent_t e : events) (void)hipEventDestroy(e);
  for (hipStream_t st : streams) (void)hipStreamDestroy(st);
  for (int i = 0; i < nkeys.load(); ++i) {
    if (keys[i].d_fourier) (void)hipFree(keys[i].d_fourier);
  }
  for (auto &p : plans) {
    cfft_plan q = p.load();
    if (q) cfft_plan_destroy(q);
  }
  if (d_pool) (void)hipFree(d_pool);
  if (h_pool) (void)hipHostFree(h_pool);
}

size_t cfft_server_s::dev_scratch(const Batch &b, size_t m) const {
  return b.kind == CFFT_REQUEST_EXTERNAL_PRODUCT ? key_scratch(keys[b.key], m) : 0;
}

hipError_t cfft_server_s::launch(Batch &b) {
  hipStream_t st = streams[b.slot];
  char *h_in = h_pool + b.host_off, *d_in = d_pool + b.dev_off;
  const size_t m = b.reqs.size();
  const size_t N = (size_t)1 << b.log2N;
  const size_t in_rows = b.kind == CFFT_REQUEST_EXTERNAL_PRODUCT ? m : 2 * m;
  const size_t in_row_bytes = b.kind == CFFT_REQUEST_EXTERNAL_PRODUCT ? b.reqs[0].in_bytes : N * 8;
  hipError_t e = hipSuccess;
  // external products on the wave-local split kernels read and write
  // through an address table: the callers' device buffers, or for a small
  // host-buffer batch the request slots of the mapped pinned staging
  const bool tables = b.kind == CFFT_REQUEST_EXTERNAL_PRODUCT &&
                      cfft_internal::external_product_tables_supported(keys[b.key].plan);
  const bool direct = tables && (b.dev || m <= kHostDirectMax);
  const size_t host_tab_off = b.dev ? 0 : round_up(b.in_bytes) + round_up(m * b.reqs[0].out_bytes);
  const uint64_t *tab_dev = reinterpret_cast<const uint64_t *>(h_pool_dev + b.host_off + host_tab_off);
  if (b.dev) {
    // address table: input rows (products: all A rows, then all B rows),
    // then output rows; the kernels read it over the bus
    uint64_t *tab = reinterpret_cast<uint64_t *>(h_in);
    for (size_t i = 0; i < m; ++i) {
      tab[i] = (uint64_t)(uintptr_t)b.reqs[i].in0;
      if (b.kind == CFFT_REQUEST_POLYMUL) tab[m + i] = (uint64_t)(uintptr_t)b.reqs[i].in1;
      tab[in_rows + i] = (uint64_t)(uintptr_t)b.reqs[i].out;
    }
    if (!direct)
      e = cfft_launch::gather_rows(tab_dev, in_rows, in_row_bytes / 8,
                                   reinterpret_cast<uint64_t *>(d_in), st);
  } else {
    // gather inputs into pinned staging
    for (size_t i = 0; i < m; ++i) {
      memcpy(h_in + i * in_row_bytes, b.reqs[i].in0, in_row_bytes);
      if (b.kind == CFFT_REQUEST_POLYMUL)
        memcpy(h_in + (m + i) * in_row_bytes, b.reqs[i].in1, in_row_bytes);
    }
    if (direct) {
      uint64_t *tab = reinterpret_cast<uint64_t *>(h_in + host_tab_off);
      const uint64_t base = (uint64_t)(uintptr_t)(h_pool_dev + b.host_off);
      for (size_t i = 0; i < m; ++i) {
        tab[i] = base + i * in_row_bytes;
        tab[m + i] = base + round_up(b.in_bytes) + i * b.reqs[0].out_bytes;
      }
    } else {
      e = hipMemcpyAsync(d_in, h_in, b.in_bytes, hipMemcpyHostToDevice, st);
    }
  }
  if (e != hipSuccess) return e;
  char *d_out = d_in + round_up(b.in_bytes);
  int rc;
  if (b.kind == CFFT_REQUEST_EXTERNAL_PRODUCT) {
    const Key &k = keys[b.key];
    // the split form's Fourier scratch sits in this batch's pool block, after
    // its outputs: no allocation and no shared buffer on the serving path
    const size_t io = round_up(b.in_bytes) + round_up(m * b.reqs[0].out_bytes);
    rc = cfft_internal::external_product_with_scratch(
        k.plan, k.d_fourier, k.k1, k.levels, k.base_log,
        direct ? nullptr : reinterpret_cast<const uint64_t *>(d_in),
        direct ? nullptr : reinterpret_cast<uint64_t *>(d_out), m, st, d_in + io, b.dev_len - io,
        false, direct ? tab_dev : nullptr, direct ? tab_dev + m : nullptr);
  } else {
    cfft_plan p = plans[b.log2N].load(std::memory_order_acquire);
    rc = cfft_negacyclic_polymul_u64(p, reinterpret_cast<const uint64_t *>(d_in),
                                     reinterpret_cast<const int64_t *>(d_in + m * N * 8),
                                     reinterpret_cast<uint64_t *>(d_out), m, st);
  }
  if (rc != CFFT_SUCCESS) return hipErrorLaunchFailure;
  const size_t per_out = b.reqs[0].out_bytes;
  if (b.dev) {
    if (!direct)
      e = cfft_launch::scatter_rows(tab_dev + in_rows, m, per_out / 8,
                                    reinterpret_cast<const uint64_t *>(d_out), st);
  } else if (!direct) {
    e = hipMemcpyAsync(h_pool + b.host_off + round_up(b.in_bytes), d_out, m * per_out,
                       hipMemcpyDeviceToHost, st);
  }
  if (e != hipSuccess) return e;
  return hipEventRecord(events[b.slot], st);
}

void cfft_server_s::retire(Batch &b, hipError_t err) {
  if (err == hipSuccess && !b.dev) {
    const char *h_out = h_pool + b.host_off + round_up(b.in_bytes);
    const size_t per = b.reqs[0].out_bytes;
    for (size_t i = 0; i < b.reqs.size(); ++i) memcpy(b.reqs[i].out, h_out + i * per, per);
  }
  dalloc.release(b.dev_off, b.dev_len);
  halloc.release(b.host_off, b.host_len);
  slot_busy[b.slot] = 0;
  const Clock::time_point now = Clock::now();
  {
    std::lock_guard<std::mutex> g(mu);
    for (const Req &r : b.reqs) {
      auto it = tickets.find(r.ticket);
      if (it == tickets.end()) continue;
      if (it->second.released) {
        tickets.erase(it);
        continue;
      }
      it->second.done = true;
      it->second.status = err == hipSuccess ? CFFT_SUCCESS : CFFT_ERROR_DEVICE;
      it->second.t_done = now;
    }
  }
  n_done += b.reqs.size();
  cv_done.notify_all();
}

// forms and launches one batch if a slot is free; false when nothing launched
bool cfft_server_s::launch_one() {
  int slot = -1;
  for (size_t i = 0; i < slot_busy.size(); ++i)
    if (!slot_busy[i]) {
      slot = (int)i;
      break;
    }
  if (slot < 0 || waiting.empty()) return false;
  const Req &head = waiting.front();
  Batch b;
  b.kind = head.kind;
  b.key = head.key;
  b.dev = head.dev;
  b.log2N = head.log2N;
  b.slot = slot;
  const size_t limit = b.dev ? std::min<size_t>(cfg.max_batch, kMaxDeviceBatch) : cfg.max_batch;
  // pinned staging of a batch: its data (host requests, inputs then
  // outputs, then for external products an address table of 2 words per
  // request) or its address table (device requests)
  const bool host_tab = !b.dev && b.kind == CFFT_REQUEST_EXTERNAL_PRODUCT;
  auto host_need = [&](size_t in_sum, size_t out_sum, size_t tab_sum, size_t count) {
    if (b.dev) return round_up(tab_sum);
    return round_up(in_sum) + round_up(out_sum) + (host_tab ? round_up(16 * count) : 0);
  };
  // collect the head's group in arrival order while memory allows
  size_t in_sum = 0, out_sum = 0, tab_sum = 0;
  std::vector<size_t> take;
  for (size_t i = 0; i < waiting.size() && take.size() < limit; ++i) {
    const Req &r = waiting[i];
    if (r.kind != b.kind || r.dev != b.dev || r.log2N != b.log2N || r.key != b.key) continue;
    const size_t need = round_up(in_sum + r.in_bytes) + round_up(out_sum + r.out_bytes);
    const size_t dneed = need + dev_scratch(b, take.size() + 1);
    const size_t hneed = host_need(in_sum + r.in_bytes, out_sum + r.out_bytes,
                                   tab_sum + r.host_bytes, take.size() + 1);
    if (dneed > dalloc.capacity() || hneed > halloc.capacity()) break;
    take.push_back(i);
    in_sum += r.in_bytes;
    out_sum += r.out_bytes;
    tab_sum += r.host_bytes;
  }
  // shrink until both pools have room (admission by free memory)
  while (!take.empty()) {
    const size_t dneed = round_up(in_sum) + round_up(out_sum) + dev_scratch(b, take.size());
    const size_t hneed = host_need(in_sum, out_sum, tab_sum, take.size());
    size_t doff = 0, hoff = 0;
    if (dalloc.alloc(dneed, &doff)) {
      if (halloc.alloc(hneed, &hoff)) {
        b.dev_off = doff;
        b.host_off = hoff;
        b.dev_len = dneed;
        b.host_len = hneed;
        break;
      }
      dalloc.release(doff, dneed);
    }
    const Req &last = waiting[take.back()];
    in_sum -= last.in_bytes;
    out_sum -= last.out_bytes;
    tab_sum -= last.host_bytes;
    take.pop_back();
  }
  if (take.empty()) {
    ++n_admission_waits;
    return false;
  }
  b.in_bytes = in_sum;
  for (size_t k = 0; k < take.size(); ++k) b.reqs.push_back(waiting[take[k]]);
  for (size_t k = take.size(); k-- > 0;) waiting.erase(waiting.begin() + (long)take[k]);
  slot_busy[slot] = 1;
  hipError_t e = launch(b);
  ++n_batches;
  if (e != hipSuccess) {
    (void)hipStreamSynchronize(streams[slot]);
    retire(b, e);
    return true;
  }
  inflight.push_back(std::move(b));
  if (inflight.size() > max_inflight) max_inflight = inflight.size();
  return true;
}

void cfft_server_s::run() {
  (void)hipSetDevice(device);
  // the 5 us poll below would otherwise sleep ~55 us (Linux's default 50 us
  // timer slack), delaying every retirement
  (void)prctl(PR_SET_TIMERSLACK, 1UL, 0UL, 0UL, 0UL);
  for (;;) {
    // 1. retire finished batches
    for (size_t i = 0; i < inflight.size();) {
      hipError_t q = hipEventQuery(events[inflight[i].slot]);
      if (q == hipErrorNotReady) {
        ++i;
        continue;
      }
      retire(inflight[i], q);
      inflight.erase(inflight.begin() + (long)i);
    }
    // 2. take new submissions, launch while slots are free
    {
      std::unique_lock<std::mutex> g(mu);
      if (stop && pending.empty() && waiting.empty() && inflight.empty()) return;
      while (!pending.empty()) {
        waiting.push_back(pending.front());
        pending.pop_front();
      }
    }
    while (launch_one()) {
    }
    // 3. wait
    if (inflight.empty() && waiting.empty()) {
      // a bounded spin catches the next submission without a wake-up
      const uint64_t seen = n_submitted.load(std::memory_order_acquire);
      const Clock::time_point spin_end = Clock::now() + std::chrono::microseconds(kIdleSpinUs);
      while (n_submitted.load(std::memory_order_acquire) == seen && Clock::now() < spin_end)
        __builtin_ia32_pause();
      std::unique_lock<std::mutex> g(mu);
      cv_submit.wait(g, [&] { return stop || !pending.empty(); });
    } else {
      std::this_thread::sleep_for(std::chrono::microseconds(5));
    }
  }
}

extern "C" {

int cfft_server_create(const cfft_server_config *config, cfft_server *out) {
  if (!out) return set_error(CFFT_ERROR_NULL_POINTER, "out is null");
  *out = nullptr;
  cfft_server_config c{};
  if (config) c = *config;
  if (c.max_batch == 0) c.max_batch = 1024;
  if (c.streams == 0) c.streams = 4;
  if (c.streams > 16) return set_error(CFFT_ERROR_INVALID_ARGUMENT, "at most 16 streams");
  size_t free_b = 0, total_b = 0;
  hipError_t e = hipMemGetInfo(&free_b, &total_b);
  if (e != hipSuccess) return hip_error(e, "hipMemGetInfo");
  if (c.hbm_budget_bytes == 0) c.hbm_budget_bytes = std::min<size_t>((size_t)1 << 30, free_b / 4);
  if (c.hbm_budget_bytes > free_b)
    return set_error(CFFT_ERROR_OUT_OF_MEMORY, "server budget exceeds the free device memory");
  if (c.pinned_bytes == 0) c.pinned_bytes = c.hbm_budget_bytes;
  cfft_server_s *s = new cfft_server_s;
  s->cfg = c;
  (void)hipGetDevice(&s->device);
  e = hipMalloc(reinterpret_cast<void **>(&s->d_pool), c.hbm_budget_bytes);
  if (e == hipSuccess)
    e = hipHostMalloc(reinterpret_cast<void **>(&s->h_pool), c.pinned_bytes, hipHostMallocMapped);
  if (e == hipSuccess)
    e = hipHostGetDevicePointer(reinterpret_cast<void **>(&s->h_pool_dev), s->h_pool, 0);
  for (uint32_t i = 0; i < c.streams && e == hipSuccess; ++i) {
    hipStream_t st = nullptr;
    hipEvent_t ev = nullptr;
    e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e == hipSuccess) s->streams.push_back(st);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
    if (e == hipSuccess) s->events.push_back(ev);
  }
  if (e != hipSuccess) {
    delete s;
    return hip_error(e, "server resources");
  }
  s->dalloc.reset(c.hbm_budget_bytes);
  s->halloc.reset(c.pinned_bytes);
  s->slot_busy.assign(c.streams, 0);
  s->worker = std::thread([s] { s->run(); });
  *out = s;
  return CFFT_SUCCESS;
}

int cfft_server_destroy(cfft_server s) {
  delete s;
  return CFFT_SUCCESS;
}

int cfft_server_add_key(cfft_server s, size_t N, int k1, int levels, int base_log,
                        const uint64_t *key, int *key_id) {
  if (!s || !key || !key_id) return set_error(CFFT_ERROR_NULL_POINTER, "null argument");
  if (N < 4 || (N & (N - 1)) || ilog2(N) > kMaxLog2N)
    return set_error(CFFT_ERROR_SIZE, "key polynomial size must be a power of two in [4, 16384]");
  if (k1 < 1 || levels < 1 || base_log < 1 || base_log > 32 || levels * base_log > 64)
    return set_error(CFFT_ERROR_INVALID_ARGUMENT, "invalid GGSW shape");
  std::lock_guard<std::mutex> g(s->plan_mu);
  const int id = s->nkeys.load();
  if (id >= kMaxKeys) return set_error(CFFT_ERROR_SIZE, "too many keys");
  Key &k = s->keys[id];
  const size_t n = N / 2;
  int rc = cfft_unordered_plan_create(n, CFFT_METHOD_USER_PROVIDED, CFFT_DIF16, n < 1024 ? n : 1024,
                                      0.0, &k.plan);
  if (rc) return rc;
  // the GGSW shape must be one the external product accepts
  rc = cfft_external_product(k.plan, nullptr, k1, levels, base_log, nullptr, nullptr, 0, nullptr);
  if (rc) {
    cfft_plan_destroy(k.plan);
    k = Key{};
    return rc;
  }
  const size_t polys = (size_t)k1 * levels * k1;
  void *d_key = nullptr;
  hipStream_t st = nullptr;
  hipError_t e = hipMalloc(&d_key, polys * N * 8);
  if (e == hipSuccess) e = hipMalloc(&k.d_fourier, polys * n * 16);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMemcpyAsync(d_key, key, polys * N * 8, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) {
    rc = cfft_ggsw_to_fourier(k.plan, reinterpret_cast<const uint64_t *>(d_key), k1, levels,
                              k.d_fourier, st);
    if (rc) e = hipErrorLaunchFailure;
  }
  if (st) {
    hipError_t e2 = hipStreamSynchronize(st);
    if (e == hipSuccess) e = e2;
    (void)hipStreamDestroy(st);
  }
  if (d_key) (void)hipFree(d_key);
  if (e != hipSuccess) {
    if (k.d_fourier) (void)hipFree(k.d_fourier);
    cfft_plan_destroy(k.plan);
    k = Key{};
    return rc ? rc : hip_error(e, "key upload");
  }
  k.N = N;
  k.k1 = k1;
  k.levels = levels;
  k.base_log = base_log;
  s->nkeys.store(id + 1, std::memory_order_release);
  *key_id = id;
  return CFFT_SUCCESS;
}

}  // extern "C"

namespace {

// a device-buffer request's buffer: 8-byte aligned device (or managed)
// memory of the server's device
int check_device_buffer(const cfft_server_s *s, const void *p) {
  if ((uintptr_t)p % 8) return set_error(CFFT_ERROR_INVALID_ARGUMENT, "request buffer is not 8-byte aligned");
  hipPointerAttribute_t a{};
  hipError_t e = hipPointerGetAttributes(&a, p);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return set_error(CFFT_ERROR_INVALID_ARGUMENT, "request buffer is not device memory");
  }
  if ((a.type != hipMemoryTypeDevice && a.type != hipMemoryTypeManaged) || a.device != s->device)
    return set_error(CFFT_ERROR_INVALID_ARGUMENT, "request buffer is not memory of the server's device");
  return CFFT_SUCCESS;
}

}  // namespace

extern "C" {

int cfft_server_submit(cfft_server s, const cfft_request *r, uint64_t *ticket) {
  if (!s || !r || !ticket) return set_error(CFFT_ERROR_NULL_POINTER, "null argument");
  Req q{};
  q.dev = (r->kind & CFFT_REQUEST_DEVICE_BUFFERS) != 0;
  q.kind = r->kind & ~CFFT_REQUEST_DEVICE_BUFFERS;
  q.in0 = r->in0;
  q.in1 = r->in1;
  q.out = r->out;
  q.key = -1;
  const Key *k = nullptr;
  if (q.kind == CFFT_REQUEST_EXTERNAL_PRODUCT) {
    if (r->key_id < 0 || r->key_id >= s->nkeys.load(std::memory_order_acquire))
      return set_error(CFFT_ERROR_INVALID_ARGUMENT, "unknown key id");
    k = &s->keys[r->key_id];
    q.key = r->key_id;
    q.log2N = ilog2(k->N);
    if (!r->in0 || !r->out) return set_error(CFFT_ERROR_NULL_POINTER, "null request buffer");
  } else if (q.kind == CFFT_REQUEST_POLYMUL) {
    if (r->N < 4 || (r->N & (r->N - 1)) || ilog2(r->N) > kMaxLog2N)
      return set_error(CFFT_ERROR_SIZE, "polynomial size must be a power of two in [4, 16384]");
    if (!r->in0 || !r->in1 || !r->out) return set_error(CFFT_ERROR_NULL_POINTER, "null request buffer");
    q.log2N = ilog2(r->N);
    cfft_plan p = nullptr;
    int rc = get_plan(s, q.log2N, &p);
    if (rc) return rc;
  } else {
    return set_error(CFFT_ERROR_INVALID_ARGUMENT, "unknown request kind");
  }
  if (q.dev) {
    for (const void *p : {q.in0, q.kind == CFFT_REQUEST_POLYMUL ? q.in1 : q.in0, (const void *)q.out}) {
      const int rc = check_device_buffer(s, p);
      if (rc) return rc;
    }
  }
  q.in_bytes = req_in_bytes(q.kind, k, q.log2N);
  q.out_bytes = req_out_bytes(q.kind, k, q.log2N);
  q.host_bytes = q.dev ? (q.kind == CFFT_REQUEST_POLYMUL ? 3 : 2) * sizeof(uint64_t)
                       : q.in_bytes + q.out_bytes;
  const size_t need = round_up(q.in_bytes) + round_up(q.out_bytes);
  const size_t dneed = need + (k ? key_scratch(*k, 1) : 0);
  const size_t hneed = q.dev ? round_up(q.host_bytes)
                            : need + (q.kind == CFFT_REQUEST_EXTERNAL_PRODUCT ? round_up(16) : 0);
  if (dneed > s->cfg.hbm_budget_bytes || hneed > s->cfg.pinned_bytes)
    return set_error(CFFT_ERROR_BUSY, "request larger than the server's memory budget");
  {
    std::lock_guard<std::mutex> g(s->mu);
    if (s->stop) return set_error(CFFT_ERROR_BUSY, "server is stopping");
    q.ticket = s->next_ticket++;
    q.t_submit = Clock::now();
    s->tickets[q.ticket] = TicketState{};
    s->pending.push_back(q);
  }
  s->n_submitted.fetch_add(1, std::memory_order_release);
  s->cv_submit.notify_one();
  *ticket = q.ticket;
  return CFFT_SUCCESS;
}

int cfft_server_poll(cfft_server s, uint64_t ticket, int *done) {
  if (!s || !done) return set_error(CFFT_ERROR_NULL_POINTER, "null argument");
  std::lock_guard<std::mutex> g(s->mu);
  auto it = s->tickets.find(ticket);
  if (it == s->tickets.end() || it->second.released)
    return set_error(CFFT_ERROR_INVALID_ARGUMENT, "unknown ticket");
  *done = it->second.done ? 1 : 0;
  return CFFT_SUCCESS;
}

int cfft_server_release(cfft_server s, uint64_t ticket) {
  if (!s) return set_error(CFFT_ERROR_NULL_POINTER, "server is null");
  std::lock_guard<std::mutex> g(s->mu);
  auto it = s->tickets.find(ticket);
  if (it == s->tickets.end() || it->second.released)
    return set_error(CFFT_ERROR_INVALID_ARGUMENT, "unknown ticket");
  if (it->second.claimed)
    return set_error(CFFT_ERROR_INVALID_ARGUMENT, "ticket is being waited on");
  if (it->second.done) s->tickets.erase(it);
  else it->second.released = true;
  return CFFT_SUCCESS;
}

}  // extern "C"

namespace {
int wait_ticket(cfft_server s, uint64_t ticket, Clock::time_point *t_done) {
  std::unique_lock<std::mutex> g(s->mu);
  auto it = s->tickets.find(ticket);
  if (it == s->tickets.end() || it->second.released)
    return set_error(CFFT_ERROR_INVALID_ARGUMENT, "unknown ticket");
  if (it->second.claimed)
    return set_error(CFFT_ERROR_INVALID_ARGUMENT, "ticket is already being waited on");
  it->second.claimed = true;
  // submissions may rehash the map while this thread sleeps: look the entry
  // up again after every wake-up (the claim keeps it from being erased)
  s->cv_done.wait(g, [&] { return s->tickets.find(ticket)->second.done; });
  it = s->tickets.find(ticket);
  const int st = it->second.status;
  if (t_done) *t_done = it->second.t_done;
  s->tickets.erase(it);
  if (st != CFFT_SUCCESS) return set_error(st, "request failed on the device");
  return CFFT_SUCCESS;
}
}  // namespace

extern "C" {

int cfft_server_wait(cfft_server s, uint64_t ticket) {
  if (!s) return set_error(CFFT_ERROR_NULL_POINTER, "server is null");
  return wait_ticket(s, ticket, nullptr);
}

int cfft_server_replay(cfft_server s, const cfft_request *reqs, size_t n, const double *arrival_s,
                       double *latency_s) {
  if (!s || (n && (!reqs || !arrival_s || !latency_s)))
    return set_error(CFFT_ERROR_NULL_POINTER, "null argument");
  std::vector<uint64_t> tickets(n);
  std::vector<Clock::time_point> due(n);
  // arrivals are kept to the microsecond: a 1 us timer slack on this thread
  // while replaying (Linux's default 50 us would blur the sleeps)
  const int old_slack = prctl(PR_GET_TIMERSLACK, 0UL, 0UL, 0UL, 0UL);
  (void)prctl(PR_SET_TIMERSLACK, 1UL, 0UL, 0UL, 0UL);
  const Clock::time_point t0 = Clock::now() + std::chrono::milliseconds(1);
  int rc = CFFT_SUCCESS;
  size_t submitted = 0;
  for (size_t i = 0; i < n; ++i) {
    due[i] = t0 + std::chrono::duration_cast<Clock::duration>(
                      std::chrono::duration<double>(arrival_s[i]));
    for (;;) {
      const auto now = Clock::now();
      if (now >= due[i]) break;
      if (due[i] - now > std::chrono::microseconds(200))
        std::this_thread::sleep_for(due[i] - now - std::chrono::microseconds(100));
    }
    rc = cfft_server_submit(s, &reqs[i], &tickets[i]);
    if (rc) break;
    ++submitted;
  }
  for (size_t i = 0; i < submitted; ++i) {
    Clock::time_point done;
    const int w = wait_ticket(s, tickets[i], &done);
    if (w && !rc) rc = w;
    latency_s[i] = std::chrono::duration<double>(done - due[i]).count();
  }
  if (old_slack > 0) (void)prctl(PR_SET_TIMERSLACK, (unsigned long)old_slack, 0UL, 0UL, 0UL);
  return rc;
}

int cfft_server_stats(cfft_server s, cfft_server_statistics *out) {
  if (!s || !out) return set_error(CFFT_ERROR_NULL_POINTER, "null argument");
  out->batches = s->n_batches.load();
  out->completed = s->n_done.load();
  out->admission_waits = s->n_admission_waits.load();
  out->max_inflight = s->max_inflight.load();
  out->hbm_budget_bytes = s->cfg.hbm_budget_bytes;
  out->streams = s->cfg.streams;
  return CFFT_SUCCESS;
}

}  // extern "C"
