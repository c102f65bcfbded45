// uneven_api.cpp -- C ABI of the uneven batch: many f64 negacyclic round-trip
// jobs of mixed N, run as one persistent launch fed by a device tile queue
// (replayed from a captured HIP graph), or in the plain form of one launch
// per size class.
//
// A schedule is built once from the job list: one unordered plan per size
// class (fixed: the caller's algorithm with base_n = min(n, 1024); measured:
// the fastest of a candidate set, timed on the device with this schedule's
// own kernels), tiles of up to 8192/n jobs of one class, queued largest class
// first so the long tiles start early and the short ones fill the tail.
// Classes whose plan is Dif16 with base_n = min(n, 1024) (256 <= n <= 8192)
// run the register-resident Dif16 body in their own persistent launch, fed
// by a device tile queue or dealt statically (measured: the queue's one
// device-scope atomic per tile is a serialised round trip); the other
// classes run the generic body in a second persistent launch fed by a
// device tile queue.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "cfft.h"
#include "host_tables.hpp"
#include "plan.hpp"
#include "uneven.hpp"

using cfft_internal::hip_error;
using cfft_internal::set_error;
using cfft_launch::dc64;
using cfft_ub::UClass;
using cfft_ub::UJob;
using cfft_ub::UTile;

namespace {

struct DevSchedule {
  UJob *jobs = nullptr;
  uint32_t *order = nullptr;
  UTile *tiles = nullptr;
  UClass *classes = nullptr;
  dc64 *tw_r16 = nullptr;       // twiddle block of the 1024-point base transform
  uint32_t *counter = nullptr;  // tile queues: Dif16 partitions, then the generic body
  uint32_t ntiles = 0;
  uint32_t ntiles_r16 = 0;      // tiles [0, ntiles_r16) are Dif16 classes
  void release() {
    if (jobs) (void)hipFree(jobs);
    if (order) (void)hipFree(order);
    if (tiles) (void)hipFree(tiles);
    if (classes) (void)hipFree(classes);
    if (tw_r16) (void)hipFree(tw_r16);
    if (counter) (void)hipFree(counter);
    *this = DevSchedule{};
  }
};

struct ClassRange {
  uint32_t log2n = 0;
  bool r16 = false;
  uint32_t first_tile = 0, ntiles = 0;
};

// tiles of `order` (job indices grouped by class)
struct HostSchedule {
  std::vector<uint32_t> order;
  std::vector<UTile> tiles;
  std::vector<ClassRange> ranges;  // per class, in queue order
  uint32_t ntiles_r16 = 0;
};

template <class T>
hipError_t upload(T **dst, const std::vector<T> &src) {
  if (src.empty()) return hipSuccess;
  hipError_t e = hipMalloc(reinterpret_cast<void **>(dst), src.size() * sizeof(T));
  if (e != hipSuccess) return e;
  return hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice);
}

// jobs_by_class[c]: job indices of class c; classes of the Dif16 body are
// queued first, each group in descending size
HostSchedule make_schedule(const std::vector<std::vector<uint32_t>> &jobs_by_class,
                           const std::vector<UClass> &classes) {
  HostSchedule h;
  std::vector<uint32_t> seq(jobs_by_class.size());
  for (uint32_t c = 0; c < seq.size(); ++c) seq[c] = c;
  std::stable_sort(seq.begin(), seq.end(), [&](uint32_t a, uint32_t b) {
    if (classes[a].r16 != classes[b].r16) return classes[a].r16 > classes[b].r16;
    return classes[a].log2n > classes[b].log2n;
  });
  for (uint32_t c : seq) {
    const auto &js = jobs_by_class[c];
    if (js.empty()) continue;
    ClassRange r;
    r.log2n = classes[c].log2n;
    r.r16 = classes[c].r16 != 0;
    r.first_tile = (uint32_t)h.tiles.size();
    const uint32_t G = cfft_ub::jobs_per_tile(classes[c].log2n, classes[c].r16 != 0);
    const uint32_t base = (uint32_t)h.order.size();
    h.order.insert(h.order.end(), js.begin(), js.end());
    for (uint32_t i = 0; i < js.size(); i += G) {
      UTile t;
      t.cls = c;
      t.first = base + i;
      t.count = std::min<uint32_t>(G, (uint32_t)js.size() - i);
      t.pad = 0;
      h.tiles.push_back(t);
    }
    r.ntiles = (uint32_t)h.tiles.size() - r.first_tile;
    if (r.r16) h.ntiles_r16 += r.ntiles;
    h.ranges.push_back(r);
  }
  return h;
}

// twiddle block of the Dif16 base transform of a plan of size n (rows
// k < 16 of T[k l tw_mul], l < base / 16, then rows of T[16 k p tw_mul],
// p < base / 256; tw_mul = n / base)
std::vector<cfft_host::C64> r16_block(size_t n, size_t base) {
  const size_t mul = n / base;
  const size_t T = base / 16, T2 = std::max<size_t>(T / 16, 1);
  const std::vector<cfft_host::C64> tw = cfft_host::twiddle_table(n);
  std::vector<cfft_host::C64> blk(16 * T + 16 * T2);
  for (size_t k = 0; k < 16; ++k) {
    for (size_t l = 0; l < T; ++l) blk[k * T + l] = tw[k * l * mul];
    for (size_t q = 0; q < T2; ++q) blk[16 * T + k * T2 + q] = tw[(16 * k * q * mul) % n];
  }
  return blk;
}

bool same_bits(const cfft_host::C64 &a, const cfft_host::C64 &b) {
  return std::memcmp(&a, &b, sizeof a) == 0;
}

// The Dif16 body reads every class's base twiddles from the 1024-point
// block at column stride 1024 / base_n; true if the plan's own table agrees
// bit for bit (it does whenever both come from sincospi of the exact
// argument -2 j / n).
bool block_covers(const std::vector<cfft_host::C64> &b1024, const cfft_plan_s *p) {
  const std::vector<cfft_host::C64> own = r16_block(p->n, p->base_n);
  const size_t T = p->base_n / 16, T2 = T / 16, cs = 1024 / p->base_n;
  for (size_t k = 0; k < 16; ++k) {
    for (size_t l = 0; l < T; ++l)
      if (!same_bits(own[k * T + l], b1024[k * 64 + l * cs])) return false;
    for (size_t q = 0; q < T2; ++q)
      if (!same_bits(own[16 * T + k * T2 + q], b1024[1024 + k * 4 + q * cs])) return false;
  }
  return true;
}

hipError_t upload_schedule(DevSchedule &d, const std::vector<UJob> &jobs, const HostSchedule &h,
                           const std::vector<UClass> &classes) {
  hipError_t e = upload(&d.jobs, jobs);
  if (e == hipSuccess) e = upload(&d.order, h.order);
  if (e == hipSuccess) e = upload(&d.tiles, h.tiles);
  if (e == hipSuccess && h.ntiles_r16) {
    const std::vector<cfft_host::C64> blk = r16_block(1024, 1024);
    e = hipMalloc(reinterpret_cast<void **>(&d.tw_r16), blk.size() * sizeof(dc64));
    if (e == hipSuccess)
      e = hipMemcpy(d.tw_r16, blk.data(), blk.size() * sizeof(dc64), hipMemcpyHostToDevice);
  }
  if (e == hipSuccess) e = upload(&d.classes, classes);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&d.counter), cfft_ub::kCounterBytes);
  d.ntiles = (uint32_t)h.tiles.size();
  d.ntiles_r16 = h.ntiles_r16;
  return e;
}

UClass class_of(const cfft_plan_s *p) {
  UClass c{};
  c.fwd = p->fwd_full;
  c.inv = p->inv_full;
  c.tw = p->d_tw;
  c.twist = p->d_twist;
  c.untwist = p->d_untwist;
  uint32_t l = 0, lb = 0;
  while (((size_t)1 << l) < p->n) ++l;
  while (((size_t)1 << lb) < p->base_n) ++lb;
  c.log2n = l;
  c.r16 = 0;
  if (cfft_ub::r16_class(p->algo == CFFT_DIF16, l, lb))
    c.r16 = block_covers(r16_block(1024, 1024), p) ? 1u : 0u;
  return c;
}

// the Dif16 classes (device queue or static deal), then the generic ones
// (device queue), each launch on its own queue counter
hipError_t launch_schedule(const DevSchedule &d, uint32_t grid, bool r16_queue, const double *in,
                           dc64 *f, double *out, hipStream_t s) {
  hipError_t e = hipSuccess;
  if (d.ntiles_r16)
    e = cfft_ub::launch_r16_persistent(d.jobs, d.order, d.tiles, d.ntiles_r16, d.classes,
                                       d.tw_r16, r16_queue ? d.counter : nullptr, grid, in, f,
                                       out, s);
  if (e == hipSuccess && d.ntiles > d.ntiles_r16)
    e = cfft_ub::launch_persistent(d.jobs, d.order, d.tiles + d.ntiles_r16,
                                   d.ntiles - d.ntiles_r16, d.classes,
                                   d.counter + cfft_ub::kGenericCounter, grid, in,
                                   f, out, s);
  return e;
}

// best time of direct launches on the given buffers and stream, after one
// warm-up: at least 3 timed launches, then more until `seconds` have passed
// (at most 1000)
hipError_t time_schedule(const DevSchedule &d, uint32_t grid, bool r16_queue, const double *x,
                         dc64 *f, double *y, hipStream_t s, double seconds, double *ms) {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  double best = 1e30;
  const auto t0 = std::chrono::steady_clock::now();
  for (int it = 0; it < 1001 && e == hipSuccess; ++it) {
    e = hipEventRecord(e0, s);
    if (e == hipSuccess) e = launch_schedule(d, grid, r16_queue, x, f, y, s);
    if (e == hipSuccess) e = hipEventRecord(e1, s);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float t = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
    if (e == hipSuccess && it > 0) best = std::min(best, (double)t);
    const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (it >= 3 && el >= seconds) break;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  *ms = best;
  return e;
}

// zero-filled stream-ordered timing buffers
struct TuneBufs {
  double *x = nullptr, *y = nullptr;
  dc64 *f = nullptr;
  hipError_t alloc(size_t nx, size_t nf, size_t ny, hipStream_t s) {
    nx = std::max<size_t>(nx, 1);
    ny = std::max<size_t>(ny, 1);
    nf = std::max<size_t>(nf, 1);
    hipError_t e = hipMallocAsync(reinterpret_cast<void **>(&x), nx * sizeof(double), s);
    if (e == hipSuccess) e = hipMallocAsync(reinterpret_cast<void **>(&y), ny * sizeof(double), s);
    if (e == hipSuccess) e = hipMallocAsync(reinterpret_cast<void **>(&f), nf * sizeof(dc64), s);
    if (e == hipSuccess) e = hipMemsetAsync(x, 0, nx * sizeof(double), s);
    return e;
  }
  // frees in stream order and waits for the stream; the first failure wins
  hipError_t free_on(hipStream_t s, hipError_t e) {
    for (void *p : {(void *)x, (void *)y, (void *)f})
      if (p) {
        const hipError_t e2 = hipFreeAsync(p, s);
        if (e == hipSuccess) e = e2;
      }
    x = y = nullptr;
    f = nullptr;
    const hipError_t e3 = hipStreamSynchronize(s);
    return e == hipSuccess ? e3 : e;
  }
};

struct Candidate {
  int algo;
  size_t base_n;
};

std::vector<Candidate> candidates(uint32_t log2n) {
  const size_t n = (size_t)1 << log2n;
  std::vector<Candidate> out;
  std::vector<size_t> bases{std::min<size_t>(n, 1024)};
  if (n >= 1024) bases.push_back(512);
  for (size_t b : bases)
    for (int a : {CFFT_DIF4, CFFT_DIF8, CFFT_DIF16, CFFT_DIT16}) out.push_back({a, b});
  return out;
}

int grid_size(uint32_t *grid) {
  int dev = 0, cus = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess) return hip_error(e, "device query");
  if (cus <= 0) return set_error(CFFT_ERROR_DEVICE, "device reports no compute units");
  int maxlds = 0;
  e = hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
  if (e != hipSuccess) return hip_error(e, "device query");
  if ((size_t)maxlds < cfft_ub::kR16LdsBytes + 64)
    return set_error(CFFT_ERROR_SIZE, "device LDS per workgroup is below the uneven kernel's need");
  *grid = (uint32_t)cus;  // the workgroup's LDS fills a CU
  return CFFT_SUCCESS;
}

}  // namespace

struct cfft_uneven_s {
  int device = 0;
  size_t in_len = 0, f_len = 0, out_len = 0;
  std::vector<UJob> jobs;
  std::vector<uint32_t> class_log2n;       // per class
  std::vector<cfft_plan> plans;            // per class
  HostSchedule host;
  DevSchedule dev;
  uint32_t grid = 0;
  bool r16_queue = true;  // Dif16 tiles: device queue (true) or static deal
  std::mutex mu;
  hipStream_t cap = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  const void *g_in = nullptr, *g_f = nullptr, *g_out = nullptr;
  ~cfft_uneven_s() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (cap) (void)hipStreamDestroy(cap);
    dev.release();
    for (cfft_plan p : plans)
      if (p) cfft_plan_destroy(p);
  }
};

namespace {

// time one class alone with a candidate plan on packed synthetic buffers
int time_candidate(uint32_t log2n, size_t njobs, const Candidate &c, uint32_t grid,
                   hipStream_t s, double seconds, double *ms) {
  cfft_plan p = nullptr;
  int rc = cfft_unordered_plan_create((size_t)1 << log2n, CFFT_METHOD_USER_PROVIDED, c.algo,
                                      c.base_n, 0.0, &p);
  if (rc) return rc;
  const size_t n = (size_t)1 << log2n;
  std::vector<UJob> jobs(njobs);
  for (size_t i = 0; i < njobs; ++i) jobs[i] = UJob{2 * n * i, n * i, 2 * n * i, (uint32_t)log2n, 0};
  std::vector<std::vector<uint32_t>> by_class(1);
  for (uint32_t i = 0; i < njobs; ++i) by_class[0].push_back(i);
  std::vector<UClass> cls{class_of(p)};
  HostSchedule h = make_schedule(by_class, cls);
  DevSchedule d;
  TuneBufs b;
  hipError_t e = upload_schedule(d, jobs, h, cls);
  if (e == hipSuccess) e = b.alloc(2 * n * njobs, n * njobs, 2 * n * njobs, s);
  double best = 1e30, t_static = 1e30;
  if (e == hipSuccess) e = time_schedule(d, grid, true, b.x, b.f, b.y, s, seconds, &best);
  if (e == hipSuccess && d.ntiles_r16)
    e = time_schedule(d, grid, false, b.x, b.f, b.y, s, seconds, &t_static);
  best = std::min(best, t_static);
  e = b.free_on(s, e);
  d.release();
  cfft_plan_destroy(p);
  if (e != hipSuccess) return hip_error(e, "uneven autotuner");
  *ms = best;
  return CFFT_SUCCESS;
}

int check_job(const cfft_uneven_job &j, size_t idx, size_t in_len, size_t f_len, size_t out_len) {
  char m[200];
  if (j.log2_N < cfft_ub::kMinLog2N || j.log2_N > cfft_ub::kMaxLog2N) {
    snprintf(m, sizeof m, "job %zu: log2_N = %u outside [%u, %u]", idx, j.log2_N,
             cfft_ub::kMinLog2N, cfft_ub::kMaxLog2N);
    return set_error(CFFT_ERROR_SIZE, m);
  }
  const uint64_t N = (uint64_t)1 << j.log2_N;
  if (j.in_offset > in_len || in_len - j.in_offset < N || j.out_offset > out_len ||
      out_len - j.out_offset < N || j.fourier_offset > f_len || f_len - j.fourier_offset < N / 2) {
    snprintf(m, sizeof m, "job %zu: buffer range outside the stated buffer lengths", idx);
    return set_error(CFFT_ERROR_INVALID_ARGUMENT, m);
  }
  return CFFT_SUCCESS;
}

// Method::Measure: times the Dif16 launch with the device queue and with the
// static deal on zeroed buffers of the schedule's own sizes, for about
// `seconds` each on the schedule's private stream (skipped, keeping the
// queue, when they would take more than a quarter of free device memory)
int choose_tile_schedule(cfft_uneven_s *u, double seconds) {
  size_t free_b = 0, total_b = 0;
  hipError_t e = hipMemGetInfo(&free_b, &total_b);
  if (e != hipSuccess) return hip_error(e, "device query");
  const size_t need = (u->in_len + u->out_len) * sizeof(double) + u->f_len * sizeof(dc64);
  if (need > free_b / 4) return CFFT_SUCCESS;
  TuneBufs b;
  e = b.alloc(u->in_len, u->f_len, u->out_len, u->cap);
  double t_queue = 0, t_static = 0;
  if (e == hipSuccess) e = time_schedule(u->dev, u->grid, true, b.x, b.f, b.y, u->cap, seconds, &t_queue);
  if (e == hipSuccess) e = time_schedule(u->dev, u->grid, false, b.x, b.f, b.y, u->cap, seconds, &t_static);
  e = b.free_on(u->cap, e);
  if (e != hipSuccess) return hip_error(e, "uneven autotuner");
  u->r16_queue = t_queue <= t_static;
  return CFFT_SUCCESS;
}

int build(cfft_uneven_s *u, const cfft_uneven_job *jobs, size_t njobs, int method, int base_algo,
          double seconds) {
  seconds = std::max(1e-3, seconds);
  std::vector<std::vector<uint32_t>> by_log2n(cfft_ub::kMaxLog2N + 1);
  u->jobs.resize(njobs);
  for (size_t i = 0; i < njobs; ++i) {
    const uint32_t l = jobs[i].log2_N - 1;
    u->jobs[i] = UJob{jobs[i].in_offset, jobs[i].fourier_offset, jobs[i].out_offset, l, 0};
    by_log2n[l].push_back((uint32_t)i);
  }
  int rc = grid_size(&u->grid);
  if (rc) return rc;
  // the schedule's private stream: graph capture, and the autotuner's timing
  hipError_t ep = hipStreamCreateWithFlags(&u->cap, hipStreamNonBlocking);
  if (ep != hipSuccess) return hip_error(ep, "hipStreamCreateWithFlags");
  ep = cfft_ub::prepare();
  if (ep != hipSuccess) return hip_error(ep, "uneven kernel attributes");
  std::vector<std::vector<uint32_t>> by_class;
  std::vector<UClass> classes;
  for (uint32_t l = 0; l < by_log2n.size(); ++l) {
    if (by_log2n[l].empty()) continue;
    const size_t n = (size_t)1 << l;
    Candidate pick{base_algo, std::min<size_t>(n, 1024)};
    if (method == CFFT_METHOD_MEASURE) {
      // time on up to 4 tiles per workgroup of this class
      const size_t sample = std::min<size_t>(
          by_log2n[l].size(), (size_t)4 * u->grid * cfft_ub::groups_per_tile(l));
      double best = 1e30;
      for (const Candidate &c : candidates(l)) {
        double ms = 0;
        rc = time_candidate(l, sample, c, u->grid, u->cap, seconds, &ms);
        if (rc) return rc;
        if (ms < best) {
          best = ms;
          pick = c;
        }
      }
    }
    cfft_plan p = nullptr;
    rc = cfft_unordered_plan_create(n, CFFT_METHOD_USER_PROVIDED, pick.algo, pick.base_n, 0.0, &p);
    if (rc) return rc;
    u->plans.push_back(p);
    u->class_log2n.push_back(l);
    classes.push_back(class_of(p));
    by_class.push_back(by_log2n[l]);
  }
  u->host = make_schedule(by_class, classes);
  // synchronous copies: the schedule is on the device when this returns
  hipError_t e = upload_schedule(u->dev, u->jobs, u->host, classes);
  if (e != hipSuccess) return hip_error(e, "uneven schedule upload");
  if (method == CFFT_METHOD_MEASURE && u->dev.ntiles_r16) return choose_tile_schedule(u, seconds);
  return CFFT_SUCCESS;
}

}  // namespace

extern "C" {

int cfft_uneven_create(const cfft_uneven_job *jobs, size_t njobs, size_t in_len,
                       size_t fourier_len, size_t out_len, int method, int base_algo,
                       double measure_seconds, cfft_uneven *out) {
  if (!out) return set_error(CFFT_ERROR_NULL_POINTER, "out is null");
  *out = nullptr;
  if (njobs && !jobs) return set_error(CFFT_ERROR_NULL_POINTER, "jobs is null");
  if (njobs > 0xffffffffull) return set_error(CFFT_ERROR_SIZE, "too many jobs for one schedule");
  if (method != CFFT_METHOD_USER_PROVIDED && method != CFFT_METHOD_MEASURE)
    return set_error(CFFT_ERROR_INVALID_ARGUMENT, "unknown method");
  if (method == CFFT_METHOD_USER_PROVIDED && (base_algo < CFFT_DIF2 || base_algo > CFFT_DIT16))
    return set_error(CFFT_ERROR_INVALID_ALGO, "unknown FftAlgo");
  for (size_t i = 0; i < njobs; ++i) {
    int rc = check_job(jobs[i], i, in_len, fourier_len, out_len);
    if (rc) return rc;
  }
  cfft_uneven_s *u = new cfft_uneven_s;
  u->in_len = in_len;
  u->f_len = fourier_len;
  u->out_len = out_len;
  (void)hipGetDevice(&u->device);
  int rc = build(u, jobs, njobs, method, base_algo, measure_seconds);
  if (rc) {
    delete u;
    return rc;
  }
  *out = u;
  return CFFT_SUCCESS;
}

int cfft_uneven_destroy(cfft_uneven u) {
  delete u;
  return CFFT_SUCCESS;
}

int cfft_uneven_class_algo(cfft_uneven u, uint32_t log2_N, int *algo, size_t *base_n) {
  if (!u) return set_error(CFFT_ERROR_NULL_POINTER, "schedule is null");
  for (size_t c = 0; c < u->class_log2n.size(); ++c) {
    if (u->class_log2n[c] + 1 != log2_N) continue;
    if (algo) *algo = u->plans[c]->algo;
    if (base_n) *base_n = u->plans[c]->base_n;
    return CFFT_SUCCESS;
  }
  return set_error(CFFT_ERROR_INVALID_ARGUMENT, "no job of this size in the schedule");
}

int cfft_uneven_info(cfft_uneven u, size_t *njobs, size_t *ntiles, size_t *nclasses,
                     uint32_t *grid) {
  if (!u) return set_error(CFFT_ERROR_NULL_POINTER, "schedule is null");
  if (njobs) *njobs = u->jobs.size();
  if (ntiles) *ntiles = u->host.tiles.size();
  if (nclasses) *nclasses = u->class_log2n.size();
  if (grid) *grid = u->grid;
  return CFFT_SUCCESS;
}

int cfft_uneven_run(cfft_uneven u, const double *in, double *fourier, double *out,
                    cfft_stream_t stream) {
  if (!u) return set_error(CFFT_ERROR_NULL_POINTER, "schedule is null");
  if (u->jobs.empty()) return CFFT_SUCCESS;
  if (!in || !fourier || !out) return set_error(CFFT_ERROR_NULL_POINTER, "null buffer");
  std::lock_guard<std::mutex> g(u->mu);
  hipError_t e = hipSuccess;
  if (!u->exec || u->g_in != in || u->g_f != fourier || u->g_out != out) {
    if (u->exec) (void)hipGraphExecDestroy(u->exec);
    if (u->graph) (void)hipGraphDestroy(u->graph);
    u->exec = nullptr;
    u->graph = nullptr;
    e = hipStreamBeginCapture(u->cap, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
      hipError_t el = launch_schedule(u->dev, u->grid, u->r16_queue, in,
                                      reinterpret_cast<dc64 *>(fourier), out, u->cap);
      e = hipStreamEndCapture(u->cap, &u->graph);
      if (el != hipSuccess) e = el;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&u->exec, u->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) return hip_error(e, "uneven graph capture");
    u->g_in = in;
    u->g_f = fourier;
    u->g_out = out;
  }
  return hip_error(hipGraphLaunch(u->exec, (hipStream_t)stream), "uneven graph launch");
}

int cfft_uneven_run_direct(cfft_uneven u, const double *in, double *fourier, double *out,
                           cfft_stream_t stream) {
  if (!u) return set_error(CFFT_ERROR_NULL_POINTER, "schedule is null");
  if (u->jobs.empty()) return CFFT_SUCCESS;
  if (!in || !fourier || !out) return set_error(CFFT_ERROR_NULL_POINTER, "null buffer");
  std::lock_guard<std::mutex> g(u->mu);  // the launches share the queue counters
  return hip_error(launch_schedule(u->dev, u->grid, u->r16_queue, in,
                                   reinterpret_cast<dc64 *>(fourier), out, (hipStream_t)stream),
                   "uneven launch");
}

int cfft_uneven_tile_schedule(cfft_uneven u, int *kind) {
  if (!u) return set_error(CFFT_ERROR_NULL_POINTER, "schedule is null");
  if (!kind) return set_error(CFFT_ERROR_NULL_POINTER, "kind is null");
  *kind = u->r16_queue ? CFFT_UNEVEN_DEVICE_QUEUE : CFFT_UNEVEN_STATIC_DEAL;
  return CFFT_SUCCESS;
}

int cfft_uneven_set_tile_schedule(cfft_uneven u, int kind) {
  if (!u) return set_error(CFFT_ERROR_NULL_POINTER, "schedule is null");
  if (kind != CFFT_UNEVEN_DEVICE_QUEUE && kind != CFFT_UNEVEN_STATIC_DEAL)
    return set_error(CFFT_ERROR_INVALID_ARGUMENT, "unknown tile schedule");
  std::lock_guard<std::mutex> g(u->mu);
  u->r16_queue = kind == CFFT_UNEVEN_DEVICE_QUEUE;
  if (u->exec) (void)hipGraphExecDestroy(u->exec);  // captured with the old choice
  if (u->graph) (void)hipGraphDestroy(u->graph);
  u->exec = nullptr;
  u->graph = nullptr;
  return CFFT_SUCCESS;
}

int cfft_uneven_run_plain(cfft_uneven u, const double *in, double *fourier, double *out,
                          cfft_stream_t stream) {
  if (!u) return set_error(CFFT_ERROR_NULL_POINTER, "schedule is null");
  if (u->jobs.empty()) return CFFT_SUCCESS;
  if (!in || !fourier || !out) return set_error(CFFT_ERROR_NULL_POINTER, "null buffer");
  for (const ClassRange &r : u->host.ranges) {
    hipError_t e = cfft_ub::launch_tiles(u->dev.jobs, u->dev.order, u->dev.tiles, r.first_tile,
                                         r.ntiles, u->dev.classes,
                                         r.r16 ? u->dev.tw_r16 : nullptr, in,
                                         reinterpret_cast<dc64 *>(fourier), out,
                                         (hipStream_t)stream);
    if (e != hipSuccess) return hip_error(e, "uneven plain launch");
  }
  return CFFT_SUCCESS;
}

}  // extern "C"
