// api.cpp -- C ABI of libcfft_amd: plans, f64 transforms, permuted-order
// helpers and the fused negacyclic product.  Kernels live in *.hip files.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "cfft.h"
#include "host_tables.hpp"
#include "launch.hpp"
#include "plan.hpp"
#include "host_stage.hpp"
#include "scratch.hpp"
#include "tune_cache.hpp"

using cfft_launch::dc64;
using cfft_prog::Pass;
using cfft_prog::Program;

namespace cfft_internal {

static thread_local std::string g_last_error;

int set_error(int status, const std::string &msg) {
  g_last_error = msg;
  return status;
}

int hip_error(hipError_t e, const char *where) {
  if (e == hipSuccess) return CFFT_SUCCESS;
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
    return set_error(CFFT_ERROR_OUT_OF_MEMORY, std::string(where) + ": " + hipGetErrorString(e));
  if (e == hipErrorNoDevice)
    return set_error(CFFT_ERROR_NO_DEVICE, std::string(where) + ": " + hipGetErrorString(e));
  return set_error(CFFT_ERROR_DEVICE, std::string(where) + ": " + hipGetErrorString(e));
}

}  // namespace cfft_internal

using cfft_internal::hip_error;
using cfft_internal::set_error;

#define CFFT_HIP(call)                                  \
  do {                                                  \
    hipError_t e__ = (call);                            \
    if (e__ != hipSuccess) return hip_error(e__, #call); \
  } while (0)

cfft_plan_s::~cfft_plan_s() {
  if (d_tw) (void)hipFree(d_tw);
  if (d_freq) (void)hipFree(d_freq);
  if (d_pos) (void)hipFree(d_pos);
  if (d_twist) (void)hipFree(d_twist);
  if (d_tw_r16) (void)hipFree(d_tw_r16);
  if (d_tw_dit16) (void)hipFree(d_tw_dit16);
  if (d_tw_base1024) (void)hipFree(d_tw_base1024);
  if (d_tw_r8) (void)hipFree(d_tw_r8);
  if (d_tw_chunk) (void)hipFree(d_tw_chunk);
  if (d_untwist) (void)hipFree(d_untwist);
  if (d_tw128) (void)hipFree(d_tw128);
}

int cfft_polymul_dispatch(const cfft_plan_s *p, const uint64_t *a, const int64_t *b, uint64_t *c,
                          size_t batch, hipStream_t s);

namespace {

bool is_pow2(size_t n) { return n != 0 && (n & (n - 1)) == 0; }
uint32_t ilog2(size_t n) {
  uint32_t l = 0;
  while (((size_t)1 << l) < n) ++l;
  return l;
}
int algo_radix(int algo) {
  switch (algo) {
    case CFFT_DIF2: case CFFT_DIT2: return 2;
    case CFFT_DIF4: case CFFT_DIT4: return 4;
    case CFFT_DIF8: case CFFT_DIT8: return 8;
    case CFFT_DIF16: case CFFT_DIT16: return 16;
    default: return 0;
  }
}
const char *algo_name(int algo) {
  static const char *names[] = {"Dif2", "Dit2", "Dif4", "Dit4", "Dif8", "Dit8", "Dif16", "Dit16"};
  return (algo >= 0 && algo < 8) ? names[algo] : "?";
}

// The plan's pass lists: cfft_prog::fixed_program (program.hpp), the same
// builder the compile-time kernels use, so their programs match by
// construction.
bool full_program(std::vector<Pass> &fwd, std::vector<Pass> &inv, size_t n, size_t base_n,
                  int algo) {
  const Program f = cfft_prog::fixed_program(algo, ilog2(n), ilog2(base_n), true);
  const Program i = cfft_prog::fixed_program(algo, ilog2(n), ilog2(base_n), false);
  if (f.npass > (uint32_t)cfft_prog::kMaxPasses || i.npass > (uint32_t)cfft_prog::kMaxPasses)
    return false;
  fwd.assign(f.p, f.p + f.npass);
  inv.assign(i.p, i.p + i.npass);
  return true;
}

bool to_program(const std::vector<Pass> &v, size_t first, size_t last, Program &P) {
  if (last - first > (size_t)cfft_prog::kMaxPasses) return false;
  P = Program{};
  P.npass = (uint32_t)(last - first);
  uint32_t chunk = 0;
  for (size_t i = first; i < last; ++i) {
    P.p[i - first] = v[i];
    chunk = std::max<uint32_t>(chunk, v[i].log2len);
  }
  P.log2n = chunk;
  return true;
}

int build_f64_plan(cfft_plan_s *p) {
  const size_t n = p->n;
  const uint32_t log2n = ilog2(n);
  std::vector<Pass> fwd, inv;
  if (!full_program(fwd, inv, n, p->base_n, p->algo))
    return set_error(CFFT_ERROR_SIZE, "transform needs more passes than the device program holds");
  // whole-transform programs for fused kernels
  p->fwd_full = Program{};
  p->inv_full = Program{};
  if (log2n <= cfft_launch::kLdsMaxLog2) {
    to_program(fwd, 0, fwd.size(), p->fwd_full);
    to_program(inv, 0, inv.size(), p->inv_full);
    p->fwd_full.log2n = log2n;
    p->inv_full.log2n = log2n;
  }
  // split off depth passes larger than an LDS chunk
  size_t fg = 0;
  while (fg < fwd.size() && fwd[fg].log2len > cfft_launch::kLdsMaxLog2) ++fg;
  size_t ig = inv.size();
  while (ig > 0 && inv[ig - 1].log2len > cfft_launch::kLdsMaxLog2) --ig;
  p->fwd_global.assign(fwd.begin(), fwd.begin() + (long)fg);
  p->inv_global.assign(inv.begin() + (long)ig, inv.end());
  if (!to_program(fwd, fg, fwd.size(), p->fwd_lds) || !to_program(inv, 0, ig, p->inv_lds))
    return set_error(CFFT_ERROR_SIZE, "transform needs more passes than the device program holds");
  p->chunk_log2_fwd = p->fwd_lds.npass ? p->fwd_lds.log2n : log2n;
  p->chunk_log2_inv = p->inv_lds.npass ? p->inv_lds.log2n : log2n;

  // tables
  std::vector<cfft_host::C64> tw = cfft_host::twiddle_table(n);
  CFFT_HIP(hipMalloc(&p->d_tw, n * sizeof(dc64)));
  CFFT_HIP(hipMemcpy(p->d_tw, tw.data(), n * sizeof(dc64), hipMemcpyHostToDevice));
  const bool dit16 = p->algo == CFFT_DIT16 && p->base_n == n && n == 1024;
  if ((p->algo == CFFT_DIF16 && p->base_n == n && cfft_launch::r16_supported(log2n)) || dit16) {
    const size_t T = n / 16, T2 = T / 16;
    std::vector<cfft_host::C64> blk(16 * T + 16 * T2);
    for (size_t k = 0; k < 16; ++k) {
      for (size_t l = 0; l < T; ++l) blk[k * T + l] = tw[(k * l) % n];
      for (size_t q = 0; q < T2; ++q) blk[16 * T + k * T2 + q] = tw[(16 * k * q) % n];
    }
    dc64 **dst = dit16 ? &p->d_tw_dit16 : &p->d_tw_r16;
    CFFT_HIP(hipMalloc(dst, blk.size() * sizeof(dc64)));
    CFFT_HIP(hipMemcpy(*dst, blk.data(), blk.size() * sizeof(dc64), hipMemcpyHostToDevice));
  }
  if (p->algo == CFFT_DIF16 && p->base_n == n && (n == 256 || n == 512)) {
    // the wave-local n = 256 / 512 kernels read the 1024-point block at
    // column stride 1024 / n: T_1024[(1024 / n) j] = T_n[j] (same sincospi
    // argument), checked here
    const std::vector<cfft_host::C64> t1024 = cfft_host::twiddle_table(1024);
    const size_t cs = 1024 / n;
    bool same = true;
    for (size_t j = 0; j < n && same; ++j)
      same = std::memcmp(&t1024[cs * j], &tw[j], sizeof(cfft_host::C64)) == 0;
    if (same) {
      std::vector<cfft_host::C64> blk(16 * 64 + 16 * 4);
      for (size_t k = 0; k < 16; ++k) {
        for (size_t l = 0; l < 64; ++l) blk[k * 64 + l] = t1024[(k * l) % 1024];
        for (size_t q = 0; q < 4; ++q) blk[1024 + k * 4 + q] = t1024[(16 * k * q) % 1024];
      }
      CFFT_HIP(hipMalloc(&p->d_tw_base1024, blk.size() * sizeof(dc64)));
      CFFT_HIP(hipMemcpy(p->d_tw_base1024, blk.data(), blk.size() * sizeof(dc64),
                         hipMemcpyHostToDevice));
    }
  }
  // Dif16, base_n = 1024, n > 8192: the global depth passes end at a chunk
  // of 2048, 4096 or 8192 points (the last depth level of radix 2, 4 or 8
  // above the base, src/unordered.rs:391-489), which the wave-local kernels
  // run with that level's twiddles
  const bool chunked = p->algo == CFFT_DIF16 && p->base_n == 1024 && n > 8192 &&
                       p->chunk_log2_fwd == p->chunk_log2_inv && p->chunk_log2_fwd >= 11 &&
                       p->chunk_log2_fwd <= 13;
  if (chunked) {
    const size_t c = (size_t)1 << p->chunk_log2_fwd, mul = n / c;
    std::vector<cfft_host::C64> tc(c);
    for (size_t i = 0; i < c; ++i) tc[i] = tw[i * mul];
    CFFT_HIP(hipMalloc(&p->d_tw_chunk, c * sizeof(dc64)));
    CFFT_HIP(hipMemcpy(p->d_tw_chunk, tc.data(), c * sizeof(dc64), hipMemcpyHostToDevice));
  }
  // the 1024-point r16 block of the plan's table, for the wave-local forms
  // with a radix-2 depth pass (n = 2 base_n), the chunks above and the
  // uneven / product kernels
  if (p->algo == CFFT_DIF16 && ((p->base_n == 1024 && n > 1024 && (n <= 8192 || chunked)) ||
                                (p->base_n == 512 && n == 1024))) {
    const size_t mul = n / 1024;
    std::vector<cfft_host::C64> blk(16 * 64 + 16 * 4);
    for (size_t k = 0; k < 16; ++k) {
      for (size_t l = 0; l < 64; ++l) blk[k * 64 + l] = tw[k * l * mul];
      for (size_t q = 0; q < 4; ++q) blk[1024 + k * 4 + q] = tw[16 * k * q * mul];
    }
    CFFT_HIP(hipMalloc(&p->d_tw_base1024, blk.size() * sizeof(dc64)));
    CFFT_HIP(hipMemcpy(p->d_tw_base1024, blk.data(), blk.size() * sizeof(dc64),
                       hipMemcpyHostToDevice));
  }
  if (p->algo == CFFT_DIF8 && p->base_n == n && cfft_launch::r8_supported(log2n)) {
    std::vector<cfft_host::C64> blk(8 * 128 + 8 * 16 + 8 * 2);
    for (size_t k = 0; k < 8; ++k) {
      for (size_t l = 0; l < 128; ++l) blk[k * 128 + l] = tw[(k * l) % n];
      for (size_t q = 0; q < 16; ++q) blk[1024 + k * 16 + q] = tw[(8 * k * q) % n];
      for (size_t q = 0; q < 2; ++q) blk[1024 + 128 + k * 2 + q] = tw[(64 * k * q) % n];
    }
    CFFT_HIP(hipMalloc(&p->d_tw_r8, blk.size() * sizeof(dc64)));
    CFFT_HIP(hipMemcpy(p->d_tw_r8, blk.data(), blk.size() * sizeof(dc64), hipMemcpyHostToDevice));
  }
  const unsigned nbits = log2n, bbits = ilog2(p->base_n);
  std::vector<uint32_t> freq(n);
  p->pos_of_freq.resize(n);
  for (size_t i = 0; i < n; ++i) {
    p->pos_of_freq[i] = (uint32_t)cfft_host::freq_position(nbits, bbits, i);
    freq[i] = (uint32_t)cfft_host::position_freq(nbits, bbits, i);
  }
  CFFT_HIP(hipMalloc(&p->d_freq, n * sizeof(uint32_t)));
  CFFT_HIP(hipMalloc(&p->d_pos, n * sizeof(uint32_t)));
  CFFT_HIP(hipMemcpy(p->d_freq, freq.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice));
  CFFT_HIP(hipMemcpy(p->d_pos, p->pos_of_freq.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice));
  if (p->kind == cfft_plan_s::UNORDERED) {
    const size_t N = 2 * n;
    std::vector<cfft_host::C64> tws = cfft_host::twist_table(N), utw = cfft_host::untwist_table(N);
    CFFT_HIP(hipMalloc(&p->d_twist, n * sizeof(dc64)));
    CFFT_HIP(hipMalloc(&p->d_untwist, n * sizeof(dc64)));
    CFFT_HIP(hipMemcpy(p->d_twist, tws.data(), n * sizeof(dc64), hipMemcpyHostToDevice));
    CFFT_HIP(hipMemcpy(p->d_untwist, utw.data(), n * sizeof(dc64), hipMemcpyHostToDevice));
  }
  return CFFT_SUCCESS;
}

// An ordered plan above one workgroup's LDS (2^13 < n <= 2^16): every pass
// of the Stockham recursion over global memory.  Core passes move the data
// between the caller's buffer and a scratch of batch x n points; the end
// butterfly writes where it reads, so it runs in place, or from the scratch
// back into the caller's buffer when the number of core passes is odd.
int run_global_stockham(const cfft_plan_s *p, const std::vector<Pass> &passes, dc64 *d,
                        size_t batch, size_t stride, hipStream_t s) {
  const uint32_t log2n = ilog2(p->n);
  size_t cores = 0;
  for (const Pass &P : passes)
    if (P.kind == cfft_prog::DIF_CORE || P.kind == cfft_prog::DIT_CORE) ++cores;
  cfft_internal::ScratchLease lease;
  CFFT_HIP(cfft_internal::scratch_acquire(batch * p->n * sizeof(dc64), s, &lease));
  dc64 *cur = d, *oth = reinterpret_cast<dc64 *>(lease.ptr);
  size_t cur_stride = stride, oth_stride = p->n;
  hipError_t e = hipSuccess;
  for (const Pass &P : passes) {
    if (e != hipSuccess) break;
    if (P.kind == cfft_prog::END && cores % 2 == 0) {
      e = cfft_launch::pass_global(P, cur, cur_stride, cur, cur_stride, batch, log2n, p->d_tw, s);
      continue;
    }
    e = cfft_launch::pass_global(P, cur, cur_stride, oth, oth_stride, batch, log2n, p->d_tw, s);
    std::swap(cur, oth);
    std::swap(cur_stride, oth_stride);
  }
  const hipError_t e2 = cfft_internal::scratch_release(lease, s);
  if (e != hipSuccess) return hip_error(e, "global Stockham pass");
  return hip_error(e2, "scratch release");
}

int run_transform(const cfft_plan_s *p, void *buf, size_t batch, size_t stride, bool fwd,
                  hipStream_t s) {
  if (!p) return set_error(CFFT_ERROR_NULL_POINTER, "plan is null");
  if (p->kind == cfft_plan_s::FFT128)
    return set_error(CFFT_ERROR_WRONG_PLAN_KIND, "an fft128 plan takes split double-double buffers");
  if (batch == 0) return CFFT_SUCCESS;
  if (!buf) return set_error(CFFT_ERROR_NULL_POINTER, "buffer is null");
  if (stride < p->n) {
    char m[160];
    snprintf(m, sizeof m, "assertion `left == right` failed\n  left: %zu\n right: %zu", p->n, stride);
    return set_error(CFFT_ERROR_LENGTH_MISMATCH, m);
  }
  const uint32_t log2n = ilog2(p->n);
  dc64 *d = reinterpret_cast<dc64 *>(buf);
  if (p->n == 512 && p->base_n == 512 && p->d_tw_base1024)  // Dif16, wave-local
    return hip_error(cfft_launch::w512_transform(fwd, d, batch, stride, p->d_tw_base1024, s),
                     "w512_transform");
  if (p->algo == CFFT_DIF16 && p->base_n == p->n && cfft_launch::r16_supported(log2n))
    return hip_error(cfft_launch::r16_transform(log2n, fwd, d, batch, stride, p->d_tw_r16, s),
                     "r16_transform");
  if (p->d_tw_r8)
    return hip_error(cfft_launch::r8_transform(fwd, d, batch, stride, p->d_tw_r8, s),
                     "r8_transform");
  if (p->d_tw_dit16)  // Dit16, n = base_n = 1024
    return hip_error(cfft_launch::dit16_transform(fwd, d, batch, stride, p->d_tw_dit16, s),
                     "dit16_transform");
  if (p->n == 1024 && p->base_n == 512 && p->d_tw_base1024)  // Dif16, base_n = 512
    return hip_error(cfft_launch::wd2_512_transform(fwd, d, batch, stride, p->d_tw,
                                                    p->d_tw_base1024, s),
                     "wd2_512_transform");
  if (p->n == 4096 && p->base_n == 1024 && p->d_tw_base1024)  // Dif16
    return hip_error(cfft_launch::depth_transform(fwd, log2n, d, batch, stride, p->d_tw,
                                                  p->d_tw_base1024, s),
                     "depth_transform");
  if (p->n == 8192 && p->base_n == 1024 && p->d_tw_base1024)  // Dif16
    return hip_error(cfft_launch::depth8k_transform(fwd, d, batch, stride, p->d_tw,
                                                    p->d_tw_base1024, s),
                     "depth8k_transform");
  if (p->n == 2048 && p->d_tw_base1024)  // Dif16, base_n = 1024
    return hip_error(cfft_launch::w2048_transform(fwd, d, batch, stride, p->d_tw,
                                                  p->d_tw_base1024, s),
                     "w2048_transform");
  // Dif16, base_n = 1024, n = 16384: one radix-8 level above chunks of 2048
  if (p->n == 16384 && p->d_tw_chunk && p->d_tw_base1024 && p->chunk_log2_fwd == 11)
    return hip_error(cfft_launch::depth16k_transform(fwd, d, batch, stride, p->d_tw,
                                                     p->d_tw_chunk, p->d_tw_base1024, s),
                     "depth16k_transform");
  const std::vector<Pass> &global = fwd ? p->fwd_global : p->inv_global;
  if (!global.empty() && global.front().kind != cfft_prog::DEPTH_FWD &&
      global.front().kind != cfft_prog::DEPTH_INV)
    return run_global_stockham(p, global, d, batch, stride, s);
  // chunks of a contiguous batch on the wave-local kernels
  const bool chunks = p->d_tw_chunk && p->d_tw_base1024 && stride == p->n;
  const uint32_t cl = p->chunk_log2_fwd;
  auto run_chunks = [&]() -> int {
    const uint64_t nc = batch << (log2n - cl);
    if (cl == 11)
      return hip_error(cfft_launch::w2048_transform(fwd, d, nc, 2048, p->d_tw_chunk,
                                                    p->d_tw_base1024, s),
                       "w2048_transform");
    if (cl == 13)
      return hip_error(cfft_launch::depth8k_transform(fwd, d, nc, 8192, p->d_tw_chunk,
                                                      p->d_tw_base1024, s),
                       "depth8k_transform");
    return hip_error(cfft_launch::depth_transform(fwd, cl, d, nc, (uint64_t)1 << cl, p->d_tw_chunk,
                                                  p->d_tw_base1024, s),
                     "depth_transform");
  };
  if (fwd) {
    for (const Pass &P : p->fwd_global)
      CFFT_HIP(cfft_launch::depth_global(P, d, batch, log2n, stride, p->d_tw, s));
    if (chunks) return run_chunks();
    if (p->fwd_lds.npass)
      CFFT_HIP(cfft_launch::fft_lds(p->fwd_lds, d, batch << (log2n - p->chunk_log2_fwd),
                                    log2n - p->chunk_log2_fwd, stride, p->d_tw, s));
  } else {
    if (chunks) {
      const int rc = run_chunks();
      if (rc) return rc;
    } else if (p->inv_lds.npass)
      CFFT_HIP(cfft_launch::fft_lds(p->inv_lds, d, batch << (log2n - p->chunk_log2_inv),
                                    log2n - p->chunk_log2_inv, stride, p->d_tw, s));
    for (const Pass &P : p->inv_global)
      CFFT_HIP(cfft_launch::depth_global(P, d, batch, log2n, stride, p->d_tw, s));
  }
  return CFFT_SUCCESS;
}

// Plan::fwd / Plan::inv on a host buffer (src/unordered.rs:826-839,
// src/ordered.rs:342-372): staged through a pinned slot (host_stage.hpp)
int run_transform_host(const cfft_plan_s *p, double *buf, size_t batch, bool fwd) {
  if (!p) return set_error(CFFT_ERROR_NULL_POINTER, "plan is null");
  if (batch == 0) return CFFT_SUCCESS;
  if (!buf) return set_error(CFFT_ERROR_NULL_POINTER, "buffer is null");
  const size_t n = p->n, per = n * sizeof(dc64);
  return cfft_internal::stage_host(
      batch, per,
      [&](void *pin, size_t b0, size_t nb) { std::memcpy(pin, buf + 2 * n * b0, nb * per); },
      [&](void *d, size_t nb, hipStream_t s) { return run_transform(p, d, nb, n, fwd, s); },
      [&](const void *pin, size_t b0, size_t nb) { std::memcpy(buf + 2 * n * b0, pin, nb * per); });
}

// ---- Method::Measure: time candidate plans on the device ----
// Timing buffers (zero-filled): the transforms run on a batch of
// kTuneTransformBytes, the fused product on a batch whose inputs and output
// fill kTuneProductBytes.
constexpr size_t kTuneTransformBytes = (size_t)32 << 20;
constexpr size_t kTuneProductBytes = (size_t)96 << 20;

struct TuneRig {
  hipStream_t s = nullptr;  // private, non-blocking: the caller's streams run on
  hipEvent_t a = nullptr, b = nullptr;
  void *buf = nullptr;
  size_t bytes = 0;
};

bool tune_product(const cfft_plan_s *p) {
  return p->kind == cfft_plan_s::UNORDERED && p->n >= 2 && p->fwd_full.npass != 0;
}

// Times one candidate on each shape a plan of its kind serves: the forward
// and the inverse transform and, for unordered plans the fused negacyclic
// product runs on (n <= 8192), that product.  Rounds of one launch per shape
// repeat until `seconds` have passed (at least 2 rounds, at most 1000); the
// score is the sum over the shapes of the best time per polynomial.
int time_plan(const cfft_plan_s *p, double seconds, const TuneRig &r, double *score) {
  const size_t n = p->n;
  const size_t tbatch = std::max<size_t>(1, kTuneTransformBytes / (n * sizeof(dc64)));
  const bool prod = tune_product(p) && r.bytes >= kTuneProductBytes;
  const size_t pbatch = prod ? std::max<size_t>(1, kTuneProductBytes / (6 * n * sizeof(uint64_t))) : 0;
  const uint64_t *pa = static_cast<const uint64_t *>(r.buf);
  const int64_t *pb = reinterpret_cast<const int64_t *>(pa + 2 * n * pbatch);
  uint64_t *pc = const_cast<uint64_t *>(pa) + 4 * n * pbatch;
  auto launch = [&](int shape) -> int {
    if (shape < 2) return run_transform(p, r.buf, tbatch, n, shape == 0, r.s);
    return cfft_polymul_dispatch(p, pa, pb, pc, pbatch, r.s);
  };
  const int shapes = prod ? 3 : 2;
  for (int k = 0; k < shapes; ++k) {  // warm-up
    int rc = launch(k);
    if (rc) return rc;
  }
  CFFT_HIP(hipStreamSynchronize(r.s));
  double best[3] = {1e30, 1e30, 1e30};
  const auto t0 = std::chrono::steady_clock::now();
  for (int round = 0; round < 1000; ++round) {
    for (int k = 0; k < shapes; ++k) {
      CFFT_HIP(hipEventRecord(r.a, r.s));
      int rc = launch(k);
      if (rc) return rc;
      CFFT_HIP(hipEventRecord(r.b, r.s));
      CFFT_HIP(hipEventSynchronize(r.b));
      float ms = 0;
      CFFT_HIP(hipEventElapsedTime(&ms, r.a, r.b));
      best[k] = std::min(best[k], (double)ms / (double)(k < 2 ? tbatch : pbatch));
    }
    const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (round >= 1 && el >= seconds) break;
  }
  *score = best[0] + best[1] + (prod ? best[2] : 0.0);
  return CFFT_SUCCESS;
}

int create_f64(cfft_plan_s::Kind kind, size_t n, int algo, size_t base_n, cfft_plan *out) {
  auto *p = new (std::nothrow) cfft_plan_s();
  if (!p) return set_error(CFFT_ERROR_OUT_OF_MEMORY, "plan allocation failed");
  p->kind = kind;
  p->n = n;
  p->base_n = base_n;
  p->algo = algo;
  (void)hipGetDevice(&p->device);
  int rc = build_f64_plan(p);
  if (rc != CFFT_SUCCESS) {
    delete p;
    return rc;
  }
  *out = p;
  return CFFT_SUCCESS;
}

// device key of a tuning-cache entry: model, ISA, compute units, HIP
// runtime and ABI version
int device_key(int dev, std::string *out) {
  static std::mutex mu;
  static std::map<int, std::string> keys;
  std::lock_guard<std::mutex> g(mu);
  auto it = keys.find(dev);
  if (it == keys.end()) {
    hipDeviceProp_t prop;
    CFFT_HIP(hipGetDeviceProperties(&prop, dev));
    int rt = 0;
    CFFT_HIP(hipRuntimeGetVersion(&rt));
    char buf[640];
    snprintf(buf, sizeof buf, "%s|%s|cu%d|hip%d|abi%d", prop.name, prop.gcnArchName,
             prop.multiProcessorCount, rt, CFFT_ABI_VERSION);
    std::string k(buf);
    for (char &c : k)
      if (c == '\t' || c == '\n' || c == '\r') c = ' ';
    it = keys.emplace(dev, k).first;
  }
  *out = it->second;
  return CFFT_SUCCESS;
}

// Method::Measure (src/unordered.rs:553-640, src/ordered.rs:99-180): the
// fastest of the candidate (algo, base_n) pairs, each timed for about
// `seconds` on the shapes the plan serves (time_plan).  Bases: n for ordered
// plans and n <= 256, else 512 and 1024 (those not above n), with all 8
// algorithms each.  The choice is kept in the tuning cache (tune_cache.hpp),
// so a later Measure of the same kind and size on the same device and stack,
// in this process or another, reuses it without timing.  Every HIP failure
// is returned.
int measure(cfft_plan_s::Kind kind, size_t n, double seconds, int *algo, size_t *base_n) {
  int dev = 0;
  CFFT_HIP(hipGetDevice(&dev));
  std::string key;
  int rc = device_key(dev, &key);
  if (rc) return rc;
  const int tk = kind == cfft_plan_s::ORDERED ? cfft_internal::TUNE_ORDERED
                                              : cfft_internal::TUNE_UNORDERED;
  cfft_internal::TuneChoice hit;
  if (cfft_internal::tune_lookup(key, tk, n, &hit)) {
    *algo = hit.algo;
    *base_n = hit.base_n;
    return CFFT_SUCCESS;
  }
  std::vector<size_t> bases;
  if (kind == cfft_plan_s::ORDERED || n <= 256) {
    bases.push_back(n);
  } else {
    for (size_t b : {(size_t)512, (size_t)1024})
      if (b <= n) bases.push_back(b);
  }
  const double per = std::max(1e-3, seconds);
  TuneRig r;
  const bool prod = kind == cfft_plan_s::UNORDERED && n >= 2 && n <= 8192;
  r.bytes = prod ? kTuneProductBytes : std::max(kTuneTransformBytes, n * sizeof(dc64));
  hipError_t e = hipStreamCreateWithFlags(&r.s, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&r.a);
  if (e == hipSuccess) e = hipEventCreate(&r.b);
  if (e == hipSuccess) e = hipMallocAsync(&r.buf, r.bytes, r.s);
  if (e == hipSuccess) e = hipMemsetAsync(r.buf, 0, r.bytes, r.s);
  rc = hip_error(e, "autotuner set-up");
  double best = 1e30;
  int best_algo = CFFT_DIF4;
  size_t best_base = bases.front();
  for (size_t b : bases) {
    for (int a = 0; a < 8 && rc == CFFT_SUCCESS; ++a) {
      cfft_plan p0 = nullptr;
      rc = create_f64(kind, n, a, b, &p0);
      if (rc) break;
      double t = 0;
      rc = time_plan(p0, per, r, &t);
      delete p0;
      if (rc == CFFT_SUCCESS && t < best) {
        best = t;
        best_algo = a;
        best_base = b;
      }
    }
  }
  // tear-down on the private stream only; a failure here is reported when
  // nothing failed before it
  e = hipSuccess;
  if (r.buf) e = hipFreeAsync(r.buf, r.s);
  if (r.s) {
    const hipError_t e2 = hipStreamSynchronize(r.s);
    if (e == hipSuccess) e = e2;
  }
  if (r.a) (void)hipEventDestroy(r.a);
  if (r.b) (void)hipEventDestroy(r.b);
  if (r.s) (void)hipStreamDestroy(r.s);
  if (rc) return rc;
  if (e != hipSuccess) return hip_error(e, "autotuner tear-down");
  *algo = best_algo;
  *base_n = best_base;
  cfft_internal::tune_store(key, tk, n, cfft_internal::TuneChoice{best_algo, best_base});
  return CFFT_SUCCESS;
}

void put_varint(std::vector<uint8_t> &o, uint64_t v) {
  // bincode 1.3 VarintEncoding (options()): < 251 one byte, then tags
  // 251 (u16), 252 (u32), 253 (u64), little endian
  if (v < 251) {
    o.push_back((uint8_t)v);
  } else if (v <= 0xffff) {
    o.push_back(251);
    for (int i = 0; i < 2; ++i) o.push_back((uint8_t)(v >> (8 * i)));
  } else if (v <= 0xffffffffull) {
    o.push_back(252);
    for (int i = 0; i < 4; ++i) o.push_back((uint8_t)(v >> (8 * i)));
  } else {
    o.push_back(253);
    for (int i = 0; i < 8; ++i) o.push_back((uint8_t)(v >> (8 * i)));
  }
}

}  // namespace

extern "C" {

const char *cfft_status_string(int status) {
  switch (status) {
    case CFFT_SUCCESS: return "success";
    case CFFT_ERROR_NOT_POWER_OF_TWO: return "size is not a power of two";
    case CFFT_ERROR_SIZE: return "size out of the supported range";
    case CFFT_ERROR_INVALID_BASE: return "invalid base size";
    case CFFT_ERROR_INVALID_ALGO: return "invalid algorithm";
    case CFFT_ERROR_NULL_POINTER: return "null pointer";
    case CFFT_ERROR_LENGTH_MISMATCH: return "buffer length does not match the plan";
    case CFFT_ERROR_WRONG_PLAN_KIND: return "operation not supported by this plan kind";
    case CFFT_ERROR_OUT_OF_MEMORY: return "out of memory";
    case CFFT_ERROR_DEVICE: return "device error";
    case CFFT_ERROR_INVALID_ARGUMENT: return "invalid argument";
    case CFFT_ERROR_SIZE_OVERFLOW: return "size overflow";
    case CFFT_ERROR_DESERIALIZE: return "malformed serialized buffer";
    case CFFT_ERROR_NO_DEVICE: return "no device";
    case CFFT_ERROR_BUSY: return "busy";
    default: return "unknown status";
  }
}

const char *cfft_last_error_message(void) { return cfft_internal::g_last_error.c_str(); }
int cfft_abi_version(void) { return CFFT_ABI_VERSION; }

// ---- device plumbing ----
int cfft_device_count(int *count) {
  if (!count) return set_error(CFFT_ERROR_NULL_POINTER, "count is null");
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  if (e != hipSuccess) {
    *count = 0;
    return hip_error(e, "hipGetDeviceCount");
  }
  *count = c;
  return CFFT_SUCCESS;
}
int cfft_set_device(int device) { return hip_error(hipSetDevice(device), "hipSetDevice"); }
int cfft_device_alloc(size_t bytes, void **ptr) {
  if (!ptr) return set_error(CFFT_ERROR_NULL_POINTER, "ptr is null");
  *ptr = nullptr;
  if (bytes == 0) return CFFT_SUCCESS;
  return hip_error(hipMalloc(ptr, bytes), "hipMalloc");
}
int cfft_device_free(void *ptr) { return ptr ? hip_error(hipFree(ptr), "hipFree") : CFFT_SUCCESS; }
int cfft_device_mem_info(size_t *free_bytes, size_t *total_bytes) {
  size_t f = 0, t = 0;
  CFFT_HIP(hipMemGetInfo(&f, &t));
  if (free_bytes) *free_bytes = f;
  if (total_bytes) *total_bytes = t;
  return CFFT_SUCCESS;
}
int cfft_memcpy_htod(void *dst, const void *src, size_t bytes, cfft_stream_t stream) {
  if (bytes == 0) return CFFT_SUCCESS;
  return hip_error(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream), "hipMemcpyAsync");
}
int cfft_memcpy_dtoh(void *dst, const void *src, size_t bytes, cfft_stream_t stream) {
  if (bytes == 0) return CFFT_SUCCESS;
  CFFT_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream));
  return hip_error(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize");
}
int cfft_memcpy_dtod(void *dst, const void *src, size_t bytes, cfft_stream_t stream) {
  if (bytes == 0) return CFFT_SUCCESS;
  return hip_error(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream), "hipMemcpyAsync");
}
int cfft_memset(void *dst, int value, size_t bytes, cfft_stream_t stream) {
  if (bytes == 0) return CFFT_SUCCESS;
  return hip_error(hipMemsetAsync(dst, value, bytes, (hipStream_t)stream), "hipMemsetAsync");
}
int cfft_stream_create(cfft_stream_t *stream) {
  if (!stream) return set_error(CFFT_ERROR_NULL_POINTER, "stream is null");
  hipStream_t s = nullptr;
  CFFT_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *stream = s;
  return CFFT_SUCCESS;
}
int cfft_stream_destroy(cfft_stream_t stream) {
  return stream ? hip_error(hipStreamDestroy((hipStream_t)stream), "hipStreamDestroy") : CFFT_SUCCESS;
}
int cfft_stream_synchronize(cfft_stream_t stream) {
  return hip_error(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize");
}
int cfft_device_synchronize(void) { return hip_error(hipDeviceSynchronize(), "hipDeviceSynchronize"); }
int cfft_event_create(cfft_event_t *ev) {
  if (!ev) return set_error(CFFT_ERROR_NULL_POINTER, "event is null");
  hipEvent_t e = nullptr;
  CFFT_HIP(hipEventCreate(&e));
  *ev = e;
  return CFFT_SUCCESS;
}
int cfft_event_destroy(cfft_event_t ev) {
  return ev ? hip_error(hipEventDestroy((hipEvent_t)ev), "hipEventDestroy") : CFFT_SUCCESS;
}
int cfft_event_record(cfft_event_t ev, cfft_stream_t stream) {
  return hip_error(hipEventRecord((hipEvent_t)ev, (hipStream_t)stream), "hipEventRecord");
}
int cfft_event_synchronize(cfft_event_t ev) {
  return hip_error(hipEventSynchronize((hipEvent_t)ev), "hipEventSynchronize");
}
int cfft_event_elapsed_ms(cfft_event_t start, cfft_event_t stop, float *ms) {
  if (!ms) return set_error(CFFT_ERROR_NULL_POINTER, "ms is null");
  return hip_error(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop), "hipEventElapsedTime");
}

// ---- tuning cache: the device key of the current device ----
int cfft_tune_device_key(char *buf, size_t cap, size_t *len) {
  int dev = 0;
  CFFT_HIP(hipGetDevice(&dev));
  std::string k;
  int rc = device_key(dev, &k);
  if (rc) return rc;
  if (len) *len = k.size();
  if (!buf) return cap ? set_error(CFFT_ERROR_NULL_POINTER, "buf is null") : CFFT_SUCCESS;
  if (cap <= k.size()) return set_error(CFFT_ERROR_SIZE, "buffer too small for the key");
  std::memcpy(buf, k.c_str(), k.size() + 1);
  return CFFT_SUCCESS;
}

// ---- plans ----
int cfft_ordered_plan_create(size_t n, int method, int algo, double measure_seconds,
                             cfft_plan *out) {
  if (!out) return set_error(CFFT_ERROR_NULL_POINTER, "out is null");
  *out = nullptr;
  if (!is_pow2(n)) return set_error(CFFT_ERROR_NOT_POWER_OF_TWO, "assertion failed: n.is_power_of_two()");
  if (ilog2(n) >= 11) return set_error(CFFT_ERROR_SIZE, "assertion failed: n.trailing_zeros() < 11");
  size_t base = n;
  if (method == CFFT_METHOD_MEASURE) {
    int rc = measure(cfft_plan_s::ORDERED, n, measure_seconds, &algo, &base);
    if (rc) return rc;
  } else if (method != CFFT_METHOD_USER_PROVIDED) {
    return set_error(CFFT_ERROR_INVALID_ARGUMENT, "unknown method");
  }
  if (algo_radix(algo) == 0) return set_error(CFFT_ERROR_INVALID_ALGO, "unknown FftAlgo");
  return create_f64(cfft_plan_s::ORDERED, n, algo, n, out);
}

int cfft_ordered_plan_create_large(size_t n, int method, int algo, double measure_seconds,
                                   cfft_plan *out) {
  if (!out) return set_error(CFFT_ERROR_NULL_POINTER, "out is null");
  *out = nullptr;
  if (!is_pow2(n)) return set_error(CFFT_ERROR_NOT_POWER_OF_TWO, "assertion failed: n.is_power_of_two()");
  // the range the reference documents for the ordered plan (src/ordered.rs:185)
  if (ilog2(n) > 16) return set_error(CFFT_ERROR_SIZE, "ordered plans run up to n = 2^16");
  size_t base = n;
  if (method == CFFT_METHOD_MEASURE) {
    int rc = measure(cfft_plan_s::ORDERED, n, measure_seconds, &algo, &base);
    if (rc) return rc;
  } else if (method != CFFT_METHOD_USER_PROVIDED) {
    return set_error(CFFT_ERROR_INVALID_ARGUMENT, "unknown method");
  }
  if (algo_radix(algo) == 0) return set_error(CFFT_ERROR_INVALID_ALGO, "unknown FftAlgo");
  return create_f64(cfft_plan_s::ORDERED, n, algo, n, out);
}

int cfft_unordered_plan_create(size_t n, int method, int base_algo, size_t base_n,
                               double measure_seconds, cfft_plan *out) {
  if (!out) return set_error(CFFT_ERROR_NULL_POINTER, "out is null");
  *out = nullptr;
  if (!is_pow2(n)) return set_error(CFFT_ERROR_NOT_POWER_OF_TWO, "assertion failed: n.is_power_of_two()");
  if (ilog2(n) > 30) return set_error(CFFT_ERROR_SIZE, "fft size too large for the device plan");
  if (method == CFFT_METHOD_MEASURE) {
    int rc = measure(cfft_plan_s::UNORDERED, n, measure_seconds, &base_algo, &base_n);
    if (rc) return rc;
  } else if (method == CFFT_METHOD_USER_PROVIDED) {
    if (!is_pow2(base_n)) return set_error(CFFT_ERROR_INVALID_BASE, "assertion failed: base_n.is_power_of_two()");
    if (base_n > n) return set_error(CFFT_ERROR_INVALID_BASE, "assertion failed: base_n <= n");
    if (base_n != n && base_n < 32) return set_error(CFFT_ERROR_INVALID_BASE, "assertion failed: base_n >= 32");
    if (ilog2(base_n) > 10) return set_error(CFFT_ERROR_INVALID_BASE, "assertion failed: base_n.trailing_zeros() <= 10");
  } else {
    return set_error(CFFT_ERROR_INVALID_ARGUMENT, "unknown method");
  }
  if (algo_radix(base_algo) == 0) return set_error(CFFT_ERROR_INVALID_ALGO, "unknown FftAlgo");
  return create_f64(cfft_plan_s::UNORDERED, n, base_algo, base_n, out);
}

int cfft_plan_destroy(cfft_plan plan) {
  delete plan;
  return CFFT_SUCCESS;
}

int cfft_plan_fft_size(cfft_plan plan, size_t *n) {
  if (!plan || !n) return set_error(CFFT_ERROR_NULL_POINTER, "null argument");
  *n = plan->n;
  return CFFT_SUCCESS;
}
int cfft_plan_algo(cfft_plan plan, int *algo, size_t *base_n) {
  if (!plan) return set_error(CFFT_ERROR_NULL_POINTER, "plan is null");
  if (plan->kind == cfft_plan_s::FFT128)
    return set_error(CFFT_ERROR_WRONG_PLAN_KIND, "fft128 plans have no algorithm choice");
  if (algo) *algo = plan->algo;
  if (base_n) *base_n = plan->base_n;
  return CFFT_SUCCESS;
}
int cfft_plan_fft_scratch(cfft_plan plan, size_t *bytes, size_t *align) {
  if (!plan) return set_error(CFFT_ERROR_NULL_POINTER, "plan is null");
  const size_t elems = plan->kind == cfft_plan_s::ORDERED ? plan->n
                       : plan->kind == cfft_plan_s::UNORDERED ? plan->base_n
                                                               : 0;
  if (bytes) *bytes = elems * sizeof(dc64);
  if (align) *align = 128;  // aligned_vec::CACHELINE_ALIGN on x86-64
  return CFFT_SUCCESS;
}
int cfft_plan_permutation(cfft_plan plan, size_t *idx) {
  if (!plan || !idx) return set_error(CFFT_ERROR_NULL_POINTER, "null argument");
  if (plan->kind == cfft_plan_s::FFT128)
    return set_error(CFFT_ERROR_WRONG_PLAN_KIND, "fft128 output order is bit-reversed");
  for (size_t i = 0; i < plan->n; ++i) idx[i] = plan->pos_of_freq[i];
  return CFFT_SUCCESS;
}

// ---- f64 transforms ----
int cfft_fwd(cfft_plan plan, void *buf, size_t batch, size_t stride, cfft_stream_t stream) {
  return run_transform(plan, buf, batch, stride, true, (hipStream_t)stream);
}
int cfft_inv(cfft_plan plan, void *buf, size_t batch, size_t stride, cfft_stream_t stream) {
  return run_transform(plan, buf, batch, stride, false, (hipStream_t)stream);
}
int cfft_fwd_host(cfft_plan plan, double *buf, size_t batch) { return run_transform_host(plan, buf, batch, true); }
int cfft_inv_host(cfft_plan plan, double *buf, size_t batch) { return run_transform_host(plan, buf, batch, false); }

int cfft_fwd_monomial(cfft_plan plan, size_t degree, void *buf, cfft_stream_t stream) {
  if (!plan || !buf) return set_error(CFFT_ERROR_NULL_POINTER, "null argument");
  if (plan->kind != cfft_plan_s::UNORDERED)
    return set_error(CFFT_ERROR_WRONG_PLAN_KIND, "fwd_monomial is a method of unordered::Plan");
  if (degree >= plan->n) return set_error(CFFT_ERROR_INVALID_ARGUMENT, "assertion failed: degree < this.fft_size()");
  return hip_error(cfft_launch::monomial(reinterpret_cast<dc64 *>(buf), ilog2(plan->n), degree,
                                         plan->d_freq, plan->d_tw, (hipStream_t)stream),
                   "monomial");
}
int cfft_fwd_monomial_host(cfft_plan plan, size_t degree, double *buf) {
  if (!plan || !buf) return set_error(CFFT_ERROR_NULL_POINTER, "null argument");
  const size_t per = plan->n * sizeof(dc64);
  return cfft_internal::stage_host(
      1, per, [](void *, size_t, size_t) {},  // the kernel writes every point
      [&](void *d, size_t, hipStream_t s) { return cfft_fwd_monomial(plan, degree, d, s); },
      [&](const void *pin, size_t, size_t) { std::memcpy(buf, pin, per); });
}

// ---- serde ----
int cfft_fourier_to_standard(cfft_plan plan, const void *permuted, void *standard, size_t batch,
                             cfft_stream_t stream) {
  if (!plan || !permuted || !standard) return set_error(CFFT_ERROR_NULL_POINTER, "null argument");
  if (plan->kind == cfft_plan_s::FFT128) return set_error(CFFT_ERROR_WRONG_PLAN_KIND, "f64 plans only");
  return hip_error(cfft_launch::gather(reinterpret_cast<const dc64 *>(permuted),
                                       reinterpret_cast<dc64 *>(standard), plan->d_pos,
                                       (uint32_t)plan->n, batch, (hipStream_t)stream),
                   "gather");
}
int cfft_fourier_from_standard(cfft_plan plan, const void *standard, void *permuted,
                               size_t batch, cfft_stream_t stream) {
  if (!plan || !permuted || !standard) return set_error(CFFT_ERROR_NULL_POINTER, "null argument");
  if (plan->kind == cfft_plan_s::FFT128) return set_error(CFFT_ERROR_WRONG_PLAN_KIND, "f64 plans only");
  return hip_error(cfft_launch::scatter(reinterpret_cast<const dc64 *>(standard),
                                        reinterpret_cast<dc64 *>(permuted), plan->d_pos,
                                        (uint32_t)plan->n, batch, (hipStream_t)stream),
                   "scatter");
}
int cfft_serialize_fourier_buffer(cfft_plan plan, const double *buf, uint8_t *out, size_t out_cap,
                                  size_t *out_len) {
  if (!plan || !buf || !out_len) return set_error(CFFT_ERROR_NULL_POINTER, "null argument");
  if (plan->kind == cfft_plan_s::FFT128) return set_error(CFFT_ERROR_WRONG_PLAN_KIND, "f64 plans only");
  std::vector<uint8_t> o;
  put_varint(o, plan->n);
  const size_t need = o.size() + plan->n * 16;
  *out_len = need;
  if (!out) return CFFT_SUCCESS;
  if (out_cap < need) return set_error(CFFT_ERROR_SIZE_OVERFLOW, "output buffer too small");
  std::memcpy(out, o.data(), o.size());
  uint8_t *w = out + o.size();
  for (size_t i = 0; i < plan->n; ++i) {
    const double *z = buf + 2 * (size_t)plan->pos_of_freq[i];
    std::memcpy(w, z, 16);  // (re, im) little-endian f64
    w += 16;
  }
  return CFFT_SUCCESS;
}
int cfft_deserialize_fourier_buffer(cfft_plan plan, const uint8_t *bytes, size_t len, double *buf) {
  if (!plan || !bytes || !buf) return set_error(CFFT_ERROR_NULL_POINTER, "null argument");
  if (plan->kind == cfft_plan_s::FFT128) return set_error(CFFT_ERROR_WRONG_PLAN_KIND, "f64 plans only");
  if (len < 1) return set_error(CFFT_ERROR_DESERIALIZE, "unexpected end of input");
  uint64_t count = 0;
  size_t off = 0;
  const uint8_t tag = bytes[0];
  int w = tag < 251 ? 0 : tag == 251 ? 2 : tag == 252 ? 4 : tag == 253 ? 8 : -1;
  if (w < 0) return set_error(CFFT_ERROR_DESERIALIZE, "invalid varint tag");
  if (w == 0) {
    count = tag;
    off = 1;
  } else {
    if (len < (size_t)(1 + w)) return set_error(CFFT_ERROR_DESERIALIZE, "unexpected end of input");
    for (int i = 0; i < w; ++i) count |= (uint64_t)bytes[1 + i] << (8 * i);
    off = 1 + (size_t)w;
  }
  if (len < off + count * 16) return set_error(CFFT_ERROR_DESERIALIZE, "unexpected end of input");
  if (count != plan->n) {
    char m[160];
    snprintf(m, sizeof m, "invalid length %llu, expected a sequence of %zu 64-bit complex numbers",
             (unsigned long long)count, plan->n);
    return set_error(CFFT_ERROR_DESERIALIZE, m);
  }
  for (size_t i = 0; i < plan->n; ++i)
    std::memcpy(buf + 2 * (size_t)plan->pos_of_freq[i], bytes + off + 16 * i, 16);
  return CFFT_SUCCESS;
}

// ---- fused negacyclic product ----
int cfft_negacyclic_polymul_u64(cfft_plan plan, const uint64_t *a, const int64_t *b, uint64_t *c,
                                size_t batch, cfft_stream_t stream) {
  if (!plan) return set_error(CFFT_ERROR_NULL_POINTER, "plan is null");
  if (plan->kind != cfft_plan_s::UNORDERED)
    return set_error(CFFT_ERROR_WRONG_PLAN_KIND, "the negacyclic product runs on an unordered plan");
  if (batch == 0) return CFFT_SUCCESS;
  if (!a || !b || !c) return set_error(CFFT_ERROR_NULL_POINTER, "null buffer");
  if (plan->n < 2 || plan->fwd_full.npass == 0)
    return set_error(CFFT_ERROR_SIZE, "negacyclic product needs 2 <= fft_size <= 8192");
  return cfft_polymul_dispatch(plan, a, b, c, batch, (hipStream_t)stream);
}

}  // extern "C"

// generic fused product (the specialised kernels hook in here)
int cfft_polymul_dispatch(const cfft_plan_s *p, const uint64_t *a, const int64_t *b, uint64_t *c,
                          size_t batch, hipStream_t s) {
  const uint32_t log2n = ilog2(p->n);
  if ((p->n == 256 || p->n == 512) && p->base_n == p->n && p->d_tw_base1024)  // Dif16
    return hip_error(cfft_launch::wsub_polymul(log2n, a, b, c, batch, p->d_tw_base1024,
                                               p->d_twist, s),
                     "wsub_polymul");
  if (p->algo == CFFT_DIF16 && p->base_n == p->n && p->n == 1024 && p->d_tw_r16)
    return hip_error(cfft_launch::r16_polymul(log2n, a, b, c, batch, p->d_tw_r16, p->d_twist, s),
                     "r16_polymul");
  if (p->algo == CFFT_DIF16 && p->base_n == 1024 && p->n >= 2048 && p->n <= 8192 &&
      p->d_tw_base1024)
    return hip_error(cfft_launch::depth_polymul(log2n, a, b, c, batch, p->d_tw, p->d_tw_base1024,
                                                p->d_twist, p->d_untwist, s),
                     "depth_polymul");
  if (p->d_tw_r8)
    return hip_error(cfft_launch::r8_polymul(a, b, c, batch, p->d_tw_r8, p->d_twist, s),
                     "r8_polymul");
  if (p->d_tw_dit16)
    return hip_error(cfft_launch::dit16_polymul(a, b, c, batch, p->d_tw_dit16, p->d_twist, s),
                     "dit16_polymul");
  if (p->n == 1024 && p->base_n == 512 && p->d_tw_base1024)  // Dif16, base_n = 512
    return hip_error(cfft_launch::wd2_512_polymul(a, b, c, batch, p->d_tw, p->d_tw_base1024,
                                                  p->d_twist, s),
                     "wd2_512_polymul");
  return hip_error(cfft_launch::polymul_generic(p->fwd_full, p->inv_full, a, b, c, batch, p->d_tw,
                                                p->d_twist, p->d_untwist, s),
                   "polymul");
}
