// plan.hpp -- internal plan object behind the opaque cfft_plan handle.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "launch.hpp"
#include "program.hpp"

struct cfft_plan_s {
  enum Kind { ORDERED = 0, UNORDERED = 1, FFT128 = 2 };
  Kind kind = UNORDERED;
  size_t n = 0;       // fft_size
  size_t base_n = 0;  // internal ordered size (== n for ordered plans)
  int algo = 0;       // (base) algorithm, cfft_algo
  int device = 0;

  // f64 path
  cfft_launch::dc64 *d_tw = nullptr;  // T[j] = (cos, sin)(-2 pi j / n)
  uint32_t *d_freq = nullptr;         // position -> frequency
  uint32_t *d_pos = nullptr;          // frequency -> position
  std::vector<uint32_t> pos_of_freq;  // host copy of d_pos
  // LDS program and the depth passes too large for one workgroup
  cfft_prog::Program fwd_lds{}, inv_lds{};
  uint32_t chunk_log2_fwd = 0, chunk_log2_inv = 0;
  std::vector<cfft_prog::Pass> fwd_global, inv_global;  // fwd: before, inv: after
  // whole-transform programs (n <= LDS chunk) used by the fused kernels
  cfft_prog::Program fwd_full{}, inv_full{};
  // r16 (register-resident Dif16) twiddle block: T[k*l] as [k][l] (16 x n/16)
  // then T[16 k p] as [k][p] (16 x n/256)
  cfft_launch::dc64 *d_tw_r16 = nullptr;
  // Dit16 plans with n = base_n = 1024: the same block, for the
  // register-resident Dit16 kernels
  cfft_launch::dc64 *d_tw_dit16 = nullptr;
  // Dif16 plans with base_n = 1024 < n <= 8192, and with n = base_n = 512:
  // the r16 twiddle block of a 1024-point transform (T[k l n/1024] and
  // T[16 k p n/1024], the same values as a 1024-point plan's d_tw_r16; at
  // n = 512 read at column stride 2), for the wave-local kernels
  cfft_launch::dc64 *d_tw_base1024 = nullptr;
  // Dif16 plans with base_n = 1024 and n > 8192: the twiddles of the last
  // depth level above the base, T_c[i] = T[i n / c] for the chunk of c =
  // 2^chunk_log2_fwd points (2048, 4096 or 8192), so that the chunks run on
  // the wave-local kernels after the global depth passes
  cfft_launch::dc64 *d_tw_chunk = nullptr;
  // r8 (register-resident Dif8, n = 1024) twiddle block: T[k*l] as [k][l]
  // (8 x 128), T[8 k p] as [k][p] (8 x 16), T[64 k p] as [k][p] (8 x 2)
  cfft_launch::dc64 *d_tw_r8 = nullptr;
  // negacyclic product tables (N = 2n)
  cfft_launch::dc64 *d_twist = nullptr, *d_untwist = nullptr;

  // fft128
  double *d_tw128 = nullptr;  // 4 x n: re0 | re1 | im0 | im1

  ~cfft_plan_s();
};

namespace cfft_internal {
int set_error(int status, const std::string &msg);
int hip_error(hipError_t e, const char *where);

// Fourier scratch the split external product of `batch` GLWEs needs: the
// spectra of one pass (large batches run in passes over the same scratch).
size_t external_product_scratch_bytes(const cfft_plan_s *plan, int k1, int levels, size_t batch);
// cfft_external_product with caller-owned scratch: `scratch` is used when it
// holds external_product_scratch_bytes(...) bytes, else the scratch cache.
// With no_alloc the batch runs in passes of as many GLWEs as `scratch` holds
// and nothing is allocated (CFFT_ERROR_SIZE if one GLWE does not fit).
// glwe_tab / out_tab (optional, device-readable, one u64 address per GLWE of
// the batch) replace the contiguous glwe / out; they need a plan for which
// external_product_tables_supported holds.
int external_product_with_scratch(cfft_plan_s *plan, const void *key_fourier, int k1, int levels,
                                  int base_log, const uint64_t *glwe, uint64_t *out,
                                  size_t batch, hipStream_t stream, void *scratch,
                                  size_t scratch_bytes, bool no_alloc,
                                  const uint64_t *glwe_tab = nullptr,
                                  const uint64_t *out_tab = nullptr);
// the plan runs the wave-local split kernels (Dif16, n = 512, 1024, 2048),
// which read and write GLWEs through address tables
bool external_product_tables_supported(const cfft_plan_s *plan);
}  // namespace cfft_internal
