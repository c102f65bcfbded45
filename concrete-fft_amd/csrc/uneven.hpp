// uneven.hpp -- device-side records of an uneven-batch schedule (shared by
// the host runtime in uneven_api.cpp and the kernels in uneven_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "launch.hpp"
#include "program.hpp"

namespace cfft_ub {

constexpr uint32_t kThreads = 512;  // one workgroup: 8 waves
// LDS of one workgroup: (8192 / n) transform groups of n + n/16 points
constexpr uint32_t kLdsPoints = 8192 + 512;
constexpr size_t kLdsBytes = kLdsPoints * 16;
// Dif16 body: tile data followed by the twiddle block of the 1024-point
// base transform (rows k < 16 of T[k l], l < 64, then rows of T[16 k p],
// p < 4); every Dif16 class reads its base twiddles from it (a class with
// base_n < 1024 at column stride 1024 / base_n)
constexpr uint32_t kTwR16Points = 16 * 64 + 16 * 4;
constexpr size_t kR16LdsBytes = (size_t)(kLdsPoints + kTwR16Points) * 16;
// device work queue of the Dif16 body: partitions (one counter each, on its
// own 128-byte line) and the counter buffer's layout
constexpr uint32_t kQueueParts = 8;
constexpr uint32_t kCounterStride = 32;  // words
constexpr uint32_t kGenericCounter = kQueueParts * kCounterStride;
constexpr size_t kCounterBytes = (kGenericCounter + kCounterStride) * 4;
constexpr uint32_t kMinLog2N = 5;   // N = 32: n = 16, one thread per transform
constexpr uint32_t kMaxLog2N = 14;  // N = 16384: n = 8192, one transform per workgroup

struct UJob {
  uint64_t in_off;   // doubles
  uint64_t f_off;    // c64
  uint64_t out_off;  // doubles
  uint32_t log2n;    // complex length n = N / 2
  uint32_t pad;
};

// `count` jobs order[first .. first + count) of class `cls`, one per group
struct UTile {
  uint32_t cls;
  uint32_t first;
  uint32_t count;
  uint32_t pad;
};

struct UClass {
  cfft_prog::Program fwd, inv;  // whole-transform programs of the class plan
  const cfft_launch::dc64 *tw;  // T[j], n entries
  const cfft_launch::dc64 *twist;
  const cfft_launch::dc64 *untwist;
  uint32_t log2n;
  uint32_t r16;  // 1: the class runs the Dif16 body
};

// classes the Dif16 body runs: Dif16 plan, base_n = min(n, 1024), 8 <= log2n <= 13
constexpr bool r16_class(bool dif16, uint32_t log2n, uint32_t log2_base) {
  return dif16 && log2n >= 8 && log2n <= 13 && log2_base == (log2n < 10 ? log2n : 10);
}

// groups (transform groups of n/16 threads) per tile for a class of
// complex length 2^log2n, and jobs per tile: one per group, jobs_per_group
// for the wave-local Dif16 classes (n <= 1024)
constexpr uint32_t groups_per_tile(uint32_t log2n) { return kThreads / ((1u << log2n) / 16); }
// jobs a group of the wave-local Dif16 body transforms together: two at
// n = 1024 (a group is a whole wavefront), one below (two would need more
// than 256 VGPRs beside the per-lane job pointers of sub-wave groups)
constexpr uint32_t jobs_per_group(uint32_t log2n) { return log2n >= 10 ? 2u : 1u; }
constexpr uint32_t jobs_per_tile(uint32_t log2n, bool r16) {
  // a Dif16 class at n = 2048 runs one job per wavefront (8 per tile)
  return r16 && log2n == 11 ? kThreads / 64
                            : groups_per_tile(log2n) * (r16 && log2n <= 10 ? jobs_per_group(log2n) : 1u);
}

// sets the kernels' LDS limit on the current device (before any launch)
hipError_t prepare();
// Generic body: resets counter[0], then runs tiles[0 .. ntiles) on `grid`
// persistent workgroups.
hipError_t launch_persistent(const UJob *jobs, const uint32_t *order, const UTile *tiles,
                             uint32_t ntiles, const UClass *classes, uint32_t *counter,
                             uint32_t grid, const double *in, cfft_launch::dc64 *fourier,
                             double *out, hipStream_t s);
// Dif16 body: runs tiles[0 .. ntiles) on `grid` persistent workgroups, fed
// by the partitioned device queue at `counter` (kQueueParts counters, reset
// by a kernel launched first) or, with counter null, dealt round-robin.
hipError_t launch_r16_persistent(const UJob *jobs, const uint32_t *order, const UTile *tiles,
                                 uint32_t ntiles, const UClass *classes,
                                 const cfft_launch::dc64 *tw_block, uint32_t *counter,
                                 uint32_t grid, const double *in, cfft_launch::dc64 *fourier,
                                 double *out, hipStream_t s);
// Plain form: one workgroup per tile; `tw_block` non-null selects the Dif16
// body (every tile's class must be a Dif16 class).
hipError_t launch_tiles(const UJob *jobs, const uint32_t *order, const UTile *tiles,
                        uint32_t first_tile, uint32_t count, const UClass *classes,
                        const cfft_launch::dc64 *tw_block, const double *in,
                        cfft_launch::dc64 *fourier, double *out, hipStream_t s);

}  // namespace cfft_ub
