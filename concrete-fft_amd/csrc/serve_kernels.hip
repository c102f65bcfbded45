// serve_kernels.hip -- row gather and scatter for the request server's
// device-buffer requests: a batch's inputs are copied from the callers'
// device buffers into the batch's contiguous block, and its results back,
// on the device.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "launch.hpp"

namespace {

constexpr uint32_t kRowThreads = 256;

// dst + r * row_words <- table[r] (row r of the batch), row_words u64 each
__global__ void __launch_bounds__(kRowThreads) k_gather_rows(const uint64_t *__restrict__ table,
                                                             uint64_t *__restrict__ dst,
                                                             uint64_t row_words) {
  const uint64_t r = blockIdx.y;
  const uint64_t *src = reinterpret_cast<const uint64_t *>(table[r]);
  uint64_t *d = dst + r * row_words;
  for (uint64_t i = (uint64_t)blockIdx.x * kRowThreads + threadIdx.x; i < row_words;
       i += (uint64_t)gridDim.x * kRowThreads)
    d[i] = src[i];
}

// table[r] <- src + r * row_words
__global__ void __launch_bounds__(kRowThreads) k_scatter_rows(const uint64_t *__restrict__ table,
                                                              const uint64_t *__restrict__ src,
                                                              uint64_t row_words) {
  const uint64_t r = blockIdx.y;
  uint64_t *d = reinterpret_cast<uint64_t *>(table[r]);
  const uint64_t *s = src + r * row_words;
  for (uint64_t i = (uint64_t)blockIdx.x * kRowThreads + threadIdx.x; i < row_words;
       i += (uint64_t)gridDim.x * kRowThreads)
    d[i] = s[i];
}

dim3 row_grid(uint64_t rows, uint64_t row_words) {
  // up to 8 words per thread per row, at most 64 workgroups per row
  uint64_t x = (row_words + 8 * kRowThreads - 1) / (8 * kRowThreads);
  if (x > 64) x = 64;
  if (x < 1) x = 1;
  return dim3((uint32_t)x, (uint32_t)rows);
}

}  // namespace

namespace cfft_launch {

hipError_t gather_rows(const uint64_t *table, uint64_t rows, uint64_t row_words, uint64_t *dst,
                       hipStream_t s) {
  if (rows == 0 || row_words == 0) return hipSuccess;
  if (rows > 65535) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_gather_rows, row_grid(rows, row_words), dim3(kRowThreads), 0, s, table,
                     dst, row_words);
  return hipGetLastError();
}

hipError_t scatter_rows(const uint64_t *table, uint64_t rows, uint64_t row_words,
                        const uint64_t *src, hipStream_t s) {
  if (rows == 0 || row_words == 0) return hipSuccess;
  if (rows > 65535) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_scatter_rows, row_grid(rows, row_words), dim3(kRowThreads), 0, s, table,
                     src, row_words);
  return hipGetLastError();
}

}  // namespace cfft_launch
