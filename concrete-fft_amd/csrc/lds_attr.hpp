// lds_attr.hpp -- launch helpers shared by the kernel translation units.
//
// ensure_lds: raise a kernel's dynamic-LDS limit above the 64 KiB default,
// per (device, kernel), to the largest size requested so far (not to the
// device maximum: a kernel's static LDS counts against the same budget).
// Thread-safe; host code shared by the kernel translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <utility>

namespace cfft_launch {

inline hipError_t ensure_lds(const void *kernel, size_t bytes) {
  if (bytes <= 64 * 1024) return hipSuccess;
  static std::mutex mu;
  static std::map<std::pair<int, const void *>, size_t> set_to;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  std::lock_guard<std::mutex> g(mu);
  auto it = set_to.find({dev, kernel});
  if (it != set_to.end() && it->second >= bytes) return hipSuccess;
  int maxb = 0;
  e = hipDeviceGetAttribute(&maxb, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
  if (e != hipSuccess) return e;
  if ((size_t)maxb < bytes) return hipErrorInvalidValue;
  e = hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  if (e != hipSuccess) return e;
  set_to[{dev, kernel}] = bytes;
  return hipSuccess;
}

// Compute units of the current device (read once per device), for the grids
// of persistent kernels.
inline hipError_t cu_count(int *cus) {
  static std::mutex mu;
  static std::map<int, int> of;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  std::lock_guard<std::mutex> g(mu);
  auto it = of.find(dev);
  if (it == of.end()) {
    int v = 0;
    e = hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return e;
    if (v <= 0) return hipErrorInvalidDevice;
    it = of.emplace(dev, v).first;
  }
  *cus = it->second;
  return hipSuccess;
}

}  // namespace cfft_launch
