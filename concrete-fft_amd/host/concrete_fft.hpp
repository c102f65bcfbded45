// concrete_fft.hpp -- C++ mirror of concrete-fft's public API (v0.5.1) over
// the C ABI of libcfft_amd (include/cfft.h).  Header-only.
//
//   concrete_fft::ordered::Plan    src/ordered.rs:186-373
//   concrete_fft::unordered::Plan  src/unordered.rs:496-1036
//   concrete_fft::fft128::Plan     src/fft128/mod.rs:1830-1961
//
// Where the reference panics, these throw concrete_fft::Panic with the same
// message.  Buffers are std::complex<double> spans (the c64 layout).  Host
// overloads copy through the device and synchronise; device overloads take a
// device pointer, a batch size and a stream.
#pragma once

#include <complex>
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "cfft.h"

namespace concrete_fft {

using c64 = std::complex<double>;

struct Panic : std::runtime_error {
  int status;
  Panic(int st, const std::string &m) : std::runtime_error(m), status(st) {}
};

inline void check(int st) {
  if (st != CFFT_SUCCESS) {
    const char *m = cfft_last_error_message();
    throw Panic(st, (m && *m) ? m : cfft_status_string(st));
  }
}

enum class FftAlgo { Dif2 = 0, Dit2, Dif4, Dit4, Dif8, Dit8, Dif16, Dit16 };

struct StackReq {
  size_t size_bytes, align_bytes;
};

namespace detail {
class PlanHandle {
 public:
  PlanHandle() = default;
  explicit PlanHandle(cfft_plan p) : p_(p) {}
  PlanHandle(const PlanHandle &) = delete;
  PlanHandle &operator=(const PlanHandle &) = delete;
  PlanHandle(PlanHandle &&o) noexcept : p_(std::exchange(o.p_, nullptr)) {}
  PlanHandle &operator=(PlanHandle &&o) noexcept {
    std::swap(p_, o.p_);
    return *this;
  }
  ~PlanHandle() {
    if (p_) cfft_plan_destroy(p_);
  }
  cfft_plan get() const { return p_; }

 private:
  cfft_plan p_ = nullptr;
};

inline size_t fft_size(cfft_plan p) {
  size_t n = 0;
  check(cfft_plan_fft_size(p, &n));
  return n;
}
inline void length_check(size_t expected, size_t got) {
  if (expected != got)
    throw Panic(CFFT_ERROR_LENGTH_MISMATCH, "assertion `left == right` failed\n  left: " +
                                                std::to_string(expected) +
                                                "\n right: " + std::to_string(got));
}
}  // namespace detail

namespace ordered {
using concrete_fft::FftAlgo;
struct Method {
  int kind;
  FftAlgo algo;
  double seconds;
  static Method user_provided(FftAlgo a) { return {CFFT_METHOD_USER_PROVIDED, a, 0.0}; }
  static Method measure(double seconds) { return {CFFT_METHOD_MEASURE, FftAlgo::Dif4, seconds}; }
};

class Plan {
 public:
  Plan(size_t n, Method m) {
    cfft_plan p = nullptr;
    check(cfft_ordered_plan_create(n, m.kind, (int)m.algo, m.seconds, &p));
    h_ = detail::PlanHandle(p);
  }
  // 2^11 <= n <= 2^13 as well: the reference's ordered recursion beyond its
  // n < 2^11 assertion (cfft_ordered_plan_create_large)
  static Plan large(size_t n, Method m) {
    cfft_plan p = nullptr;
    check(cfft_ordered_plan_create_large(n, m.kind, (int)m.algo, m.seconds, &p));
    return Plan(detail::PlanHandle(p));
  }
  size_t fft_size() const { return detail::fft_size(h_.get()); }
  FftAlgo algo() const {
    int a = 0;
    size_t b = 0;
    check(cfft_plan_algo(h_.get(), &a, &b));
    return (FftAlgo)a;
  }
  StackReq fft_scratch() const {
    StackReq r{};
    check(cfft_plan_fft_scratch(h_.get(), &r.size_bytes, &r.align_bytes));
    return r;
  }
  void fwd(std::vector<c64> &buf) const {
    detail::length_check(fft_size(), buf.size());
    check(cfft_fwd_host(h_.get(), reinterpret_cast<double *>(buf.data()), 1));
  }
  void inv(std::vector<c64> &buf) const {
    detail::length_check(fft_size(), buf.size());
    check(cfft_inv_host(h_.get(), reinterpret_cast<double *>(buf.data()), 1));
  }
  void fwd(void *device_buf, size_t batch, cfft_stream_t stream) const {
    check(cfft_fwd(h_.get(), device_buf, batch, fft_size(), stream));
  }
  void inv(void *device_buf, size_t batch, cfft_stream_t stream) const {
    check(cfft_inv(h_.get(), device_buf, batch, fft_size(), stream));
  }
  cfft_plan raw() const { return h_.get(); }

 private:
  explicit Plan(detail::PlanHandle h) : h_(std::move(h)) {}
  detail::PlanHandle h_;
};
}  // namespace ordered

namespace unordered {
using concrete_fft::FftAlgo;
struct Method {
  int kind;
  FftAlgo base_algo;
  size_t base_n;
  double seconds;
  static Method user_provided(FftAlgo a, size_t base_n) {
    return {CFFT_METHOD_USER_PROVIDED, a, base_n, 0.0};
  }
  static Method measure(double seconds) { return {CFFT_METHOD_MEASURE, FftAlgo::Dif4, 0, seconds}; }
};

class Plan {
 public:
  Plan(size_t n, Method m) {
    cfft_plan p = nullptr;
    check(cfft_unordered_plan_create(n, m.kind, (int)m.base_algo, m.base_n, m.seconds, &p));
    h_ = detail::PlanHandle(p);
  }
  size_t fft_size() const { return detail::fft_size(h_.get()); }
  std::pair<FftAlgo, size_t> algo() const {
    int a = 0;
    size_t b = 0;
    check(cfft_plan_algo(h_.get(), &a, &b));
    return {(FftAlgo)a, b};
  }
  StackReq fft_scratch() const {
    StackReq r{};
    check(cfft_plan_fft_scratch(h_.get(), &r.size_bytes, &r.align_bytes));
    return r;
  }
  void fwd(std::vector<c64> &buf) const {
    detail::length_check(fft_size(), buf.size());
    check(cfft_fwd_host(h_.get(), reinterpret_cast<double *>(buf.data()), 1));
  }
  void inv(std::vector<c64> &buf) const {
    detail::length_check(fft_size(), buf.size());
    check(cfft_inv_host(h_.get(), reinterpret_cast<double *>(buf.data()), 1));
  }
  void fwd(void *device_buf, size_t batch, cfft_stream_t stream) const {
    check(cfft_fwd(h_.get(), device_buf, batch, fft_size(), stream));
  }
  void inv(void *device_buf, size_t batch, cfft_stream_t stream) const {
    check(cfft_inv(h_.get(), device_buf, batch, fft_size(), stream));
  }
  void fwd_monomial(size_t degree, std::vector<c64> &buf) const {
    detail::length_check(fft_size(), buf.size());
    check(cfft_fwd_monomial_host(h_.get(), degree, reinterpret_cast<double *>(buf.data())));
  }
  std::vector<uint8_t> serialize_fourier_buffer(const std::vector<c64> &buf) const {
    detail::length_check(fft_size(), buf.size());
    size_t len = 0;
    const double *p = reinterpret_cast<const double *>(buf.data());
    check(cfft_serialize_fourier_buffer(h_.get(), p, nullptr, 0, &len));
    std::vector<uint8_t> out(len);
    check(cfft_serialize_fourier_buffer(h_.get(), p, out.data(), out.size(), &len));
    return out;
  }
  void deserialize_fourier_buffer(const std::vector<uint8_t> &bytes, std::vector<c64> &buf) const {
    detail::length_check(fft_size(), buf.size());
    check(cfft_deserialize_fourier_buffer(h_.get(), bytes.data(), bytes.size(),
                                          reinterpret_cast<double *>(buf.data())));
  }
  cfft_plan raw() const { return h_.get(); }

 private:
  detail::PlanHandle h_;
};
}  // namespace unordered

namespace fft128 {
class Plan {
 public:
  explicit Plan(size_t n) {
    cfft_plan p = nullptr;
    check(cfft128_plan_create(n, &p));
    h_ = detail::PlanHandle(p);
  }
  size_t fft_size() const { return detail::fft_size(h_.get()); }
  void fwd(std::vector<double> &re0, std::vector<double> &re1, std::vector<double> &im0,
           std::vector<double> &im1) const {
    lengths(re0, re1, im0, im1);
    check(cfft128_fwd_host(h_.get(), re0.data(), re1.data(), im0.data(), im1.data(), 1));
  }
  void inv(std::vector<double> &re0, std::vector<double> &re1, std::vector<double> &im0,
           std::vector<double> &im1) const {
    lengths(re0, re1, im0, im1);
    check(cfft128_inv_host(h_.get(), re0.data(), re1.data(), im0.data(), im1.data(), 1));
  }
  cfft_plan raw() const { return h_.get(); }

 private:
  void lengths(const std::vector<double> &a, const std::vector<double> &b,
               const std::vector<double> &c, const std::vector<double> &d) const {
    const size_t n = fft_size();
    detail::length_check(n, a.size());
    detail::length_check(n, b.size());
    detail::length_check(n, c.size());
    detail::length_check(n, d.size());
  }
  detail::PlanHandle h_;
};
using Plan128 = Plan;
}  // namespace fft128

}  // namespace concrete_fft
