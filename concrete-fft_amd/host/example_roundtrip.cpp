// example_roundtrip.cpp -- the reference's doctest / test_roundtrip relations
// (src/lib.rs:35-64, src/unordered.rs:1141-1172) written against the C++
// mirror.  Exits non-zero on any failure; prints one line per check.
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#include "concrete_fft.hpp"

using namespace concrete_fft;

static double max_err(const std::vector<c64> &a, const std::vector<c64> &b, double scale) {
  double e = 0;
  for (size_t i = 0; i < a.size(); ++i) e = std::fmax(e, std::abs(a[i] * scale - b[i]));
  return e;
}

int main() {
  std::mt19937_64 rng(7);
  std::uniform_real_distribution<double> u(0.0, 1.0);
  int failures = 0;
  // doctest of src/lib.rs:35-64: n = 4 ordered roundtrip
  {
    ordered::Plan plan(4, ordered::Method::user_provided(FftAlgo::Dif4));
    std::vector<c64> data = {{1, 0}, {2, 0}, {3, 0}, {4, 0}}, x = data;
    plan.fwd(x);
    plan.inv(x);
    const double e = max_err(x, data, 0.25);
    std::printf("ordered n=4 roundtrip err %.3g\n", e);
    failures += e >= 1e-9;
  }
  // an ordered plan beyond the reference's n < 2^11 assertion
  {
    const size_t n = 4096;
    auto plan = ordered::Plan::large(n, ordered::Method::user_provided(FftAlgo::Dif16));
    std::vector<c64> z(n);
    for (auto &v : z) v = c64(u(rng), u(rng));
    std::vector<c64> orig = z;
    plan.fwd(z);
    plan.inv(z);
    const double e = max_err(z, orig, 1.0 / (double)n);
    std::printf("ordered n=%zu (large) roundtrip err %.3g\n", n, e);
    failures += e >= 1e-12;
  }
  // test_roundtrip over unordered plans
  for (size_t n : {32, 64, 256, 512, 1024}) {
    unordered::Plan plan(n, unordered::Method::user_provided(FftAlgo::Dif4, 32));
    std::vector<c64> z(n);
    for (auto &v : z) v = c64(u(rng), u(rng));
    std::vector<c64> orig = z;
    plan.fwd(z);
    plan.inv(z);
    const double e = max_err(z, orig, 1.0 / (double)n);
    std::printf("unordered n=%zu roundtrip err %.3g\n", n, e);
    failures += e >= 1e-12;
  }
  // the error path: a non power of two panics with the reference's message
  try {
    unordered::Plan bad(48, unordered::Method::user_provided(FftAlgo::Dif4, 32));
    failures += 1;
  } catch (const Panic &p) {
    std::printf("panic ok: %s\n", p.what());
  }
  // fft128 roundtrip
  {
    const size_t n = 64;
    fft128::Plan plan(n);
    std::vector<double> r0(n), r1(n, 0.0), i0(n), i1(n, 0.0);
    for (size_t i = 0; i < n; ++i) r0[i] = u(rng), i0[i] = u(rng);
    auto a = r0, b = i0;
    plan.fwd(r0, r1, i0, i1);
    plan.inv(r0, r1, i0, i1);
    double e = 0;
    for (size_t i = 0; i < n; ++i)
      e = std::fmax(e, std::fabs((r0[i] + r1[i]) / n - a[i]) + std::fabs((i0[i] + i1[i]) / n - b[i]));
    std::printf("fft128 n=%zu roundtrip err %.3g\n", n, e);
    failures += e >= 1e-28;
  }
  std::printf("%s\n", failures ? "FAILED" : "OK");
  return failures ? 1 : 0;
}
