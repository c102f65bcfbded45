// Dumps cfft_prog::fixed_program's pass lists (the builder the plans use) for
// every algorithm and size given on the command line, one line per program:
//   algo log2n log2bn fwd npass kind:radix:fwd:log2len:s:tw_mul ...
// Used by tests/test_program_builder.py to check the lists against the
// reference's plan structure on the CPU.
#include <cstdio>
#include <cstdlib>

#include "program.hpp"

int main(int argc, char **argv) {
  const uint32_t max_log2n = argc > 1 ? (uint32_t)std::atoi(argv[1]) : 20;
  for (int algo = 0; algo < 8; ++algo)
    for (uint32_t l2n = 0; l2n <= max_log2n; ++l2n)
      for (uint32_t l2b = 0; l2b <= l2n && l2b <= 16; ++l2b)
        for (int fwd = 0; fwd < 2; ++fwd) {
          const cfft_prog::Program P = cfft_prog::fixed_program(algo, l2n, l2b, fwd != 0);
          std::printf("%d %u %u %d %u", algo, l2n, l2b, fwd, P.npass);
          const uint32_t np = P.npass <= (uint32_t)cfft_prog::kMaxPasses ? P.npass : 0;
          for (uint32_t i = 0; i < np; ++i)
            std::printf(" %u:%u:%u:%u:%u:%u", P.p[i].kind, P.p[i].radix, P.p[i].fwd,
                        P.p[i].log2len, P.p[i].s, P.p[i].tw_mul);
          std::printf("\n");
        }
  return 0;
}
