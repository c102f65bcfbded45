// tune_cache.hpp -- the Method::Measure choices, kept per process and in a
// file shared by later processes.
//
// An entry maps (device key, plan kind, n) to (algo, base_n).  The device key
// names the device model, its ISA, its compute-unit count, the HIP runtime
// version and this library's ABI version, so a choice measured on another
// device or stack is never reused.  The file is a tab-separated text table:
//
//   # cfft_amd tuning cache v1: device kind n algo base_n
//   AMD Instinct MI355X|gfx950:sramecc+:xnack-|cu256|hip70226015|abi3<TAB>1<TAB>1024<TAB>6<TAB>1024
//
// Lines that do not parse, or name a choice no plan of that kind accepts, are
// skipped.  The default file is $CFFT_TUNE_CACHE, else
// $XDG_CACHE_HOME/cfft_amd/tune-cache.tsv, else
// $HOME/.cache/cfft_amd/tune-cache.tsv; CFFT_TUNE_CACHE set to "" or "off"
// keeps the choices in the process only.  The default file is read once, on
// the first lookup, and rewritten (its entries merged with the process's,
// through a temporary file and a rename) after each new measurement.
#pragma once

#include <cstddef>
#include <string>

namespace cfft_internal {

enum TuneKind { TUNE_ORDERED = 0, TUNE_UNORDERED = 1 };

struct TuneChoice {
  int algo = 0;
  size_t base_n = 0;
};

// true when the process table (after reading the default file once) holds
// a choice for this key
bool tune_lookup(const std::string &device_key, int kind, size_t n, TuneChoice *out);
// records a measured choice and rewrites the default file; a file that
// cannot be written leaves the choice in the process only
void tune_store(const std::string &device_key, int kind, size_t n, const TuneChoice &c);
// whether (kind, n, algo, base_n) is a choice a plan of that kind accepts
bool tune_choice_valid(int kind, size_t n, int algo, size_t base_n);

}  // namespace cfft_internal
