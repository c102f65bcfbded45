// tune_cache.cpp -- process table and file of Method::Measure choices
// (format and default path in tune_cache.hpp).  Host code only.
#include "tune_cache.hpp"

#include <unistd.h>

#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <sys/stat.h>
#include <tuple>

#include "cfft.h"

namespace cfft_internal {

int set_error(int status, const std::string &msg);

namespace {

using Key = std::tuple<std::string, int, size_t>;
using Table = std::map<Key, TuneChoice>;

constexpr const char *kHeader = "# cfft_amd tuning cache v1: device kind n algo base_n";

std::mutex g_mu;
Table g_table;
bool g_default_read = false;

bool pow2(size_t v) { return v && !(v & (v - 1)); }

std::string default_path() {
  const char *env = std::getenv("CFFT_TUNE_CACHE");
  if (env) {
    if (!*env || std::strcmp(env, "off") == 0) return std::string();
    return env;
  }
  const char *xdg = std::getenv("XDG_CACHE_HOME");
  if (xdg && *xdg) return std::string(xdg) + "/cfft_amd/tune-cache.tsv";
  const char *home = std::getenv("HOME");
  if (home && *home) return std::string(home) + "/.cache/cfft_amd/tune-cache.tsv";
  return std::string();
}

// one table line: device \t kind \t n \t algo \t base_n
bool parse_line(const std::string &line, Key *k, TuneChoice *c) {
  if (line.empty() || line[0] == '#') return false;
  size_t tab[4];
  size_t from = 0;
  for (int i = 0; i < 4; ++i) {
    tab[i] = line.find('\t', from);
    if (tab[i] == std::string::npos) return false;
    from = tab[i] + 1;
  }
  if (line.find('\t', from) != std::string::npos || tab[0] == 0) return false;
  unsigned long long v[4];
  for (int i = 0; i < 4; ++i) {
    const size_t b = tab[i] + 1, e = i < 3 ? tab[i + 1] : line.size();
    if (e <= b || e - b > 19) return false;
    for (size_t j = b; j < e; ++j)
      if (line[j] < '0' || line[j] > '9') return false;
    v[i] = std::strtoull(line.c_str() + b, nullptr, 10);
  }
  if (v[0] > 1 || v[2] > 7 || !tune_choice_valid((int)v[0], (size_t)v[1], (int)v[2], (size_t)v[3]))
    return false;
  *k = Key{line.substr(0, tab[0]), (int)v[0], (size_t)v[1]};
  *c = TuneChoice{(int)v[2], (size_t)v[3]};
  return true;
}

// reads `path` into `t` (entries already in `t` are kept when `keep` is
// true); false when the file cannot be opened
bool read_file(const std::string &path, Table &t, bool keep, size_t *read) {
  FILE *f = std::fopen(path.c_str(), "r");
  if (!f) return false;
  size_t count = 0;
  std::string line;
  bool overlong = false;  // dropping the rest of a line above 4 KiB
  char buf[512];
  while (std::fgets(buf, sizeof buf, f)) {
    const bool eol = buf[0] && buf[std::strlen(buf) - 1] == '\n';
    if (overlong) {
      overlong = !eol;
      continue;
    }
    line += buf;
    if (!eol) {
      if (line.size() > 4096) {
        line.clear();
        overlong = true;
      }
      continue;
    }
    line.pop_back();
    Key k;
    TuneChoice c;
    if (parse_line(line, &k, &c)) {
      if (!keep || t.find(k) == t.end()) t[k] = c;
      ++count;
    }
    line.clear();
  }
  std::fclose(f);
  if (read) *read = count;
  return true;
}

void make_parent_dirs(const std::string &path) {
  for (size_t i = 1; i < path.size(); ++i) {
    if (path[i] != '/') continue;
    const std::string dir = path.substr(0, i);
    if (::mkdir(dir.c_str(), 0755) != 0 && errno != EEXIST) return;
  }
}

// writes `t` to `path` through a temporary file in the same directory
bool write_file(const std::string &path, const Table &t) {
  make_parent_dirs(path);
  std::ostringstream tmpname;
  tmpname << path << ".tmp." << ::getpid();
  const std::string tmp = tmpname.str();
  FILE *f = std::fopen(tmp.c_str(), "w");
  if (!f) return false;
  bool ok = std::fprintf(f, "%s\n", kHeader) > 0;
  for (const auto &kv : t) {
    if (!ok) break;
    ok = std::fprintf(f, "%s\t%d\t%zu\t%d\t%zu\n", std::get<0>(kv.first).c_str(),
                      std::get<1>(kv.first), std::get<2>(kv.first), kv.second.algo,
                      kv.second.base_n) > 0;
  }
  ok = (std::fclose(f) == 0) && ok;
  if (ok) ok = std::rename(tmp.c_str(), path.c_str()) == 0;
  if (!ok) std::remove(tmp.c_str());
  return ok;
}

bool key_ok(const std::string &device_key) {
  return !device_key.empty() && device_key.find_first_of("\t\n\r") == std::string::npos &&
         device_key[0] != '#';
}

}  // namespace

bool tune_choice_valid(int kind, size_t n, int algo, size_t base_n) {
  if (algo < CFFT_DIF2 || algo > CFFT_DIT16 || !pow2(n) || n > ((size_t)1 << 30)) return false;
  if (kind == TUNE_ORDERED) return base_n == n && n <= ((size_t)1 << 16);
  if (kind != TUNE_UNORDERED || !pow2(base_n) || base_n > n || base_n > 1024) return false;
  return base_n == n || base_n >= 32;
}

bool tune_lookup(const std::string &device_key, int kind, size_t n, TuneChoice *out) {
  std::lock_guard<std::mutex> g(g_mu);
  if (!g_default_read) {
    g_default_read = true;
    const std::string path = default_path();
    if (!path.empty()) read_file(path, g_table, true, nullptr);
  }
  auto it = g_table.find(Key{device_key, kind, n});
  if (it == g_table.end()) return false;
  *out = it->second;
  return true;
}

void tune_store(const std::string &device_key, int kind, size_t n, const TuneChoice &c) {
  if (!key_ok(device_key) || !tune_choice_valid(kind, n, c.algo, c.base_n)) return;
  std::lock_guard<std::mutex> g(g_mu);
  g_table[Key{device_key, kind, n}] = c;
  const std::string path = default_path();
  if (path.empty()) return;
  // the file keeps what other processes wrote since this one read it; the
  // process table is not refilled from it (a cleared table stays cleared)
  Table merged;
  read_file(path, merged, false, nullptr);
  for (const auto &kv : g_table) merged[kv.first] = kv.second;
  (void)write_file(path, merged);
}

}  // namespace cfft_internal

using cfft_internal::set_error;

extern "C" {

int cfft_tune_cache_path(char *buf, size_t cap, size_t *len) {
  const std::string p = cfft_internal::default_path();
  if (len) *len = p.size();
  if (!buf) return cap ? set_error(CFFT_ERROR_NULL_POINTER, "buf is null") : CFFT_SUCCESS;
  if (cap <= p.size()) return set_error(CFFT_ERROR_SIZE, "buffer too small for the path");
  std::memcpy(buf, p.c_str(), p.size() + 1);
  return CFFT_SUCCESS;
}

int cfft_tune_cache_import(const char *path, size_t *entries) {
  if (!path) return set_error(CFFT_ERROR_NULL_POINTER, "path is null");
  if (entries) *entries = 0;
  std::lock_guard<std::mutex> g(cfft_internal::g_mu);
  if (!cfft_internal::read_file(path, cfft_internal::g_table, false, entries))
    return set_error(CFFT_ERROR_INVALID_ARGUMENT, std::string("cannot read ") + path);
  return CFFT_SUCCESS;
}

int cfft_tune_cache_export(const char *path) {
  if (!path) return set_error(CFFT_ERROR_NULL_POINTER, "path is null");
  std::lock_guard<std::mutex> g(cfft_internal::g_mu);
  if (!cfft_internal::write_file(path, cfft_internal::g_table))
    return set_error(CFFT_ERROR_INVALID_ARGUMENT, std::string("cannot write ") + path);
  return CFFT_SUCCESS;
}

int cfft_tune_cache_clear(void) {
  std::lock_guard<std::mutex> g(cfft_internal::g_mu);
  cfft_internal::g_table.clear();
  cfft_internal::g_default_read = true;  // the default file is not read back
  return CFFT_SUCCESS;
}

int cfft_tune_cache_size(size_t *entries) {
  if (!entries) return set_error(CFFT_ERROR_NULL_POINTER, "entries is null");
  std::lock_guard<std::mutex> g(cfft_internal::g_mu);
  *entries = cfft_internal::g_table.size();
  return CFFT_SUCCESS;
}

int cfft_tune_cache_lookup(const char *device_key, int kind, size_t n, int *found, int *algo,
                           size_t *base_n) {
  if (!device_key || !found) return set_error(CFFT_ERROR_NULL_POINTER, "null argument");
  cfft_internal::TuneChoice c;
  *found = cfft_internal::tune_lookup(device_key, kind, n, &c) ? 1 : 0;
  if (*found) {
    if (algo) *algo = c.algo;
    if (base_n) *base_n = c.base_n;
  }
  return CFFT_SUCCESS;
}

}  // extern "C"
