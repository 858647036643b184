// Autotuner for kernel variants (replaces Tuner.hpp:73-358).
//
// A kernel with several launch variants (occupancy target, block shape,
// schedule) asks pick_variant() which one to use for a given field shape.
// With tuning enabled (KLFT_TUNING / klft_tuning) and no cached entry, every
// variant is timed with HIP events on the library stream (one warm-up, then
// the minimum of a few runs) through a caller-supplied launcher that must not
// change the live field (read-only kernels, or a scratch copy), unlike the
// reference, whose tuner re-runs the real functor on live data.
// The cache file keeps the reference's line shape "<rank> <key> v0 .. v_{rank-1}"
// (Tuner.hpp:268-358) with v0 = variant, v1 = time in ns, the rest 0.
// KLFT_VARIANT="name=v,name2=v" forces variants (benchmarking all variants).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <functional>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <utility>
#include <vector>

#include "internal.hpp"

namespace klft_internal {

struct TuneEntry {
  int rank;
  int variant;
  long long ns;
};

static std::mutex g_tmu;
static std::map<std::string, TuneEntry> g_table;
static int g_tuning = -1;

int cu_count() {
  static const int n = [] {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0)
      return 256;
    return v;
  }();
  return n;
}

int blocks_per_cu(const void *kern, int block) {
  static std::mutex mu;
  static std::map<std::pair<const void *, int>, int> cache;
  std::lock_guard<std::mutex> lk(mu);
  const auto key = std::make_pair(kern, block);
  auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kern, block, 0) != hipSuccess || n < 1) n = 1;
  cache[key] = n;
  return n;
}

bool tuning_enabled() {
  if (g_tuning < 0) {
    const char *e = std::getenv("KLFT_TUNING");
    g_tuning = e ? (std::atoi(e) != 0) : 0;
  }
  return g_tuning != 0;
}

std::string tune_key(const char *kernel, const klft_field_s *f) {
  std::string k = std::string(kernel) + "_nc_" + std::to_string(f->nc) + "_storage_" +
                  std::to_string(f->storage) + "_end_";
  for (int d = 0; d < f->nd; ++d) k += std::to_string(f->dims[d]) + "_";
  return k;
}

// KLFT_VARIANT is parsed once per distinct value (the launchers ask on every
// call, and the value may change between calls)
static std::mutex g_vmu;
static std::string g_venv;
static bool g_venv_set = false;
static std::map<std::string, int> g_forced;

int forced_variant(const char *kernel) {
  const char *e = std::getenv("KLFT_VARIANT");
  std::lock_guard<std::mutex> lk(g_vmu);
  if (!e) {
    g_venv_set = false;
    return -1;
  }
  if (!g_venv_set || g_venv != e) {
    g_venv = e;
    g_venv_set = true;
    g_forced.clear();
    size_t pos = 0;
    while (pos < g_venv.size()) {
      size_t end = g_venv.find(',', pos);
      if (end == std::string::npos) end = g_venv.size();
      const std::string item = g_venv.substr(pos, end - pos);
      const size_t eq = item.find('=');
      if (eq != std::string::npos) g_forced[item.substr(0, eq)] = std::atoi(item.c_str() + eq + 1);
      pos = end + 1;
    }
  }
  auto it = g_forced.find(kernel);
  return it == g_forced.end() ? -1 : it->second;
}

int chosen_variant(const char *kernel, const klft_field_s *f, int nvar, int dflt) {
  const int forced = forced_variant(kernel);
  if (forced >= 0 && forced < nvar) return forced;
  const std::string key = tune_key(kernel, f);
  std::lock_guard<std::mutex> lk(g_tmu);
  auto it = g_table.find(key);
  if (it != g_table.end() && it->second.variant >= 0 && it->second.variant < nvar) return it->second.variant;
  return dflt;
}

int pick_variant(const char *kernel, const klft_field_s *f, int nvar, int dflt,
                 const std::function<int(int)> &run) {
  const int forced = forced_variant(kernel);
  if (forced >= 0 && forced < nvar) return forced;
  const std::string key = tune_key(kernel, f);
  {
    std::lock_guard<std::mutex> lk(g_tmu);
    auto it = g_table.find(key);
    if (it != g_table.end() && it->second.variant >= 0 && it->second.variant < nvar) return it->second.variant;
  }
  if (!tuning_enabled() || nvar <= 1) return dflt;
  hipEvent_t a, b;
  if (hipEventCreate(&a) != hipSuccess) return dflt;
  if (hipEventCreate(&b) != hipSuccess) {
    (void)hipEventDestroy(a);
    return dflt;
  }
  int best = dflt;
  float best_ms = 1e30f;
  for (int v = 0; v < nvar; ++v) {
    if (run(v) != KLFT_OK) continue;  // warm-up
    float vmin = 1e30f;
    for (int rep = 0; rep < 3; ++rep) {
      (void)hipEventRecord(a, stream());
      if (run(v) != KLFT_OK) break;
      (void)hipEventRecord(b, stream());
      (void)hipEventSynchronize(b);
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, a, b);
      if (ms < vmin) vmin = ms;
    }
    if (vmin < best_ms) {
      best_ms = vmin;
      best = v;
    }
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  std::lock_guard<std::mutex> lk(g_tmu);
  g_table[key] = TuneEntry{f->nd, best, (long long)(best_ms * 1e6)};
  return best;
}

// ---- kernel timing ----
struct ProfRec {
  std::string name;
  hipEvent_t a, b;
};
static std::mutex g_pmu;
static std::vector<ProfRec> g_prof;
static bool g_prof_on = false;

bool prof_enabled() { return g_prof_on; }

void prof_record(const char *name, hipEvent_t start, hipEvent_t stop) {
  std::lock_guard<std::mutex> lk(g_pmu);
  g_prof.push_back(ProfRec{name, start, stop});
}

}  // namespace klft_internal

namespace ki = klft_internal;

extern "C" int klft_profile(int on) {
  std::lock_guard<std::mutex> lk(ki::g_pmu);
  for (auto &r : ki::g_prof) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  ki::g_prof.clear();
  ki::g_prof_on = on != 0;
  return KLFT_OK;
}

extern "C" int klft_profile_read(const char *name, double *total_ms, int64_t *count) {
  if (!name || !total_ms || !count) return ki::set_error(KLFT_E_ARG, "null argument");
  KLFT_HIP(hipStreamSynchronize(ki::stream()));
  std::lock_guard<std::mutex> lk(ki::g_pmu);
  double ms = 0.0;
  int64_t n = 0;
  std::vector<ki::ProfRec> keep;
  for (auto &r : ki::g_prof) {
    if (r.name == name) {
      float t = 0.f;
      if (hipEventElapsedTime(&t, r.a, r.b) == hipSuccess) {
        ms += t;
        ++n;
      }
      (void)hipEventDestroy(r.a);
      (void)hipEventDestroy(r.b);
    } else {
      keep.push_back(r);
    }
  }
  ki::g_prof.swap(keep);
  *total_ms = ms;
  *count = n;
  return KLFT_OK;
}

extern "C" int klft_tuning(int on) {
  ki::g_tuning = on ? 1 : 0;
  return KLFT_OK;
}

extern "C" int klft_tune_cache_read(const char *path) {
  if (!path) return ki::set_error(KLFT_E_ARG, "null path");
  std::ifstream in(path);
  if (!in.is_open()) {
    printf("Could not open cache file %s\n", path);  // Tuner.hpp:308-311: not fatal
    return ki::set_error(KLFT_E_IO, "could not open cache file %s", path);
  }
  std::string line;
  std::lock_guard<std::mutex> lk(ki::g_tmu);
  while (std::getline(in, line)) {
    std::istringstream iss(line);
    int rank = 0;
    std::string key;
    if (!(iss >> rank >> key)) continue;
    long long v[4] = {0, 0, 0, 0};
    for (int i = 0; i < rank && i < 4; ++i) iss >> v[i];
    ki::g_table[key] = ki::TuneEntry{rank, (int)v[0], v[1]};
  }
  return KLFT_OK;
}

extern "C" int klft_tune_cache_write(const char *path) {
  if (!path) return ki::set_error(KLFT_E_ARG, "null path");
  std::ofstream out(path, std::ios::out);
  if (!out.is_open()) return ki::set_error(KLFT_E_IO, "could not open cache file %s", path);
  std::lock_guard<std::mutex> lk(ki::g_tmu);
  for (const auto &e : ki::g_table) {
    out << e.second.rank << " " << e.first << " " << e.second.variant << " " << e.second.ns << " ";
    for (int i = 2; i < e.second.rank; ++i) out << 0 << " ";
    out << "\n";
  }
  return KLFT_OK;
}

extern "C" int klft_tune_lookup(const char *key, int *variant) {
  if (!key || !variant) return ki::set_error(KLFT_E_ARG, "null argument");
  std::lock_guard<std::mutex> lk(ki::g_tmu);
  auto it = ki::g_table.find(key);
  if (it == ki::g_table.end()) return ki::set_error(KLFT_E_STATE, "no tuning entry for %s", key);
  *variant = it->second.variant;
  return KLFT_OK;
}
