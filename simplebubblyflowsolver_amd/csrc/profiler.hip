// Kernel timing with HIP events recorded on the library stream around each
// launch while profiling is on (sbfs_profile_start / sbfs_profile_stop).
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "sbfs.h"
#include "sbfs_internal.hpp"

namespace sbfs {

struct ProfRec {
  const char* name;
  hipEvent_t a, b;
};

static bool g_on = false;
static std::vector<ProfRec> g_recs;
static std::vector<hipEvent_t> g_pool;
static size_t g_used = 0;
static std::vector<std::pair<std::string, std::pair<int, double>>> g_report;

static hipEvent_t take_event() {
  if (g_used == g_pool.size()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    g_pool.push_back(e);
  }
  return g_pool[g_used++];
}

bool prof_active() { return g_on; }

void prof_begin(const char* name) {
  if (!g_on) return;
  ProfRec r{name, take_event(), take_event()};
  if (!r.a || !r.b) return;
  (void)hipEventRecord(r.a, stream());
  g_recs.push_back(r);
}

void prof_end() {
  if (!g_on || g_recs.empty()) return;
  (void)hipEventRecord(g_recs.back().b, stream());
}

}  // namespace sbfs

using namespace sbfs;

extern "C" {

int sbfs_profile_start(void) {
  if (init_runtime()) return 1;
  SBFS_HIP(hipStreamSynchronize(stream()));
  g_recs.clear();
  g_used = 0;
  g_report.clear();
  g_on = true;
  return 0;
}

int sbfs_profile_stop(void) {
  if (init_runtime()) return 1;
  g_on = false;
  SBFS_HIP(hipStreamSynchronize(stream()));
  std::map<std::string, std::pair<int, double>> agg;
  std::vector<std::string> order;
  for (const ProfRec& r : g_recs) {
    float ms = 0.f;
    SBFS_HIP(hipEventElapsedTime(&ms, r.a, r.b));
    auto it = agg.find(r.name);
    if (it == agg.end()) {
      order.push_back(r.name);
      agg[r.name] = {1, (double)ms};
    } else {
      it->second.first += 1;
      it->second.second += ms;
    }
  }
  g_report.clear();
  for (const std::string& n : order) g_report.push_back({n, agg[n]});
  g_recs.clear();
  g_used = 0;
  return 0;
}

int sbfs_profile_count(void) { return (int)g_report.size(); }

int sbfs_profile_entry(int k, char* name, size_t len, int* count, double* total_ms) {
  SBFS_REQUIRE(k >= 0 && k < (int)g_report.size(), "profile entry %d out of range", k);
  snprintf(name, len, "%s", g_report[k].first.c_str());
  *count = g_report[k].second.first;
  *total_ms = g_report[k].second.second;
  return 0;
}

}  // extern "C"
