// Library-level entry points: ABI version, error text, device queries.
#include <string.h>

#include <string>

#include "common.h"

namespace {
thread_local std::string g_last_error;
}

namespace fmlh {
int record_error(hipError_t e, const char *where) {
  g_last_error = std::string(where) + ": " + hipGetErrorString(e);
  return FMLH_EHIP;
}
}  // namespace fmlh

extern "C" {

int fmlh_abi_version(void) { return 1; }

const char *fmlh_last_error(void) { return g_last_error.c_str(); }

int fmlh_device_cu_count(void) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return 0;
  return cus;
}

}  // extern "C"

// ---- main-kernel timing --------------------------------------------------
// When enabled, each entry point brackets its dominant kernel with a pair of
// HIP events on the caller's stream.  fmlh_timing_collect() waits for the
// recorded pairs, returns the summed milliseconds and the number of pairs,
// and recycles the events.
#include <mutex>
#include <vector>

namespace {
std::mutex g_tmu;
bool g_timing = false;
std::vector<std::pair<hipEvent_t, hipEvent_t>> g_pairs;  // recorded
std::vector<std::pair<hipEvent_t, hipEvent_t>> g_free;   // reusable
hipEvent_t g_open = nullptr;                             // begin of the current pair
}  // namespace

namespace fmlh {
void timing_begin(hipStream_t st) {
  std::lock_guard<std::mutex> lk(g_tmu);
  if (!g_timing) return;
  hipEvent_t a = nullptr, b = nullptr;
  if (!g_free.empty()) {
    a = g_free.back().first;
    b = g_free.back().second;
    g_free.pop_back();
  } else {
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return;
  }
  g_pairs.emplace_back(a, b);
  g_open = a;
  hipEventRecord(a, st);
}
void timing_end(hipStream_t st) {
  std::lock_guard<std::mutex> lk(g_tmu);
  if (!g_timing || g_open == nullptr) return;
  hipEventRecord(g_pairs.back().second, st);
  g_open = nullptr;
}
}  // namespace fmlh

extern "C" {

void fmlh_timing_enable(int on) {
  std::lock_guard<std::mutex> lk(g_tmu);
  g_timing = on != 0;
}

int fmlh_timing_collect(double *total_ms, int *count) {
  std::lock_guard<std::mutex> lk(g_tmu);
  double tot = 0.0;
  int n = 0;
  for (auto &pr : g_pairs) {
    if (hipEventSynchronize(pr.second) != hipSuccess) continue;
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, pr.first, pr.second) == hipSuccess) {
      tot += ms;
      ++n;
    }
    g_free.push_back(pr);
  }
  g_pairs.clear();
  if (total_ms) *total_ms = tot;
  if (count) *count = n;
  return 0;
}

}  // extern "C"
