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
  // the error is reported through the return code: clear the runtime's
  // per-thread last-error state so it does not surface in a later,
  // unrelated call of the caller (torch checks it after its own calls)
  (void)hipGetLastError();
  return FMLH_EHIP;
}

int record_message(int rc, const char *text) {
  g_last_error = text;
  return rc;
}
}  // namespace fmlh

extern "C" {

int fmlh_abi_version(void) { return 7; }

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
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(st, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) {
    (void)hipGetLastError();
    return;  // no timing events inside a graph being captured
  }
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
    float ms = 0.f;
    if (hipEventSynchronize(pr.second) == hipSuccess &&
        hipEventElapsedTime(&ms, pr.first, pr.second) == hipSuccess) {
      tot += ms;
      ++n;
    } else {
      (void)hipGetLastError();  // a pair that never ran: skipped, not reported
    }
    g_free.push_back(pr);
  }
  g_pairs.clear();
  if (total_ms) *total_ms = tot;
  if (count) *count = n;
  return 0;
}

}  // extern "C"

// ---- HIP graphs ------------------------------------------------------------
// Thread-local stream capture: only this thread's calls on the capturing
// stream are recorded, and other threads (a serving loop's second stream,
// torch's own work) keep running eagerly.

struct fmlh_graph {
  hipGraph_t graph;
  hipGraphExec_t exec;
};

extern "C" {

int fmlh_graph_begin(void *stream) {
  int rc = fmlh_init();
  if (rc) return rc;
  hipError_t e = hipStreamBeginCapture(fmlh::as_stream(stream), hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return fmlh::record_error(e, "hipStreamBeginCapture");
  return 0;
}

int fmlh_graph_end(void *stream, fmlh_graph **out) {
  FMLH_ARG_CHECK(out != nullptr, 2);
  *out = nullptr;
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(fmlh::as_stream(stream), &g);
  if (e != hipSuccess) return fmlh::record_error(e, "hipStreamEndCapture");
  hipGraphExec_t x = nullptr;
  e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    hipGraphDestroy(g);
    return fmlh::record_error(e, "hipGraphInstantiate");
  }
  *out = new fmlh_graph{g, x};
  return 0;
}

int fmlh_graph_launch(fmlh_graph *graph, void *stream) {
  FMLH_ARG_CHECK(graph != nullptr, 1);
  hipError_t e = hipGraphLaunch(graph->exec, fmlh::as_stream(stream));
  if (e != hipSuccess) return fmlh::record_error(e, "hipGraphLaunch");
  return 0;
}

int fmlh_graph_destroy(fmlh_graph *graph) {
  if (!graph) return 0;
  hipError_t e1 = hipGraphExecDestroy(graph->exec);
  hipError_t e2 = hipGraphDestroy(graph->graph);
  delete graph;
  if (e1 != hipSuccess) return fmlh::record_error(e1, "hipGraphExecDestroy");
  if (e2 != hipSuccess) return fmlh::record_error(e2, "hipGraphDestroy");
  return 0;
}

}  // extern "C"

// ---- CU-partitioned streams ---------------------------------------------------
// A stream whose kernels may only occupy the first `cu_count` compute units
// of the device (all of them for cu_count <= 0).  Each such stream owns its
// hardware queue, so work on two of them never waits in one queue; a serving
// loop gives its fit work a stream that leaves some CUs to its predictions.

#include <vector>

extern "C" {

}  // extern "C"

namespace {
std::mutex g_mask_mu;
std::vector<std::pair<hipStream_t, std::vector<uint32_t>>> g_masked;  // streams made by fmlh_stream_create
}  // namespace

namespace fmlh {
bool stream_cu_mask(hipStream_t st, std::vector<uint32_t> &mask) {
  if (st == nullptr) return false;
  std::lock_guard<std::mutex> lk(g_mask_mu);
  for (const auto &e : g_masked)
    if (e.first == st) {
      mask = e.second;
      return true;
    }
  return false;
}
}  // namespace fmlh

extern "C" {

int fmlh_stream_create(int cu_count, void **stream) {
  FMLH_ARG_CHECK(stream != nullptr, 2);
  *stream = nullptr;
  const int cus = fmlh_device_cu_count();
  if (cus <= 0) return fmlh::record_error(hipErrorInvalidDevice, "fmlh_stream_create(device)");
  const int n = (cu_count <= 0 || cu_count > cus) ? cus : cu_count;
  std::vector<uint32_t> mask((size_t)(cus + 31) / 32, 0u);
  for (int i = 0; i < n; ++i) mask[(size_t)i / 32] |= 1u << (i % 32);
  hipStream_t s = nullptr;
  hipError_t e = hipExtStreamCreateWithCUMask(&s, (uint32_t)mask.size(), mask.data());
  if (e != hipSuccess) return fmlh::record_error(e, "hipExtStreamCreateWithCUMask");
  {
    std::lock_guard<std::mutex> lk(g_mask_mu);
    g_masked.emplace_back(s, mask);
  }
  *stream = s;
  return 0;
}

int fmlh_stream_destroy(void *stream) {
  FMLH_ARG_CHECK(stream != nullptr, 1);
  {
    std::lock_guard<std::mutex> lk(g_mask_mu);
    for (size_t i = 0; i < g_masked.size(); ++i)
      if (g_masked[i].first == fmlh::as_stream(stream)) {
        g_masked.erase(g_masked.begin() + (long)i);
        break;
      }
  }
  hipError_t e = hipStreamDestroy(fmlh::as_stream(stream));
  if (e != hipSuccess) return fmlh::record_error(e, "hipStreamDestroy");
  return 0;
}

}  // extern "C"
