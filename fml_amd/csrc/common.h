// Shared helpers for libfmlhip (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "fmlh.h"

namespace fmlh {

// Records a HIP error for fmlh_last_error() and returns FMLH_EHIP.
int record_error(hipError_t e, const char *where);

// Records a message for fmlh_last_error() and returns rc.
int record_message(int rc, const char *text);

// Brackets an entry point's dominant kernel with timing events (no-op
// unless fmlh_timing_enable(1) was called).
void timing_begin(hipStream_t st);
void timing_end(hipStream_t st);

// The CU mask of a stream made by fmlh_stream_create (false for any other
// stream).  Work the library forks off such a stream onto a stream of its
// own uses a stream with the same mask.
bool stream_cu_mask(hipStream_t st, std::vector<uint32_t> &mask);

// Checks the launch that was just queued.
inline int check_launch(const char *where) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return record_error(e, where);
  return 0;
}

inline int ceil_div(int a, int b) { return (a + b - 1) / b; }
inline int64_t ceil_div64(int64_t a, int64_t b) { return (a + b - 1) / b; }

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

// 256-byte aligned carve-out from a caller workspace.
struct WsCarver {
  char *base;
  size_t size;
  size_t used = 0;
  WsCarver(void *p, size_t n) : base(static_cast<char *>(p)), size(n) {}
  template <typename T>
  T *take(size_t count) {
    size_t off = (used + 255) & ~size_t(255);
    size_t bytes = count * sizeof(T);
    if (off + bytes > size || base == nullptr) return nullptr;
    used = off + bytes;
    return reinterpret_cast<T *>(base + off);
  }
  static size_t need(size_t bytes) { return (bytes + 255) & ~size_t(255); }
};

}  // namespace fmlh

#define FMLH_ARG_CHECK(cond, idx) \
  do {                            \
    if (!(cond)) return -(idx);   \
  } while (0)

#define FMLH_LAUNCH_CHECK(where)                  \
  do {                                            \
    int rc__ = ::fmlh::check_launch(where);       \
    if (rc__) return rc__;                        \
  } while (0)
