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
