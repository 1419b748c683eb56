// runtime.cpp — host-side helpers of the C ABI: error text, version, and
// the launch-size query shared by the streaming kernels.
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstring>

#include "common.hpp"

namespace ysql {

static thread_local char g_last_error[256] = "";

void set_last_error(const char *msg) {
  strncpy(g_last_error, msg ? msg : "", sizeof(g_last_error) - 1);
  g_last_error[sizeof(g_last_error) - 1] = 0;
}

int grid_cap() {
  // CU count of the current device; 256 on MI355X.  Cached after the first
  // query (the value is a property of the device, not mutable state).
  static std::atomic<int> cached{0};
  int v = cached.load(std::memory_order_relaxed);
  if (v > 0) return v;
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
    cus = 256;
  cached.store(cus, std::memory_order_relaxed);
  return cus;
}

}  // namespace ysql

extern "C" {

const char *ysql_version(void) { return "yallasql_amd 0.1 (gfx950)"; }

const char *ysql_last_error(void) { return ysql::g_last_error; }

}  // extern "C"
