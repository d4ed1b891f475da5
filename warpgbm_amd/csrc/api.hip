// Library-level C ABI: error reporting, version, device queries.
#include "common.h"
#include <string.h>

namespace wgbm {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

}  // namespace wgbm

extern "C" const char* wgbm_last_error(void) { return wgbm::g_err; }

extern "C" int wgbm_abi_version(void) { return 4; }

// Name of the GPU architecture the library was built for.
extern "C" const char* wgbm_build_arch(void) { return "gfx950"; }

// Checks that the current device is the one the kernels were built for.
extern "C" int wgbm_check_device(char* name_out, int32_t name_len) {
  int dev = 0;
  WGBM_HIP(hipGetDevice(&dev));
  hipDeviceProp_t p;
  WGBM_HIP(hipGetDeviceProperties(&p, dev));
  if (name_out && name_len > 0) {
    strncpy(name_out, p.gcnArchName, (size_t)name_len - 1);
    name_out[name_len - 1] = 0;
  }
  if (strncmp(p.gcnArchName, "gfx950", 6) != 0) {
    wgbm::set_error("device architecture %s is not gfx950", p.gcnArchName);
    return wgbm::kBadArg;
  }
  return wgbm::kOk;
}
