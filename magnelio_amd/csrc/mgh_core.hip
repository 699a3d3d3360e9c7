// Library-wide entry points: error reporting, version and device queries.
#include "mgh_common.h"

#include <cstdarg>

namespace mgh {

static thread_local char g_last_error[1024] = {0};

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
}

}  // namespace mgh

extern "C" {

const char* mgh_last_error(void) { return mgh::g_last_error; }

int mgh_version(void) { return 10000; }  // 1.0.0

// Fills name (NUL-terminated, at most name_len bytes), CU count, HBM bytes
// and the gfx architecture string of the current device.
int mgh_device_info(char* name, int name_len, int* cu_count, long long* hbm_bytes, char* arch,
                    int arch_len) {
  int dev = 0;
  MGH_CHECK_HIP(hipGetDevice(&dev));
  hipDeviceProp_t p;
  MGH_CHECK_HIP(hipGetDeviceProperties(&p, dev));
  if (name && name_len > 0) {
    strncpy(name, p.name, name_len - 1);
    name[name_len - 1] = 0;
  }
  if (arch && arch_len > 0) {
    strncpy(arch, p.gcnArchName, arch_len - 1);
    arch[arch_len - 1] = 0;
  }
  if (cu_count) *cu_count = p.multiProcessorCount;
  if (hbm_bytes) *hbm_bytes = (long long)p.totalGlobalMem;
  return MGH_OK;
}

}  // extern "C"
