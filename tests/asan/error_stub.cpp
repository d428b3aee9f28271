// Stand-in for the library's error recorder (engine.cpp), so the fuzz build
// links io.cpp alone, without the HIP runtime.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>

namespace segtaco {
int32_t set_error(int32_t status, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  return status;
}
int32_t hip_error(hipError_t, const char *) { return 2; }
}  // namespace segtaco
