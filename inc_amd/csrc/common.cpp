// Error state and small host helpers of libinc_amd.
#include "common.hpp"

namespace inc {

namespace {
thread_local std::string g_last_error;
}

void set_error(const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
}

void clear_error() { g_last_error.clear(); }

const char* last_error() { return g_last_error.c_str(); }

int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("%s: kernel launch failed: %s", what, hipGetErrorString(e));
    return INC_EHIP;
  }
  return INC_OK;
}

}  // namespace inc

extern "C" {
int inc_version(void) { return 100; }
const char* inc_last_error(void) { return inc::last_error(); }
}
