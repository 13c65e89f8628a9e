// Error reporting and version of the C ABI.
#include "sn_common.h"

namespace sn {

static thread_local char g_err[512] = {0};

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

void clear_error() { g_err[0] = 0; }

}  // namespace sn

extern "C" {

const char* sn_last_error(void) { return sn::g_err; }

int sn_abi_version(void) { return 2; }

}  // extern "C"
