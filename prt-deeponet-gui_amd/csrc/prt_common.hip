// Error reporting and version of the PRT-LB HIP library.
#include <cstdarg>
#include <cstdio>

#include "prt_common.h"

namespace prt {

static thread_local char g_err[512] = "";

void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

}  // namespace prt

extern "C" const char *prt_last_error(void) { return prt::g_err; }

extern "C" const char *prt_version(void) { return "prtlb-hip 0.1 gfx950"; }

extern "C" size_t prt_adr_params_size(void) { return sizeof(prt_adr_params); }
