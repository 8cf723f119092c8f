// C-ABI plumbing shared by all entry points: version and per-thread last error.
#include <cstdarg>
#include <cstdio>
#include <cstring>

#include "jm_common.h"

namespace jm {
static thread_local char g_last_error[1024] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
}
}  // namespace jm

extern "C" int jm_version(void) { return JM_VERSION; }

extern "C" int jm_last_error(char* buf, size_t len) {
  if (!buf || len == 0) return JM_ERR_ARG;
  std::strncpy(buf, jm::g_last_error, len - 1);
  buf[len - 1] = '\0';
  return JM_OK;
}
