// Library-wide host state: error messages and the version query.
#include <cstdarg>
#include <cstdio>

#include "common.h"
#include "msc_hip.h"

namespace msc {

static thread_local char g_error[1024] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_error, sizeof(g_error), fmt, ap);
  va_end(ap);
}

const char* last_error() { return g_error; }

}  // namespace msc

extern "C" int msc_version(void) { return 100; }
extern "C" const char* msc_last_error(void) { return msc::last_error(); }
