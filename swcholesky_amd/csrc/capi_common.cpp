// Thread-local last-error message of the C ABI.
#include "capi_common.h"

#include <cstdio>
#include <cstring>

namespace {
thread_local char g_msg[512] = "";
}

int swc_set_error(int code, const std::string& msg) {
  std::snprintf(g_msg, sizeof g_msg, "%s", msg.c_str());
  return code;
}

extern "C" const char* swc_last_error(void) { return g_msg; }
