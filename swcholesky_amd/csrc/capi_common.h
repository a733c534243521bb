// Shared error reporting for the C ABI.
#pragma once
#include <string>

#include "../../include/swcholesky_amd.h"

// Records `msg` as the calling thread's last error and returns `code`.
int swc_set_error(int code, const std::string& msg);
