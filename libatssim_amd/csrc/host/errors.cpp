// errors.cpp -- see errors.h.
#include "errors.h"

#include "../../../include/libatssim_hip.h"

namespace {
thread_local std::string g_last_error;
}

namespace lats {
void set_error(const std::string& msg) { g_last_error = msg; }
}  // namespace lats

extern "C" const char* lats_last_error(void) { return g_last_error.c_str(); }
