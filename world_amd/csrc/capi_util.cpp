#include "capi_util.h"

namespace {
thread_local std::string g_last_error;
}

namespace wd {

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

void clear_error() { g_last_error.clear(); }

}  // namespace wd

extern "C" const char* wd_last_error(void) { return g_last_error.c_str(); }

extern "C" const char* wd_version(void) { return "world_amd 0.1 (gfx950)"; }

extern "C" size_t wd_config_size(void) { return sizeof(wd_config); }
