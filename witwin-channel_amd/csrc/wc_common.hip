// Library identity and error channel of the C ABI.
#include "wc_api_internal.h"

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>

namespace {
thread_local std::string g_last_error;
}

int wc_fail(const char* format, ...) {
    if (!format) {
        g_last_error = "unknown error";
        return 1;
    }
    char buffer[512];
    va_list args;
    va_start(args, format);
    std::vsnprintf(buffer, sizeof(buffer), format, args);
    va_end(args);
    g_last_error = buffer;
    return 1;
}

int wc_fail_hip(const char* where, hipError_t err) {
    char buffer[512];
    std::snprintf(buffer, sizeof(buffer), "%s: HIP error %d (%s)", where, (int)err, hipGetErrorString(err));
    g_last_error = buffer;
    return 2;
}

extern "C" const char* wc_last_error(void) { return g_last_error.c_str(); }

extern "C" void wc_clear_error(void) { g_last_error.clear(); }

extern "C" int wc_abi_version(void) { return 1; }

extern "C" const char* wc_build_target(void) { return "gfx950"; }
