// host_common.hpp — error reporting shared by the host-side sources.
#pragma once

#include <cstdarg>
#include <cstdio>
#include <string>

namespace vba {

// Thread-local message for calls without a context (vba_last_error(NULL)).
inline std::string& thread_error() {
    static thread_local std::string msg;
    return msg;
}
inline void set_thread_error(const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    thread_error() = buf;
}

}  // namespace vba
