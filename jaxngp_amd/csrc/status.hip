// status.hip — per-thread last-error state and version of libngp_hip.so.
#include <cstdarg>
#include <cstdio>

#include "ngp_common.h"

namespace ngp {

namespace {
thread_local int g_code = kOk;
thread_local char g_message[512] = {0};
}  // namespace

void set_error(int code, const char *fmt, ...) {
    g_code = code;
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(g_message, sizeof(g_message), fmt, ap);
    va_end(ap);
}

bool has_error() { return g_code != kOk; }

}  // namespace ngp

extern "C" int ngp_get_last_error(char *buf, size_t buf_len) {
    int code = ngp::g_code;
    if (buf != nullptr && buf_len > 0) {
        std::snprintf(buf, buf_len, "%s", code == ngp::kOk ? "" : ngp::g_message);
    }
    ngp::g_code = ngp::kOk;
    ngp::g_message[0] = '\0';
    return code;
}

extern "C" const char *ngp_version(void) { return "ngp-hip 0.1.0 gfx950"; }
