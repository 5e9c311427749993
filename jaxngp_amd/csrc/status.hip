// status.hip — last-error state and version of libngp_hip.so.
//
// The error state is process-wide, not per thread: a host framework may run the
// targets on its own executor threads (XLA runs custom calls there), while the
// caller that polls ngp_get_last_error() sits on another thread.  The code is an
// atomic so polling is cheap; the message is guarded by a mutex.  The first error
// recorded since the last poll is kept (later ones are usually its consequences).
#include <atomic>
#include <cstdarg>
#include <cstdio>
#include <mutex>

#include "ngp_common.h"

namespace ngp {

namespace {
std::atomic<int> g_code{kOk};
std::mutex g_mutex;
char g_message[512] = {0};
}  // namespace

void set_error(int code, const char *fmt, ...) {
    std::lock_guard<std::mutex> lock(g_mutex);
    if (g_code.load(std::memory_order_relaxed) != kOk) return;
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(g_message, sizeof(g_message), fmt, ap);
    va_end(ap);
    g_code.store(code, std::memory_order_release);
}

bool has_error() { return g_code.load(std::memory_order_acquire) != kOk; }

}  // namespace ngp

extern "C" int ngp_get_last_error(char *buf, size_t buf_len) {
    if (ngp::g_code.load(std::memory_order_acquire) == ngp::kOk) {
        if (buf != nullptr && buf_len > 0) buf[0] = '\0';
        return ngp::kOk;
    }
    std::lock_guard<std::mutex> lock(ngp::g_mutex);
    int code = ngp::g_code.load(std::memory_order_relaxed);
    if (buf != nullptr && buf_len > 0) {
        std::snprintf(buf, buf_len, "%s", code == ngp::kOk ? "" : ngp::g_message);
    }
    ngp::g_code.store(ngp::kOk, std::memory_order_relaxed);
    ngp::g_message[0] = '\0';
    return code;
}

extern "C" const char *ngp_version(void) { return "ngp-hip 0.2.0 gfx950"; }
