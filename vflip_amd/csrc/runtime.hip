// Library-wide runtime helpers: error reporting, version, device query.
#include "common.h"

#include <cstring>

namespace vflip {
static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}
}  // namespace vflip

extern "C" {

const char* vflip_last_error(void) { return vflip::g_err; }

const char* vflip_version(void) { return "vflip-mi355x 0.1 (gfx950)"; }

int vflip_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

}  // extern "C"
