// sf_core.hip -- error state, device probing and version of the C ABI.
#include <cstdarg>
#include <cstddef>
#include <cstring>

#include "sf_common.hpp"

namespace sf {

static thread_local char g_last_error[512] = "";

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
    va_end(ap);
}

const char* last_error() { return g_last_error; }

void require_device() {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n < 1) {
        set_error("no HIP device available (%s); the soapfilm library runs only on gfx950 GPUs",
                  e == hipSuccess ? "0 devices" : hipGetErrorString(e));
        throw Error{SF_ERR_DEVICE};
    }
}

}  // namespace sf

namespace sf {

}  // namespace sf

static_assert(sizeof(sf_params) == 16 * 8 + 2 * 4, "sf_params layout");
static_assert(offsetof(sf_params, flags) == 128, "sf_params layout");
static_assert(sizeof(sf_stats) == 5 * 8 + 8 + 4 + 4 + 8, "sf_stats layout");
static_assert(sizeof(sf_camera) == 24, "sf_camera layout");

extern "C" const char* sf_last_error(void) { return sf::last_error(); }

extern "C" const char* sf_version(void) { return "soapfilm-mi355x 0.1.0 (gfx950)"; }

extern "C" int sf_device_available(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n > 0 ? 1 : 0;
}
