// Error state and device selection shared by every entry point.
#include "common.hpp"

#include <cstdio>

namespace sysmat {

static thread_local std::string g_error;

int set_error(int code, const std::string& msg) {
    g_error = msg;
    return code;
}

void clear_error() { g_error.clear(); }

int use_device(int device) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return set_error(SM_E_DEVICE, "no HIP device visible");
    if (device < 0 || device >= n) return set_error(SM_E_DEVICE, "device id out of range");
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return set_error(SM_E_DEVICE, std::string("hipSetDevice: ") + hipGetErrorString(e));
    return 0;
}

}  // namespace sysmat

extern "C" {

int sm_version(void) { return 100; }

const char* sm_last_error(void) { return sysmat::g_error.c_str(); }

int sm_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

}  // extern "C"
