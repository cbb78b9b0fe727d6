// common.cpp -- thread-local last error and device selection for the C-ABI.
#include "common.hpp"

static thread_local std::string g_last_error;

void chili_set_error(const std::string& msg) { g_last_error = msg; }

void select_device(int device) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) throw chili_error(CHILI_ENODEV, "no HIP device available");
    if (device < 0 || device >= n) throw chili_error(CHILI_EINVAL, "device index out of range");
    HIP_CHECK(hipSetDevice(device));
}

extern "C" {
const char* chili_last_error(void) { return g_last_error.c_str(); }
int chili_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}
}
