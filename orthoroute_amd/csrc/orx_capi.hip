// orx_capi.hip -- library init, version and error reporting of liborx_hip.
#include "orx_common.hpp"
#include "../../include/orthoroute_hip.h"

#include <cstring>
#include <string>

namespace orx_host {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

int check_hip(hipError_t e, const char* what) {
    g_last_error = std::string(what) + ": " + hipGetErrorString(e);
    return ORX_E_HIP;
}

}  // namespace orx_host

extern "C" {

int orx_version(void) { return 10000; }

const char* orx_last_error(void) { return orx_host::g_last_error.c_str(); }

int orx_init(void) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n < 1) return orx_host::fail(ORX_E_NODEVICE, "no HIP device visible");
    int dev = 0;
    ORX_HIP_CHECK(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    ORX_HIP_CHECK(hipGetDeviceProperties(&prop, dev));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return orx_host::fail(ORX_E_NODEVICE, std::string("device is ") + prop.gcnArchName +
                                                   ", this library is built for gfx950 only");
    return ORX_OK;
}

}  // extern "C"
