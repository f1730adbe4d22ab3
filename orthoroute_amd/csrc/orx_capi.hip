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
    int rc = orx_host::prepare_lattice_kernels();
    if (!rc) rc = orx_host::prepare_tile_kernels();
    return rc;
}

// Launch-sequence capture: a private non-blocking stream records every orx_*
// call enqueued on it between orx_graph_begin and orx_graph_end into one
// executable graph.  Nothing recorded runs until orx_graph_launch.
int orx_graph_begin(orx_stream_t* capture_stream) {
    if (!capture_stream) return orx_host::fail(ORX_E_ARG, "NULL argument");
    hipStream_t ps = nullptr;
    ORX_HIP_CHECK(hipStreamCreateWithFlags(&ps, hipStreamNonBlocking));
    hipError_t e = hipStreamBeginCapture(ps, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        (void)hipStreamDestroy(ps);
        return orx_host::check_hip(e, "hipStreamBeginCapture");
    }
    *capture_stream = ps;
    return ORX_OK;
}

int orx_graph_end(orx_stream_t capture_stream, orx_graph_t* graph) {
    if (!capture_stream || !graph) return orx_host::fail(ORX_E_ARG, "NULL argument");
    hipStream_t ps = (hipStream_t)capture_stream;
    hipGraph_t g = nullptr;
    hipGraphExec_t exec = nullptr;
    hipError_t e = hipStreamEndCapture(ps, &g);
    if (e == hipSuccess) e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
    if (g) (void)hipGraphDestroy(g);
    (void)hipStreamDestroy(ps);
    if (e != hipSuccess) return orx_host::check_hip(e, "orx_graph_end");
    *graph = exec;
    return ORX_OK;
}

int orx_graph_launch(orx_graph_t graph, orx_stream_t stream) {
    if (!graph) return orx_host::fail(ORX_E_ARG, "NULL graph");
    ORX_HIP_CHECK(hipGraphLaunch((hipGraphExec_t)graph, (hipStream_t)stream));
    return ORX_OK;
}

int orx_graph_destroy(orx_graph_t graph) {
    if (!graph) return ORX_OK;
    ORX_HIP_CHECK(hipGraphExecDestroy((hipGraphExec_t)graph));
    return ORX_OK;
}

}  // extern "C"
