// Device, stream, memory and event layer of the thin C ABI (qgtl_hip_abi.h).
// One process drives one GPU through one non-blocking stream of default
// priority; every copy of the library is ordered on it, and so is every
// kernel unless a caller routes launches to the second stream (lane 1, also
// of default priority) for work that may overlap, ordered by events.
#include <stdio.h>
#include <string.h>

#include "qgtl_hip_common.h"

namespace {
hipStream_t g_stream = nullptr;
hipStream_t g_side = nullptr;
int g_lane = 0;
int g_device = -1;
thread_local char g_msg[256] = "";
unsigned long long g_allocs = 0;   // hipMalloc / hipHostMalloc / hipFree / hipHostFree calls made

int code_for(hipError_t e) {
    switch (e) {
    case hipErrorOutOfMemory: return -119;          // QGT_ERROR_GPU_OUT_OF_MEMORY
    case hipErrorInvalidValue: return -118;         // QGT_ERROR_GPU_INVALID_VALUE
    case hipErrorLaunchFailure:
    case hipErrorInvalidConfiguration: return -121; // QGT_ERROR_GPU_LAUNCH_FAILED
    default: return -120;                           // QGT_ERROR_GPU_INTERNAL
    }
}
}  // namespace

namespace qgtl {
hipStream_t stream() { return g_lane ? g_side : g_stream; }

int set_error(hipError_t e, const char *what) {
    snprintf(g_msg, sizeof g_msg, "%s: %s", what, hipGetErrorString(e));
    return code_for(e);
}

int set_error_msg(int code, const char *msg) {
    snprintf(g_msg, sizeof g_msg, "%s", msg);
    return code;
}
}  // namespace qgtl

extern "C" {

const char *qgtl_hip_last_error(void) { return g_msg; }

int qgtl_hip_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int qgtl_hip_init(int device) {
    if (g_stream && device == g_device) return 0;
    int count = 0;
    QGTL_HIP_CHECK(hipGetDeviceCount(&count));
    if (count <= 0) return qgtl::set_error_msg(-6, "no HIP device visible");
    if (device < 0 || device >= count) return qgtl::set_error_msg(-118, "device index out of range");
    QGTL_HIP_CHECK(hipSetDevice(device));
    if (g_stream) {
        (void)hipStreamDestroy(g_stream);
        (void)hipStreamDestroy(g_side);
        g_stream = g_side = nullptr;
    }
    g_lane = 0;
    QGTL_HIP_CHECK(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
    QGTL_HIP_CHECK(hipStreamCreateWithFlags(&g_side, hipStreamNonBlocking));
    g_device = device;
    return qgtl_hip_noalloc_init();   // the no-allocation solve's scratch address on this device
}

void *qgtl_hip_stream(void) { return (void *)g_stream; }

int qgtl_hip_malloc(void **ptr, size_t bytes) {
    *ptr = nullptr;
    if (bytes == 0) return 0;
    g_allocs++;
    QGTL_HIP_CHECK(hipMalloc(ptr, bytes));
    return 0;
}

int qgtl_hip_free(void *ptr) {
    if (!ptr) return 0;
    g_allocs++;
    QGTL_HIP_CHECK(hipFree(ptr));
    return 0;
}

unsigned long long qgtl_hip_alloc_calls(void) { return g_allocs; }

int qgtl_hip_mem_free(size_t *free_bytes) {
    size_t total = 0;
    *free_bytes = 0;
    QGTL_HIP_CHECK(hipMemGetInfo(free_bytes, &total));
    return 0;
}

int qgtl_hip_host_alloc(void **ptr, size_t bytes) {
    *ptr = nullptr;
    if (bytes == 0) return 0;
    g_allocs++;
    QGTL_HIP_CHECK(hipHostMalloc(ptr, bytes, hipHostMallocDefault));
    return 0;
}

int qgtl_hip_host_free(void *ptr) {
    if (!ptr) return 0;
    g_allocs++;
    QGTL_HIP_CHECK(hipHostFree(ptr));
    return 0;
}

int qgtl_hip_h2d(void *dst, const void *src, size_t bytes) {
    if (bytes) QGTL_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, g_stream));
    return 0;
}

int qgtl_hip_d2h(void *dst, const void *src, size_t bytes) {
    if (bytes) QGTL_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, g_stream));
    return 0;
}

int qgtl_hip_d2d(void *dst, const void *src, size_t bytes) {
    if (bytes) QGTL_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, g_stream));
    return 0;
}

int qgtl_hip_sync(void) {
    QGTL_HIP_CHECK(hipStreamSynchronize(g_stream));
    QGTL_HIP_CHECK(hipGetLastError());
    return 0;
}

struct qgtl_hip_event_s { hipEvent_t e; };

int qgtl_hip_event_create(qgtl_hip_event_t *ev) {
    *ev = new qgtl_hip_event_s;
    QGTL_HIP_CHECK(hipEventCreate(&(*ev)->e));
    return 0;
}

int qgtl_hip_event_record(qgtl_hip_event_t ev) {
    QGTL_HIP_CHECK(hipEventRecord(ev->e, g_stream));
    return 0;
}

int qgtl_hip_event_create_sync(qgtl_hip_event_t *ev) {
    *ev = new qgtl_hip_event_s;
    QGTL_HIP_CHECK(hipEventCreateWithFlags(&(*ev)->e, hipEventDisableTiming));
    return 0;
}

int qgtl_hip_set_lane(int lane) {
    if (lane != 0 && lane != 1) return qgtl::set_error_msg(-118, "qgtl_hip_set_lane: lane is 0 or 1");
    g_lane = lane;
    return 0;
}

int qgtl_hip_lane_signal(qgtl_hip_event_t ev) {
    QGTL_HIP_CHECK(hipEventRecord(ev->e, qgtl::stream()));
    return 0;
}

int qgtl_hip_lane_wait(qgtl_hip_event_t ev) {
    QGTL_HIP_CHECK(hipStreamWaitEvent(qgtl::stream(), ev->e, 0));
    return 0;
}

int qgtl_hip_event_elapsed_ms(qgtl_hip_event_t a, qgtl_hip_event_t b, float *ms) {
    QGTL_HIP_CHECK(hipEventElapsedTime(ms, a->e, b->e));
    return 0;
}

int qgtl_hip_event_destroy(qgtl_hip_event_t ev) {
    if (!ev) return 0;
    (void)hipEventDestroy(ev->e);
    delete ev;
    return 0;
}

struct qgtl_hip_graph_s {
    hipGraph_t graph;
    hipGraphExec_t exec;
};

int qgtl_hip_capture_begin(void) {
    QGTL_HIP_CHECK(hipStreamBeginCapture(g_stream, hipStreamCaptureModeThreadLocal));
    return 0;
}

int qgtl_hip_capture_end(qgtl_hip_graph_t *out) {
    *out = nullptr;
    hipGraph_t g = nullptr;
    QGTL_HIP_CHECK(hipStreamEndCapture(g_stream, &g));
    hipGraphExec_t x = nullptr;
    hipError_t e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        return qgtl::set_error(e, "hipGraphInstantiate");
    }
    *out = new qgtl_hip_graph_s{g, x};
    return 0;
}

// End a capture into *graph: an existing graph's executable is updated in
// place when the new sequence has the same shape (kernel arguments changed),
// else it is replaced.
int qgtl_hip_capture_end_update(qgtl_hip_graph_t *graph) {
    hipGraph_t g = nullptr;
    QGTL_HIP_CHECK(hipStreamEndCapture(g_stream, &g));
    if (*graph) {
        hipGraphNode_t bad = nullptr;
        hipGraphExecUpdateResult res = hipGraphExecUpdateError;
        if (hipGraphExecUpdate((*graph)->exec, g, &bad, &res) == hipSuccess && res == hipGraphExecUpdateSuccess) {
            (void)hipGraphDestroy((*graph)->graph);
            (*graph)->graph = g;
            return 0;
        }
        (void)hipGetLastError();
        qgtl_hip_graph_destroy(*graph);
        *graph = nullptr;
    }
    hipGraphExec_t x = nullptr;
    hipError_t e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        return qgtl::set_error(e, "hipGraphInstantiate");
    }
    *graph = new qgtl_hip_graph_s{g, x};
    return 0;
}

// Abandon a capture that failed part way (the stream leaves capture mode).
void qgtl_hip_capture_abort(void) {
    hipGraph_t g = nullptr;
    if (hipStreamEndCapture(g_stream, &g) == hipSuccess && g) (void)hipGraphDestroy(g);
    (void)hipGetLastError();
}

int qgtl_hip_graph_launch(qgtl_hip_graph_t graph) {
    QGTL_HIP_CHECK(hipGraphLaunch(graph->exec, g_stream));
    return 0;
}

void qgtl_hip_graph_destroy(qgtl_hip_graph_t graph) {
    if (!graph) return;
    (void)hipGraphExecDestroy(graph->exec);
    (void)hipGraphDestroy(graph->graph);
    delete graph;
}

int qgtl_hip_is_device_ptr(const void *p) {
    if (!p) return 0;
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();   // an unregistered host pointer: not an error of the caller
        return 0;
    }
    return a.type == hipMemoryTypeDevice ? 1 : 0;
}

// Fills on the library stream with a kernel-free HIP call; used only outside
// captured graphs (the compute-backend memset slot).
int qgtl_hip_memset(void *dst, int value, size_t bytes) {
    if (bytes) QGTL_HIP_CHECK(hipMemsetAsync(dst, value, bytes, g_stream));
    return 0;
}

int qgtl_hip_copy(void *dst, const void *src, size_t bytes) {
    if (bytes) QGTL_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, g_stream));
    return 0;
}

int qgtl_hip_device_total_mem(size_t *bytes) {
    size_t free_b = 0;
    *bytes = 0;
    QGTL_HIP_CHECK(hipMemGetInfo(&free_b, bytes));
    return 0;
}

int qgtl_hip_cu_count(void) {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    return cus;
}

}  // extern "C"
