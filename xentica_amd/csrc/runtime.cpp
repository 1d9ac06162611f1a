// Runtime half of libxentica_hip.so: device, hiprtc, modules, memory,
// streams, events and graphs behind the C ABI of include/xentica_hip.h.
// Replaces the PyCUDA layer of the reference (xentica/core/base.py:527-588).
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "xentica_hip.h"
#include "common.h"

namespace xen {

static thread_local std::string g_last_error;

int fail(const std::string &msg, int code) {
    g_last_error = msg;
    return code ? code : -1;
}

int check(hipError_t err, const char *what) {
    if (err == hipSuccess) return 0;
    return fail(std::string(what) + ": " + hipGetErrorString(err),
                static_cast<int>(err));
}

}  // namespace xen

using xen::check;
using xen::fail;

extern "C" {

const char *xen_last_error(void) { return xen::g_last_error.c_str(); }

int xen_init(int device) {
    int count = 0;
    if (int rc = check(hipGetDeviceCount(&count), "hipGetDeviceCount")) return rc;
    if (device < 0 || device >= count)
        return fail("xen_init: device " + std::to_string(device) +
                    " out of range (" + std::to_string(count) + " devices)", 1);
    if (int rc = check(hipSetDevice(device), "hipSetDevice")) return rc;
    return check(hipFree(nullptr), "hipFree(0) context init");
}

int xen_device_count(int *count) {
    return check(hipGetDeviceCount(count), "hipGetDeviceCount");
}

int xen_device_name(char *buf, int len) {
    int dev = 0;
    if (int rc = check(hipGetDevice(&dev), "hipGetDevice")) return rc;
    hipDeviceProp_t prop;
    if (int rc = check(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties"))
        return rc;
    std::snprintf(buf, len, "%s (%s, %d CUs)", prop.name, prop.gcnArchName,
                  prop.multiProcessorCount);
    return 0;
}

int xen_compile(const char *source, const char *name,
                const char *const *options, int num_options,
                void **code, size_t *size) {
    *code = nullptr;
    *size = 0;
    hiprtcProgram prog;
    hiprtcResult r = hiprtcCreateProgram(&prog, source, name, 0, nullptr, nullptr);
    if (r != HIPRTC_SUCCESS)
        return fail(std::string("hiprtcCreateProgram: ") + hiprtcGetErrorString(r), 2);
    r = hiprtcCompileProgram(prog, num_options, const_cast<const char **>(options));
    if (r != HIPRTC_SUCCESS) {
        size_t log_size = 0;
        hiprtcGetProgramLogSize(prog, &log_size);
        std::string log(log_size + 1, '\0');
        if (log_size) hiprtcGetProgramLog(prog, &log[0]);
        hiprtcDestroyProgram(&prog);
        return fail(std::string("hiprtc compile failed (") + hiprtcGetErrorString(r) +
                    "):\n" + log.c_str(), 3);
    }
    size_t code_size = 0;
    hiprtcGetCodeSize(prog, &code_size);
    void *buf = std::malloc(code_size);
    if (!buf) {
        hiprtcDestroyProgram(&prog);
        return fail("xen_compile: out of host memory", 4);
    }
    hiprtcGetCode(prog, static_cast<char *>(buf));
    hiprtcDestroyProgram(&prog);
    *code = buf;
    *size = code_size;
    return 0;
}

void xen_code_free(void *code) { std::free(code); }

int xen_module_load(const void *code, xen_module *module) {
    hipModule_t mod;
    if (int rc = check(hipModuleLoadData(&mod, code), "hipModuleLoadData")) return rc;
    *module = reinterpret_cast<xen_module>(mod);
    return 0;
}

int xen_module_unload(xen_module module) {
    return check(hipModuleUnload(reinterpret_cast<hipModule_t>(module)),
                 "hipModuleUnload");
}

int xen_module_function(xen_module module, const char *name, xen_function *function) {
    hipFunction_t fn;
    if (int rc = check(hipModuleGetFunction(&fn, reinterpret_cast<hipModule_t>(module), name),
                       (std::string("hipModuleGetFunction ") + name).c_str()))
        return rc;
    *function = reinterpret_cast<xen_function>(fn);
    return 0;
}

int xen_launch(xen_function function, unsigned grid_x, unsigned grid_y, unsigned grid_z,
               unsigned block_x, unsigned block_y, unsigned block_z,
               unsigned shared_bytes, xen_stream stream, void **args) {
    if (grid_x == 0 || grid_y == 0 || grid_z == 0) return 0;
    if (block_x * block_y * block_z > 1024)
        return fail("xen_launch: block larger than 1024 threads", 5);
    return check(hipModuleLaunchKernel(reinterpret_cast<hipFunction_t>(function),
                                       grid_x, grid_y, grid_z, block_x, block_y, block_z,
                                       shared_bytes, reinterpret_cast<hipStream_t>(stream),
                                       args, nullptr),
                 "hipModuleLaunchKernel");
}

int xen_occupancy(xen_function function, unsigned block, unsigned shared_bytes,
                  int *blocks_per_cu, int *cu_count) {
    int blocks = 0;
    if (int rc = check(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(
                           &blocks, reinterpret_cast<hipFunction_t>(function),
                           static_cast<int>(block), shared_bytes),
                       "hipModuleOccupancyMaxActiveBlocksPerMultiprocessor"))
        return rc;
    int device = 0, cus = 0;
    if (int rc = check(hipGetDevice(&device), "hipGetDevice")) return rc;
    if (int rc = check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount,
                                             device),
                       "hipDeviceGetAttribute"))
        return rc;
    *blocks_per_cu = blocks;
    *cu_count = cus;
    return 0;
}

int xen_malloc(void **ptr, size_t bytes) {
    *ptr = nullptr;
    if (bytes == 0) bytes = 1;
    return check(hipMalloc(ptr, bytes), "hipMalloc");
}

int xen_free(void *ptr) { return ptr ? check(hipFree(ptr), "hipFree") : 0; }

int xen_memcpy_h2d(void *dst, const void *src, size_t bytes, xen_stream s) {
    if (!bytes) return 0;
    auto st = reinterpret_cast<hipStream_t>(s);
    if (int rc = check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st),
                       "hipMemcpyAsync H2D"))
        return rc;
    return check(hipStreamSynchronize(st), "hipStreamSynchronize after H2D");
}

int xen_memcpy_d2h(void *dst, const void *src, size_t bytes, xen_stream s) {
    if (!bytes) return 0;
    auto st = reinterpret_cast<hipStream_t>(s);
    if (int rc = check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st),
                       "hipMemcpyAsync D2H"))
        return rc;
    return check(hipStreamSynchronize(st), "hipStreamSynchronize after D2H");
}

int xen_memcpy_d2d(void *dst, const void *src, size_t bytes, xen_stream s) {
    if (!bytes) return 0;
    return check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice,
                                reinterpret_cast<hipStream_t>(s)),
                 "hipMemcpyAsync D2D");
}

int xen_memset(void *dst, int value, size_t bytes, xen_stream s) {
    if (!bytes) return 0;
    return check(hipMemsetAsync(dst, value, bytes, reinterpret_cast<hipStream_t>(s)),
                 "hipMemsetAsync");
}

int xen_mem_info(size_t *free_bytes, size_t *total_bytes) {
    return check(hipMemGetInfo(free_bytes, total_bytes), "hipMemGetInfo");
}

int xen_stream_create(xen_stream *stream) {
    hipStream_t s;
    if (int rc = check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate"))
        return rc;
    *stream = reinterpret_cast<xen_stream>(s);
    return 0;
}

int xen_stream_destroy(xen_stream stream) {
    return check(hipStreamDestroy(reinterpret_cast<hipStream_t>(stream)), "hipStreamDestroy");
}

int xen_stream_sync(xen_stream stream) {
    return check(hipStreamSynchronize(reinterpret_cast<hipStream_t>(stream)),
                 "hipStreamSynchronize");
}

int xen_device_sync(void) { return check(hipDeviceSynchronize(), "hipDeviceSynchronize"); }

int xen_event_create(xen_event *event) {
    hipEvent_t e;
    if (int rc = check(hipEventCreate(&e), "hipEventCreate")) return rc;
    *event = reinterpret_cast<xen_event>(e);
    return 0;
}

int xen_event_destroy(xen_event event) {
    return check(hipEventDestroy(reinterpret_cast<hipEvent_t>(event)), "hipEventDestroy");
}

int xen_event_record(xen_event event, xen_stream stream) {
    return check(hipEventRecord(reinterpret_cast<hipEvent_t>(event),
                                reinterpret_cast<hipStream_t>(stream)),
                 "hipEventRecord");
}

int xen_event_sync(xen_event event) {
    return check(hipEventSynchronize(reinterpret_cast<hipEvent_t>(event)),
                 "hipEventSynchronize");
}

int xen_event_elapsed(xen_event start, xen_event stop, float *ms) {
    return check(hipEventElapsedTime(ms, reinterpret_cast<hipEvent_t>(start),
                                     reinterpret_cast<hipEvent_t>(stop)),
                 "hipEventElapsedTime");
}

int xen_graph_begin(xen_stream stream) {
    return check(hipStreamBeginCapture(reinterpret_cast<hipStream_t>(stream),
                                       hipStreamCaptureModeThreadLocal),
                 "hipStreamBeginCapture");
}

int xen_graph_end(xen_stream stream, xen_graph *graph) {
    hipGraph_t g;
    if (int rc = check(hipStreamEndCapture(reinterpret_cast<hipStream_t>(stream), &g),
                       "hipStreamEndCapture"))
        return rc;
    hipGraphExec_t exec;
    int rc = check(hipGraphInstantiate(&exec, g, nullptr, nullptr, 0), "hipGraphInstantiate");
    (void)hipGraphDestroy(g);
    if (rc) return rc;
    *graph = reinterpret_cast<xen_graph>(exec);
    return 0;
}

int xen_graph_launch(xen_graph graph, xen_stream stream) {
    return check(hipGraphLaunch(reinterpret_cast<hipGraphExec_t>(graph),
                                reinterpret_cast<hipStream_t>(stream)),
                 "hipGraphLaunch");
}

int xen_graph_destroy(xen_graph graph) {
    return check(hipGraphExecDestroy(reinterpret_cast<hipGraphExec_t>(graph)),
                 "hipGraphExecDestroy");
}

}  // extern "C"
