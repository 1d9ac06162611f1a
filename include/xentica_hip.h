/*
 * xentica_hip.h -- C ABI of the MI355X (gfx950) runtime of xentica_amd.
 *
 * The reference reaches the GPU through PyCUDA: SourceModule compiles the
 * generated source with nvcc and get_function() looks kernels up
 * (xentica/core/base.py:527-541, class GPUKernels); gpuarray.to_gpu /
 * zeros / .get() hold the cells, colours and image
 * (xentica/core/base.py:544-569, class GPUArrays); kernels are launched
 * positionally from step() and render() (xentica/core/base.py:703-753).
 * This library replaces that whole layer: hiprtc compiles the generated
 * HIP source for gfx950, the HIP runtime loads and launches it, and the
 * hand-written kernels of the fast engines are exported beside it.
 *
 * Conventions: every function returns 0 on success and a nonzero code on
 * failure (never throws across the ABI); xen_last_error() gives the
 * message of the calling thread's last failure.  Handles are opaque and
 * owned by the caller, who frees them explicitly.  Device pointers are
 * plain void pointers.  A stream handle of NULL means the default stream.
 */
#ifndef XENTICA_HIP_H
#define XENTICA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct xen_module_s *xen_module;
typedef struct xen_function_s *xen_function;
typedef struct xen_stream_s *xen_stream;
typedef struct xen_event_s *xen_event;
typedef struct xen_graph_s *xen_graph;

/* ---- device and errors (replaces pycuda.autoinit, base.py:94) ---- */
int xen_init(int device);
int xen_device_count(int *count);
int xen_device_name(char *buf, int len);
const char *xen_last_error(void);

/* ---- compilation (replaces SourceModule(src, ["-w"]), base.py:538) ---- */
/* Compile HIP source text for gfx950 with hiprtc.  On success *code holds
 * a malloc'ed code object of *size bytes (free with xen_code_free); on a
 * compile error the hiprtc log is in xen_last_error(). */
int xen_compile(const char *source, const char *name,
                const char *const *options, int num_options,
                void **code, size_t *size);
void xen_code_free(void *code);

/* ---- modules and kernels (replaces get_function, base.py:539-541) ---- */
int xen_module_load(const void *code, xen_module *module);
int xen_module_unload(xen_module module);
int xen_module_function(xen_module module, const char *name,
                        xen_function *function);
/* Launch with kernelParams semantics: args[i] points at argument i. */
int xen_launch(xen_function function,
               unsigned grid_x, unsigned grid_y, unsigned grid_z,
               unsigned block_x, unsigned block_y, unsigned block_z,
               unsigned shared_bytes, xen_stream stream, void **args);

/* Workgroups of `block` threads one CU holds at once for `function` (its
 * registers and LDS counted), and the device's CU count: the size of a
 * persistent grid.  No counterpart in the reference, whose kernels are
 * grid-stride launches sized by PyCUDA's GPUArray (base.py:713-721). */
int xen_occupancy(xen_function function, unsigned block, unsigned shared_bytes,
                  int *blocks_per_cu, int *cu_count);

/* ---- memory (replaces gpuarray.to_gpu / zeros / .get(), base.py:544-569) */
int xen_malloc(void **ptr, size_t bytes);
int xen_free(void *ptr);
int xen_memcpy_h2d(void *dst, const void *src, size_t bytes, xen_stream s);
int xen_memcpy_d2h(void *dst, const void *src, size_t bytes, xen_stream s);
int xen_memcpy_d2d(void *dst, const void *src, size_t bytes, xen_stream s);
int xen_memset(void *dst, int value, size_t bytes, xen_stream s);
int xen_mem_info(size_t *free_bytes, size_t *total_bytes);

/* ---- streams, events and graphs (no counterpart in the reference, which
 *      launches synchronously on PyCUDA's default stream) ---- */
int xen_stream_create(xen_stream *stream);
int xen_stream_destroy(xen_stream stream);
int xen_stream_sync(xen_stream stream);
int xen_device_sync(void);
int xen_event_create(xen_event *event);
int xen_event_destroy(xen_event event);
int xen_event_record(xen_event event, xen_stream stream);
int xen_event_sync(xen_event event);
int xen_event_elapsed(xen_event start, xen_event stop, float *ms);
int xen_graph_begin(xen_stream stream);
int xen_graph_end(xen_stream stream, xen_graph *graph);
int xen_graph_launch(xen_graph graph, xen_stream stream);
int xen_graph_destroy(xen_graph graph);

#ifdef __cplusplus
}
#endif

#endif /* XENTICA_HIP_H */
