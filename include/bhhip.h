/* bhhip — the thin C-ABI layer between the hip vector engine's C++ host code
 * and the HIP runtime / hiprtc.  Plain pointers and sizes only.
 *
 * It replaces, for this engine, the device-facing calls the reference CUDA
 * engine makes directly on the CUDA driver API:
 *   bhhip_malloc / bhhip_free        <- cuMemAlloc / cuMemFree via MallocCache
 *                                       (ve/cuda/engine_cuda.hpp:87-96,116-131)
 *   bhhip_memcpy_h2d / _d2h          <- cuMemcpyHtoD / cuMemcpyDtoH
 *                                       (ve/cuda/engine_cuda.hpp:134-174)
 *   bhhip_compile                    <- nvcc subprocess + cuModuleLoad + cuModuleGetFunction
 *                                       (ve/cuda/engine_cuda.cpp:184-241)
 *   bhhip_launch                     <- cuLaunchKernel + cuCtxSynchronize
 *                                       (ve/cuda/engine_cuda.cpp:276-326)
 * and adds what the reference has no counterpart for: stream-ordered
 * execution without a per-launch synchronize, events for per-launch timing,
 * and graph capture/replay for repeated flushes.
 *
 * Every function returns 0 on success and a non-zero HIP/hiprtc status on
 * failure; bhhip_last_error() then describes the failure.  All work goes to
 * one HIP stream per process (default priority).
 */
#ifndef BHHIP_H
#define BHHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Select the device and create the engine stream.  Idempotent. */
int bhhip_init(int device);
const char *bhhip_last_error(void);
/* Human-readable device description (name, CUs, memory, arch). */
const char *bhhip_device_info(void);
/* Number of compute units and total device memory in bytes. */
int bhhip_device_props(int *num_cus, uint64_t *total_mem);

int bhhip_malloc(void **dptr, size_t bytes);
int bhhip_free(void *dptr);
/* Host<->device copies, ordered on the engine stream; the d2h call returns when the data is on the host. */
int bhhip_memcpy_h2d(void *dst, const void *src, size_t bytes);
int bhhip_memcpy_d2h(void *dst, const void *src, size_t bytes);
int bhhip_memcpy_d2d(void *dst, const void *src, size_t bytes);
/* Host-to-device copy from page-locked memory (bhhip_host_alloc), ordered on the engine stream;
 * returns at once.  The source must stay unchanged until an event recorded after it completes. */
int bhhip_memcpy_h2d_async(void *dst, const void *src, size_t bytes);
/* Zero device memory (used only outside graph capture, e.g. right after allocation). */
int bhhip_memzero(void *dptr, size_t bytes);
/* Page-locked host memory for staging copies. */
int bhhip_host_alloc(void **ptr, size_t bytes);
int bhhip_host_free(void *ptr);

/* Without a device: set the target architecture (e.g. "gfx950") for compile-only use. */
int bhhip_set_arch(const char *arch);

/* Compile `source` (prepended by `prelude`) for the current device with hiprtc
 * and return a handle to the kernel `name`.  Code objects are cached in memory
 * and, when `cache_dir` is non-empty, on disk keyed by a hash of the full
 * source and options.  On a compile error the hiprtc log is available from
 * bhhip_last_error().  With `function` == NULL the code object is only
 * produced (no device needed after bhhip_set_arch). */
int bhhip_compile(const char *prelude, const char *source, const char *name, const char *cache_dir,
                  void **function);
/* Number of hiprtc compilations and cache hits so far. */
void bhhip_compile_stats(uint64_t *compiles, uint64_t *memory_hits, uint64_t *disk_hits);

/* Launch `function` on the engine stream with a packed argument buffer
 * (each argument at its natural alignment). */
int bhhip_launch(void *function, uint64_t grid, uint32_t block, const void *args, size_t args_bytes);
/* The same with a 3-D grid (workgroup counts per dimension) and a 3-D workgroup shape; used for user
 * kernels, whose launch shape the caller chooses (reference ve/opencl/engine_opencl.cpp:632-662). */
int bhhip_launch_3d(void *function, const uint32_t grid[3], const uint32_t block[3], const void *args,
                    size_t args_bytes);
/* Launch `function` cooperatively (every workgroup of the grid resident at once, or an error) with
 * `params` pointing at each argument in turn; for kernels whose workgroups wait for each other. */
int bhhip_launch_cooperative(void *function, uint32_t grid, uint32_t block, void **params);
int bhhip_synchronize(void);
/* Workgroups of `block` threads of `function` that can be resident on one compute unit at once. */
int bhhip_occupancy(void *function, uint32_t block, int *blocks_per_cu);
/* The engine stream (a hipStream_t), for device code launched outside this layer (extension methods). */
void *bhhip_stream(void);

/* Timing: an event pair around stream work. */
int bhhip_event_create(void **event);
int bhhip_event_record(void *event);
int bhhip_event_elapsed_ms(void *start, void *stop, float *ms);
/* Without waiting: *done = 1 when the work recorded before `event` has finished, else 0. */
int bhhip_event_done(void *event, int *done);
/* Wait until the work recorded before `event` has finished. */
int bhhip_event_synchronize(void *event);
int bhhip_event_destroy(void *event);

/* Graph capture of the engine stream, instantiation and replay. */
int bhhip_capture_begin(void);
int bhhip_capture_end(void **graph_exec);
int bhhip_graph_launch(void *graph_exec);
int bhhip_graph_destroy(void *graph_exec);

#ifdef __cplusplus
}
#endif

#endif
