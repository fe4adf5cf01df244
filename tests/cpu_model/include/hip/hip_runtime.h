// CPU model of the HIP device language and runtime, for tests only.
//
// The library's own sources (csrc/*.hip, *.cpp) compile against this header
// with g++ instead of hipcc, so the same kernels and host runtime run on the
// CPU under AddressSanitizer / UndefinedBehaviorSanitizer:
//  - the threads of a workgroup run as fibers, one at a time, each until it
//    reaches __syncthreads or returns; a barrier is passed once every thread
//    still running has reached it (a thread that returned has left it, as a
//    finished wave leaves a hardware barrier);
//  - dynamic LDS is one array per kernel namespace, filled with a poison
//    pattern before every workgroup and ASan-poisoned past the launch's LDS
//    size, so reads of unwritten LDS give wrong results and reads or writes
//    past the allocation are reported;
//  - buffer-resource loads and stores check the offset against the
//    resource's extent (the hardware returns 0 / drops the store out of
//    range, which would hide the error) and count violations;
//  - 16-byte vector types carry their hardware alignment, so UBSan reports a
//    misaligned uint4 access;
//  - streams, events and graphs run synchronously on the calling thread;
//    kernel launches are serialised;
//  - copies whose device side is not inside one hipMalloc allocation fail.
// SFNTT_MODEL_CUS=<n> reports n compute units instead of 256.
// Fault injection: SFNTT_MODEL_FAIL="<what>:<n>[,<what>:<n>...]" makes the
// n-th (1-based) call of `what` fail, what in {malloc, host_malloc, launch,
// memcpy}.  sfn_model_* functions (below) read and reset counters.
// This file is not part of the library: nothing outside tests/ includes it.
#pragma once
#include <stddef.h>
#include <stdint.h>
#include <string.h>

#include <new>
#include <tuple>
#include <utility>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__
#define __align__(n) __attribute__((aligned(n)))

struct dim3 {
  uint32_t x, y, z;
  constexpr dim3(uint32_t x_ = 1, uint32_t y_ = 1, uint32_t z_ = 1) : x(x_), y(y_), z(z_) {}
};
extern thread_local dim3 threadIdx, blockIdx, blockDim, gridDim;

struct uint2 {
  uint32_t x, y;
} __attribute__((aligned(8)));
struct uint4 {
  uint32_t x, y, z, w;
} __attribute__((aligned(16)));
inline uint2 make_uint2(uint32_t x, uint32_t y) { return uint2{x, y}; }
inline uint4 make_uint4(uint32_t x, uint32_t y, uint32_t z, uint32_t w) { return uint4{x, y, z, w}; }

inline uint32_t min(uint32_t a, uint32_t b) { return a < b ? a : b; }
inline uint32_t max(uint32_t a, uint32_t b) { return a > b ? a : b; }
inline uint32_t __umulhi(uint32_t a, uint32_t b) { return (uint32_t)(((uint64_t)a * b) >> 32); }
inline uint32_t __brev(uint32_t x) {
  uint32_t r = 0;
  for (int i = 0; i < 32; ++i) r |= ((x >> i) & 1u) << (31 - i);
  return r;
}
template <class T>
inline T __ldg(const T *p) { return *p; }

void __syncthreads();

// ---- buffer resources (raw, stride 0: num_records counts bytes) ----
struct sfn_model_rsrc {
  const unsigned char *base;
  uint32_t bytes;
};
#define __amdgpu_buffer_rsrc_t sfn_model_rsrc
#define __builtin_amdgcn_make_buffer_rsrc sfn_model_make_rsrc
#define __builtin_amdgcn_raw_buffer_load_b32 sfn_model_buffer_load
#define __builtin_amdgcn_raw_buffer_store_b32 sfn_model_buffer_store
#define __builtin_amdgcn_sched_barrier(x) ((void)(x))
inline sfn_model_rsrc sfn_model_make_rsrc(void *base, short stride, int num, int flags) {
  (void)stride;
  (void)flags;
  return sfn_model_rsrc{(const unsigned char *)base, (uint32_t)num};
}
uint32_t sfn_model_buffer_load(sfn_model_rsrc r, int voff, int soff, int aux);
void sfn_model_buffer_store(uint32_t v, sfn_model_rsrc r, int voff, int soff, int aux);

// ---- runtime API ----
typedef enum {
  hipSuccess = 0,
  hipErrorInvalidValue = 1,
  hipErrorOutOfMemory = 2,
  hipErrorMemoryAllocation = 2,
  hipErrorInvalidConfiguration = 9,
  hipErrorInvalidDevice = 101,
  hipErrorNoDevice = 100,
  hipErrorLaunchFailure = 719,
  hipErrorUnknown = 999,
} hipError_t;

typedef enum {
  hipMemcpyHostToHost = 0,
  hipMemcpyHostToDevice = 1,
  hipMemcpyDeviceToHost = 2,
  hipMemcpyDeviceToDevice = 3,
  hipMemcpyDefault = 4
} hipMemcpyKind;
typedef enum { hipFuncAttributeMaxDynamicSharedMemorySize = 8 } hipFuncAttribute;
typedef enum { hipStreamCaptureModeGlobal = 0, hipStreamCaptureModeThreadLocal = 1 } hipStreamCaptureMode;
#define hipStreamNonBlocking 0x1
#define hipEventDisableTiming 0x2
#define hipHostMallocDefault 0x0

struct sfn_model_stream;
struct sfn_model_event;
struct sfn_model_graph;
typedef sfn_model_stream *hipStream_t;
typedef sfn_model_event *hipEvent_t;
typedef sfn_model_graph *hipGraph_t;
typedef sfn_model_graph *hipGraphExec_t;
typedef void *hipGraphNode_t;

typedef struct {
  char name[256];
  char gcnArchName[256];
  int multiProcessorCount;
  size_t maxSharedMemoryPerMultiProcessor;
  size_t totalGlobalMem;
  int clockRate;
  int memoryClockRate;
  int memoryBusWidth;
} hipDeviceProp_t;

hipError_t hipGetDeviceCount(int *count);
hipError_t hipSetDevice(int device);
hipError_t hipGetDeviceProperties(hipDeviceProp_t *p, int device);
hipError_t hipGetLastError();
const char *hipGetErrorString(hipError_t e);
hipError_t hipDeviceSynchronize();
hipError_t hipMalloc(void **p, size_t bytes);
template <class T>
inline hipError_t hipMalloc(T **p, size_t bytes) { return hipMalloc((void **)p, bytes); }
hipError_t hipFree(void *p);
hipError_t hipHostMalloc(void **p, size_t bytes, unsigned flags);
hipError_t hipHostFree(void *p);
hipError_t hipMemcpy(void *dst, const void *src, size_t bytes, hipMemcpyKind k);
hipError_t hipMemcpyAsync(void *dst, const void *src, size_t bytes, hipMemcpyKind k, hipStream_t s);
hipError_t hipMemsetAsync(void *dst, int v, size_t bytes, hipStream_t s);
hipError_t hipMemcpy2DAsync(void *dst, size_t dpitch, const void *src, size_t spitch, size_t width, size_t height,
                            hipMemcpyKind k, hipStream_t s);
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned flags);
hipError_t hipStreamDestroy(hipStream_t s);
hipError_t hipStreamSynchronize(hipStream_t s);
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned flags);
hipError_t hipEventCreate(hipEvent_t *e);
hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned flags);
hipError_t hipEventDestroy(hipEvent_t e);
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s);
hipError_t hipEventSynchronize(hipEvent_t e);
hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b);
hipError_t hipFuncSetAttribute(const void *f, hipFuncAttribute attr, int value);
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode mode);
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t *g);
hipError_t hipGraphInstantiate(hipGraphExec_t *e, hipGraph_t g, hipGraphNode_t *err_node, char *log, size_t n);
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t s);
hipError_t hipGraphDestroy(hipGraph_t g);
hipError_t hipGraphExecDestroy(hipGraphExec_t e);
#define HIP_SYMBOL(x) ((const void *)&(x))
hipError_t hipGetSymbolAddress(void **p, const void *symbol);

// ---- kernel launch ----
namespace sfn_model {
using Body = void (*)(void *);
// Runs body(ctx) for every thread of every workgroup of the grid, or records it
// when s is being captured.  Returns the launch status (also kept for
// hipGetLastError).
hipError_t launch(const void *kernel, dim3 grid, dim3 block, size_t lds, hipStream_t s, Body body, void *ctx,
                  void (*copy)(void *dst, const void *src), void (*destroy)(void *), size_t ctx_bytes);

template <class... P>
struct Closure {
  void (*k)(P...);
  std::tuple<P...> args;
};
template <class... P>
void run_closure(void *c) {
  auto *cl = static_cast<Closure<P...> *>(c);
  std::apply(cl->k, cl->args);
}
template <class... P>
void copy_closure(void *dst, const void *src) {
  new (dst) Closure<P...>(*static_cast<const Closure<P...> *>(src));
}
template <class... P>
void destroy_closure(void *c) {
  static_cast<Closure<P...> *>(c)->~Closure<P...>();
}
template <class... P, class... A>
hipError_t launch_kernel(void (*k)(P...), dim3 grid, dim3 block, size_t lds, hipStream_t s, A &&...a) {
  Closure<P...> cl{k, std::tuple<P...>(static_cast<P>(a)...)};
  return launch((const void *)k, grid, block, lds, s, &run_closure<P...>, &cl, &copy_closure<P...>,
                &destroy_closure<P...>, sizeof(cl));
}
}  // namespace sfn_model

#define hipLaunchKernelGGL(k, grid, block, lds, stream, ...) \
  ((void)sfn_model::launch_kernel(k, dim3(grid), dim3(block), (size_t)(lds), stream, __VA_ARGS__))

// ---- test hooks (C linkage, read by the test runner over ctypes) ----
extern "C" {
// buffer-resource accesses past the resource's extent since the last reset
uint64_t sfn_model_oob_count();
// kernel launches since the last reset; hipMalloc / hipHostMalloc calls ever
uint64_t sfn_model_launch_count();
uint64_t sfn_model_malloc_count();
// bytes held by live hipMalloc allocations, and their peak since the last reset
uint64_t sfn_model_device_bytes();
uint64_t sfn_model_peak_device_bytes();
void sfn_model_reset_counters();
// re-reads SFNTT_MODEL_FAIL and restarts the call counters it refers to
void sfn_model_set_faults(const char *spec);
}
