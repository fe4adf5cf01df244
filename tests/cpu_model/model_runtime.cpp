// CPU model of the HIP runtime (see include/hip/hip_runtime.h).  Test
// infrastructure: linked only into the sanitizer build of the library that
// tests/cpu_model/run_model.py drives.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <random>
#include <set>
#include <string>
#include <thread>
#include <vector>

#include <sys/mman.h>
#include <ucontext.h>

#if defined(__SANITIZE_ADDRESS__)
#include <sanitizer/asan_interface.h>
#define MODEL_POISON(p, n) ASAN_POISON_MEMORY_REGION(p, n)
#define MODEL_UNPOISON(p, n) ASAN_UNPOISON_MEMORY_REGION(p, n)
#else
#define MODEL_POISON(p, n) ((void)(p), (void)(n))
#define MODEL_UNPOISON(p, n) ((void)(p), (void)(n))
#endif

thread_local dim3 threadIdx, blockIdx, blockDim, gridDim;

// Dynamic LDS of the kernels: one array per namespace that declares
// `extern __shared__ smem[]` (kernels.hip: sfn, fast.hip: sfn::fast).
static constexpr size_t kLdsBytes = 160 * 1024;
namespace sfn {
alignas(64) unsigned char smem[kLdsBytes];
namespace fast {
alignas(64) unsigned char smem[kLdsBytes];
}
}  // namespace sfn

namespace {

std::atomic<uint64_t> g_oob{0}, g_launches{0}, g_mallocs{0};
std::mutex g_launch_mu;  // one kernel at a time
thread_local hipError_t t_last = hipSuccess;

// ---- fault injection ----
std::mutex g_fault_mu;
std::map<std::string, uint64_t> g_fault_at, g_fault_calls;

bool inject(const char *what) {
  std::lock_guard<std::mutex> lk(g_fault_mu);
  auto it = g_fault_at.find(what);
  if (it == g_fault_at.end()) return false;
  return ++g_fault_calls[what] == it->second;
}

void parse_faults(const char *spec) {
  std::lock_guard<std::mutex> lk(g_fault_mu);
  g_fault_at.clear();
  g_fault_calls.clear();
  if (!spec) return;
  std::string s(spec);
  size_t pos = 0;
  while (pos < s.size()) {
    size_t end = s.find(',', pos);
    if (end == std::string::npos) end = s.size();
    std::string item = s.substr(pos, end - pos);
    size_t c = item.find(':');
    if (c != std::string::npos) g_fault_at[item.substr(0, c)] = strtoull(item.c_str() + c + 1, nullptr, 10);
    pos = end + 1;
  }
}

struct FaultInit {
  FaultInit() { parse_faults(getenv("SFNTT_MODEL_FAIL")); }
} g_fault_init;

// ---- fiber order ----
// SFNTT_MODEL_ORDER=reverse runs a workgroup's threads last to first between
// barriers; shuffle[:seed] draws a new order for every sweep.  A kernel whose
// threads exchange LDS data without a barrier in between, or overwrite data
// another thread still has to read, then gives different results than in
// thread order.
struct FiberOrder {
  int mode = 0;  // 0 thread order, 1 reverse, 2 shuffle
  uint64_t seed = 1;
  FiberOrder() {
    const char *o = getenv("SFNTT_MODEL_ORDER");
    if (!o) return;
    std::string s(o);
    if (s == "reverse") mode = 1;
    if (s.rfind("shuffle", 0) == 0) {
      mode = 2;
      if (s.size() > 8 && s[7] == ':') seed = strtoull(s.c_str() + 8, nullptr, 10);
    }
  }
} g_fiber_order;

// ---- device memory ----
std::mutex g_mem_mu;
std::map<const void *, size_t> g_allocs;
uint64_t g_dev_bytes = 0, g_dev_peak = 0;

// ---- workgroup threads as fibers ----
// The threads of one workgroup run as fibers on the launching host thread:
// each runs until it reaches __syncthreads or returns, then the next one
// resumes; a barrier is passed once every fiber still running has reached it
// (a fiber that returned has left the barrier, as a finished wave does).
// AddressSanitizer is told about every stack switch.
#if defined(__SANITIZE_ADDRESS__)
#include <sanitizer/common_interface_defs.h>
#define FIBER_START(save, bottom, size) __sanitizer_start_switch_fiber(save, bottom, size)
#define FIBER_FINISH(save, bottom, size) __sanitizer_finish_switch_fiber(save, bottom, size)
#else
#define FIBER_START(save, bottom, size) ((void)0)
#define FIBER_FINISH(save, bottom, size) ((void)0)
#endif

constexpr size_t kFiberStack = 256 * 1024;

struct Fiber {
  ucontext_t ctx;
  void *stack = nullptr;
  void *fake_stack = nullptr;
  bool done = false;
  uint32_t barriers = 0;
  dim3 tid;
};

struct Sched {
  ucontext_t main;
  const void *main_bottom = nullptr;
  size_t main_size = 0;
  std::vector<Fiber> fibers;
  Fiber *cur = nullptr;
  sfn_model::Body body = nullptr;
  void *arg = nullptr;
};
thread_local Sched *t_sched = nullptr;

void fiber_entry() {
  Sched *s = t_sched;
  FIBER_FINISH(nullptr, &s->main_bottom, &s->main_size);
  s->body(s->arg);
  s->cur->done = true;
  FIBER_START(nullptr, s->main_bottom, s->main_size);  // this fiber's stack ends here
  setcontext(&s->main);
}

std::mutex g_attr_mu;
std::set<const void *> g_big_lds;  // kernels allowed more than 64 KiB of dynamic LDS

void set_err(hipError_t e) {
  if (e != hipSuccess) t_last = e;
}

}  // namespace

struct sfn_model_graph {
  std::vector<std::function<hipError_t()>> ops;
};
struct sfn_model_stream {
  sfn_model_graph *capture = nullptr;
};
struct sfn_model_event {
  std::chrono::steady_clock::time_point t{};
};

void __syncthreads() {
  Sched *s = t_sched;
  Fiber *f = s->cur;
  ++f->barriers;
  FIBER_START(&f->fake_stack, s->main_bottom, s->main_size);
  swapcontext(&f->ctx, &s->main);
  FIBER_FINISH(f->fake_stack, &s->main_bottom, &s->main_size);
}

uint32_t sfn_model_buffer_load(sfn_model_rsrc r, int voff, int soff, int aux) {
  (void)aux;
  const uint64_t off = (uint64_t)(uint32_t)voff + (uint64_t)(uint32_t)soff;
  if (off + 4 > r.bytes) {
    if (g_oob.fetch_add(1) < 8)
      fprintf(stderr, "[cpu model] buffer load out of range: offset %llu, extent %u (workgroup %u, thread %u)\n",
              (unsigned long long)off, r.bytes, blockIdx.x, threadIdx.x);
    return 0;
  }
  uint32_t v;
  memcpy(&v, r.base + off, 4);
  return v;
}

void sfn_model_buffer_store(uint32_t v, sfn_model_rsrc r, int voff, int soff, int aux) {
  (void)aux;
  const uint64_t off = (uint64_t)(uint32_t)voff + (uint64_t)(uint32_t)soff;
  if (off + 4 > r.bytes) {
    if (g_oob.fetch_add(1) < 8)
      fprintf(stderr, "[cpu model] buffer store out of range: offset %llu, extent %u (workgroup %u, thread %u)\n",
              (unsigned long long)off, r.bytes, blockIdx.x, threadIdx.x);
    return;
  }
  memcpy((unsigned char *)r.base + off, &v, 4);
}

namespace {

hipError_t run_grid(const void *kernel, dim3 grid, dim3 block, size_t lds, sfn_model::Body body, void *ctx) {
  const uint64_t nthreads = (uint64_t)block.x * block.y * block.z;
  const uint64_t ngroups = (uint64_t)grid.x * grid.y * grid.z;
  if (nthreads == 0 || nthreads > 1024 || lds > kLdsBytes || grid.x > 0x7FFFFFFFu) return hipErrorInvalidConfiguration;
  if (lds > 64 * 1024) {
    std::lock_guard<std::mutex> lk(g_attr_mu);
    if (!g_big_lds.count(kernel)) return hipErrorInvalidConfiguration;
  }
  if (ngroups == 0) return hipSuccess;
  std::lock_guard<std::mutex> lk(g_launch_mu);
  g_launches.fetch_add(1);
  unsigned char *lds_arrays[2] = {sfn::smem, sfn::fast::smem};
  static std::vector<void *> stacks;  // reused across launches (under g_launch_mu)
  while (stacks.size() < nthreads) {
    void *st = mmap(nullptr, kFiberStack, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE, -1, 0);
    if (st == MAP_FAILED) return hipErrorOutOfMemory;
    stacks.push_back(st);
  }
  Sched sched;
  sched.body = body;
  sched.arg = ctx;
  sched.fibers.resize(nthreads);
  Sched *prev = t_sched;
  t_sched = &sched;
  const dim3 prev_idx = threadIdx, prev_blk = blockIdx, prev_dim = blockDim, prev_grid = gridDim;
  blockDim = block;
  gridDim = grid;
  for (uint64_t g = 0; g < ngroups; ++g) {
    for (unsigned char *a : lds_arrays) {
      MODEL_UNPOISON(a, kLdsBytes);
      memset(a, 0xCD, lds);  // unwritten LDS reads as garbage
      MODEL_POISON(a + ((lds + 7) & ~(size_t)7), kLdsBytes - ((lds + 7) & ~(size_t)7));
    }
    blockIdx = dim3((uint32_t)(g % grid.x), (uint32_t)((g / grid.x) % grid.y), (uint32_t)(g / ((uint64_t)grid.x * grid.y)));
    for (uint32_t t = 0; t < nthreads; ++t) {
      Fiber &f = sched.fibers[t];
      f.done = false;
      f.barriers = 0;
      f.fake_stack = nullptr;
      f.tid = dim3(t % block.x, (t / block.x) % block.y, t / (block.x * block.y));
      f.stack = stacks[t];
      getcontext(&f.ctx);
      f.ctx.uc_stack.ss_sp = f.stack;
      f.ctx.uc_stack.ss_size = kFiberStack;
      f.ctx.uc_link = nullptr;
      makecontext(&f.ctx, fiber_entry, 0);
    }
    // sweep until every fiber has returned; each sweep takes every fiber
    // still running to its next barrier, in the order SFNTT_MODEL_ORDER asks
    std::vector<uint32_t> order(nthreads);
    std::mt19937_64 rng(g_fiber_order.seed * 0x9E3779B97F4A7C15ull + g_launches.load() * 1315423911ull + g);
    for (bool any = true; any;) {
      any = false;
      for (uint32_t t = 0; t < nthreads; ++t) order[t] = g_fiber_order.mode == 1 ? (uint32_t)(nthreads - 1 - t) : t;
      if (g_fiber_order.mode == 2) std::shuffle(order.begin(), order.end(), rng);
      for (uint32_t k = 0; k < nthreads; ++k) {
        const uint32_t t = order[k];
        Fiber &f = sched.fibers[t];
        if (f.done) continue;
        any = true;
        sched.cur = &f;
        threadIdx = f.tid;
        void *main_fake = nullptr;
        FIBER_START(&main_fake, f.stack, kFiberStack);
        swapcontext(&sched.main, &f.ctx);
        FIBER_FINISH(main_fake, nullptr, nullptr);
      }
    }
  }
  t_sched = prev;
  threadIdx = prev_idx;
  blockIdx = prev_blk;
  blockDim = prev_dim;
  gridDim = prev_grid;
  for (unsigned char *a : lds_arrays) MODEL_UNPOISON(a, kLdsBytes);
  return hipSuccess;
}

hipError_t run_or_record(hipStream_t s, std::function<hipError_t()> op) {
  if (s && s->capture) {
    s->capture->ops.push_back(std::move(op));
    return hipSuccess;
  }
  return op();
}

bool known_device(const void *p, size_t bytes) {
  std::lock_guard<std::mutex> lk(g_mem_mu);
  auto it = g_allocs.upper_bound(p);
  if (it == g_allocs.begin()) return false;
  --it;
  const char *b = (const char *)it->first;
  return (const char *)p >= b && (const char *)p + bytes <= b + it->second;
}

}  // namespace

namespace sfn_model {
hipError_t launch(const void *kernel, dim3 grid, dim3 block, size_t lds, hipStream_t s, Body body, void *ctx,
                  void (*copy)(void *, const void *), void (*destroy)(void *), size_t ctx_bytes) {
  if (inject("launch")) {
    set_err(hipErrorLaunchFailure);
    return hipErrorLaunchFailure;
  }
  hipError_t e;
  if (s && s->capture) {
    // keep a copy of the arguments for the replays
    std::shared_ptr<unsigned char> keep(new unsigned char[ctx_bytes], [destroy](unsigned char *p) {
      destroy(p);
      delete[] p;
    });
    copy(keep.get(), ctx);
    s->capture->ops.push_back([=]() { return run_grid(kernel, grid, block, lds, body, keep.get()); });
    e = hipSuccess;
  } else {
    e = run_grid(kernel, grid, block, lds, body, ctx);
  }
  set_err(e);
  return e;
}
}  // namespace sfn_model

// ---------------------------------------------------------------- runtime API
hipError_t hipGetDeviceCount(int *count) {
  *count = 1;
  return hipSuccess;
}
hipError_t hipSetDevice(int device) { return device == 0 ? hipSuccess : hipErrorInvalidDevice; }
hipError_t hipGetDeviceProperties(hipDeviceProp_t *p, int device) {
  if (device != 0) return hipErrorInvalidDevice;
  memset(p, 0, sizeof(*p));
  strcpy(p->name, "CPU model of gfx950");
  strcpy(p->gcnArchName, "gfx950:sramecc+:xnack-");
  // SFNTT_MODEL_CUS: fewer compute units, so persistent launches (one
  // workgroup per resident slot) loop over several tiles at small sizes
  const char *cus = getenv("SFNTT_MODEL_CUS");
  p->multiProcessorCount = cus && atoi(cus) > 0 ? atoi(cus) : 256;
  p->maxSharedMemoryPerMultiProcessor = kLdsBytes;
  p->totalGlobalMem = (size_t)288 << 30;
  p->clockRate = 2400000;
  return hipSuccess;
}
hipError_t hipGetLastError() {
  hipError_t e = t_last;
  t_last = hipSuccess;
  return e;
}
const char *hipGetErrorString(hipError_t e) {
  switch (e) {
    case hipSuccess: return "hipSuccess";
    case hipErrorOutOfMemory: return "hipErrorOutOfMemory";
    case hipErrorInvalidValue: return "hipErrorInvalidValue";
    case hipErrorInvalidConfiguration: return "hipErrorInvalidConfiguration";
    case hipErrorLaunchFailure: return "hipErrorLaunchFailure";
    default: return "hipError (CPU model)";
  }
}
hipError_t hipDeviceSynchronize() { return hipSuccess; }

hipError_t hipMalloc(void **p, size_t bytes) {
  *p = nullptr;
  g_mallocs.fetch_add(1);
  if (inject("malloc")) {
    set_err(hipErrorOutOfMemory);
    return hipErrorOutOfMemory;
  }
  if (bytes == 0) return hipSuccess;
  void *q = nullptr;
  if (posix_memalign(&q, 256, bytes) != 0) {
    set_err(hipErrorOutOfMemory);
    return hipErrorOutOfMemory;
  }
  std::lock_guard<std::mutex> lk(g_mem_mu);
  g_allocs[q] = bytes;
  g_dev_bytes += bytes;
  g_dev_peak = std::max(g_dev_peak, g_dev_bytes);
  *p = q;
  return hipSuccess;
}
hipError_t hipFree(void *p) {
  if (!p) return hipSuccess;
  {
    std::lock_guard<std::mutex> lk(g_mem_mu);
    auto it = g_allocs.find(p);
    if (it == g_allocs.end()) return hipErrorInvalidValue;
    g_dev_bytes -= it->second;
    g_allocs.erase(it);
  }
  free(p);
  return hipSuccess;
}
hipError_t hipHostMalloc(void **p, size_t bytes, unsigned flags) {
  (void)flags;
  *p = nullptr;
  g_mallocs.fetch_add(1);
  if (inject("host_malloc")) return hipErrorOutOfMemory;
  if (posix_memalign(p, 256, bytes ? bytes : 1) != 0) return hipErrorOutOfMemory;
  return hipSuccess;
}
hipError_t hipHostFree(void *p) {
  free(p);
  return hipSuccess;
}
hipError_t hipMemcpy(void *dst, const void *src, size_t bytes, hipMemcpyKind k) {
  return hipMemcpyAsync(dst, src, bytes, k, nullptr);
}
hipError_t hipMemcpyAsync(void *dst, const void *src, size_t bytes, hipMemcpyKind k, hipStream_t s) {
  if (inject("memcpy")) return hipErrorInvalidValue;
  // the device side of a copy must lie inside one allocation
  if ((k == hipMemcpyHostToDevice || k == hipMemcpyDeviceToDevice) && bytes && !known_device(dst, bytes))
    return hipErrorInvalidValue;
  if ((k == hipMemcpyDeviceToHost || k == hipMemcpyDeviceToDevice) && bytes && !known_device(src, bytes))
    return hipErrorInvalidValue;
  return run_or_record(s, [=]() {
    memmove(dst, src, bytes);
    return hipSuccess;
  });
}
hipError_t hipMemcpy2DAsync(void *dst, size_t dpitch, const void *src, size_t spitch, size_t width, size_t height,
                            hipMemcpyKind k, hipStream_t s) {
  if (inject("memcpy")) return hipErrorInvalidValue;
  if (width > dpitch || width > spitch) return hipErrorInvalidValue;
  const size_t dspan = height ? (height - 1) * dpitch + width : 0, sspan = height ? (height - 1) * spitch + width : 0;
  if ((k == hipMemcpyHostToDevice || k == hipMemcpyDeviceToDevice) && dspan && !known_device(dst, dspan))
    return hipErrorInvalidValue;
  if ((k == hipMemcpyDeviceToHost || k == hipMemcpyDeviceToDevice) && sspan && !known_device(src, sspan))
    return hipErrorInvalidValue;
  return run_or_record(s, [=]() {
    for (size_t r = 0; r < height; ++r)
      memmove((char *)dst + r * dpitch, (const char *)src + r * spitch, width);
    return hipSuccess;
  });
}

hipError_t hipMemsetAsync(void *dst, int v, size_t bytes, hipStream_t s) {
  if (bytes && !known_device(dst, bytes)) return hipErrorInvalidValue;
  return run_or_record(s, [=]() {
    memset(dst, v, bytes);
    return hipSuccess;
  });
}
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned flags) {
  (void)flags;
  *s = new sfn_model_stream();
  return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t s) {
  delete s;
  return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t s) {
  (void)s;
  return hipSuccess;
}
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned flags) {
  (void)s;
  (void)e;
  (void)flags;
  return hipSuccess;
}
hipError_t hipEventCreate(hipEvent_t *e) {
  *e = new sfn_model_event();
  return hipSuccess;
}
hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned flags) {
  (void)flags;
  return hipEventCreate(e);
}
hipError_t hipEventDestroy(hipEvent_t e) {
  delete e;
  return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s) {
  return run_or_record(s, [=]() {
    e->t = std::chrono::steady_clock::now();
    return hipSuccess;
  });
}
hipError_t hipEventSynchronize(hipEvent_t e) {
  (void)e;
  return hipSuccess;
}
hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b) {
  *ms = std::chrono::duration<float, std::milli>(b->t - a->t).count();
  return hipSuccess;
}
hipError_t hipFuncSetAttribute(const void *f, hipFuncAttribute attr, int value) {
  if (attr == hipFuncAttributeMaxDynamicSharedMemorySize && value > 64 * 1024) {
    if ((size_t)value > kLdsBytes) return hipErrorInvalidValue;
    std::lock_guard<std::mutex> lk(g_attr_mu);
    g_big_lds.insert(f);
  }
  return hipSuccess;
}
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode mode) {
  (void)mode;
  if (!s || s->capture) return hipErrorInvalidValue;
  s->capture = new sfn_model_graph();
  return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t *g) {
  if (!s || !s->capture) return hipErrorInvalidValue;
  *g = s->capture;
  s->capture = nullptr;
  return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t *e, hipGraph_t g, hipGraphNode_t *err_node, char *log, size_t n) {
  (void)err_node;
  (void)log;
  (void)n;
  *e = new sfn_model_graph(*g);
  return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t s) {
  return run_or_record(s, [=]() {
    for (auto &op : e->ops) {
      hipError_t r = op();
      if (r != hipSuccess) return r;
    }
    return hipSuccess;
  });
}
hipError_t hipGraphDestroy(hipGraph_t g) {
  delete g;
  return hipSuccess;
}
hipError_t hipGraphExecDestroy(hipGraphExec_t e) {
  delete e;
  return hipSuccess;
}
hipError_t hipGetSymbolAddress(void **p, const void *symbol) {
  *p = const_cast<void *>(symbol);
  return hipSuccess;
}

extern "C" {
uint64_t sfn_model_oob_count() { return g_oob.load(); }
uint64_t sfn_model_launch_count() { return g_launches.load(); }
uint64_t sfn_model_malloc_count() { return g_mallocs.load(); }
uint64_t sfn_model_device_bytes() {
  std::lock_guard<std::mutex> lk(g_mem_mu);
  return g_dev_bytes;
}
uint64_t sfn_model_peak_device_bytes() {
  std::lock_guard<std::mutex> lk(g_mem_mu);
  return g_dev_peak;
}
void sfn_model_reset_counters() {
  g_oob = 0;
  g_launches = 0;
  std::lock_guard<std::mutex> lk(g_mem_mu);
  g_dev_peak = g_dev_bytes;
}
void sfn_model_set_faults(const char *spec) { parse_faults(spec); }
}
