// CPU execution model of the HIP subset the library uses -- test
// infrastructure only.  The library's sources are compiled unchanged as host
// C++ against this header (tests/emu/Makefile): every workgroup runs on its
// own OS thread, every work-item is a fiber on that thread, wave64 ballots /
// shuffles and __syncthreads are fiber rendezvous, LDS is a thread-local
// buffer and device memory is host memory.  A stuck kernel (a spin wait that
// never ends, a barrier some work-items never reach, a wave collective called
// from divergent code) is reported with the source line of every work-item.
// Nothing in the product links or loads this.
#pragma once

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <functional>
#include <tuple>
#include <type_traits>
#include <utility>

#define __device__
#define __host__
#define __global__
#define __forceinline__ inline
#define __noinline__
#define __shared__ thread_local
#define __launch_bounds__(...)
#define __align__(n) __attribute__((aligned(n)))
#define __HIP_MEMORY_SCOPE_WAVEFRONT 1
#define __HIP_MEMORY_SCOPE_WORKGROUP 2
#define __HIP_MEMORY_SCOPE_AGENT 3
#define __HIP_MEMORY_SCOPE_SYSTEM 4

struct dim3 {
  uint32_t x, y, z;
  constexpr dim3(uint32_t x_ = 1, uint32_t y_ = 1, uint32_t z_ = 1) : x(x_), y(y_), z(z_) {}
};
struct __attribute__((aligned(16))) uint4 { uint32_t x, y, z, w; };
inline uint4 make_uint4(uint32_t x, uint32_t y, uint32_t z, uint32_t w) { return uint4{x, y, z, w}; }
struct __attribute__((aligned(8))) uint2 { uint32_t x, y; };
inline uint2 make_uint2(uint32_t x, uint32_t y) { return uint2{x, y}; }

namespace emu {
extern thread_local dim3 t_block_idx, t_block_dim, t_grid_dim;
const dim3 &thread_idx();
uint64_t ballot(int line, bool pred);
int shfl(int line, int v, int src, int width = 64);
int shfl_up(int line, int v, unsigned delta, int width = 64);
void syncthreads(int line);
void spin_yield(int line);
unsigned lane_id();
uint32_t first_lane(int line, uint32_t v);
}  // namespace emu

#define threadIdx (::emu::thread_idx())
#define blockIdx (::emu::t_block_idx)
#define blockDim (::emu::t_block_dim)
#define gridDim (::emu::t_grid_dim)
#define __syncthreads() ::emu::syncthreads(__LINE__)
#define __builtin_amdgcn_s_barrier() ::emu::syncthreads(__LINE__)
#define __builtin_amdgcn_s_waitcnt(imm) ((void)(imm))
#define __builtin_amdgcn_sched_barrier(m) ((void)(m))
#define __ballot(p) ::emu::ballot(__LINE__, (p))
#define __shfl(...) ::emu::shfl(__LINE__, __VA_ARGS__)
#define __shfl_up(...) ::emu::shfl_up(__LINE__, __VA_ARGS__)
inline unsigned __lane_id() { return ::emu::lane_id(); }
#define __builtin_amdgcn_s_sleep(n) ::emu::spin_yield(__LINE__)
#define __builtin_amdgcn_readfirstlane(v) ::emu::first_lane(__LINE__, (v))
#define __builtin_amdgcn_fence(order, scope) __atomic_thread_fence(order)
#define __hip_atomic_load(p, order, scope) __atomic_load_n((p), (order))
#define __hip_atomic_store(p, v, order, scope) __atomic_store_n((p), (v), (order))
#define __hip_atomic_fetch_add(p, v, order, scope) __atomic_fetch_add((p), (v), (order))

template <class T, class U>
inline T atomicAdd(T *p, U v) { return __atomic_fetch_add(p, (T)v, __ATOMIC_SEQ_CST); }
template <class T, class U>
inline T atomicOr(T *p, U v) { return __atomic_fetch_or(p, (T)v, __ATOMIC_SEQ_CST); }
template <class T, class U, class V>
inline T atomicCAS(T *p, U cmp, V val) {
  T e = (T)cmp;
  __atomic_compare_exchange_n(p, &e, (T)val, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST);
  return e;
}
template <class T, class U>
inline T atomicMin(T *p, U v) {
  T cur = __atomic_load_n(p, __ATOMIC_SEQ_CST);
  while ((T)v < cur && !__atomic_compare_exchange_n(p, &cur, (T)v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {}
  return cur;
}
template <class T, class U>
inline T atomicMax(T *p, U v) {
  T cur = __atomic_load_n(p, __ATOMIC_SEQ_CST);
  while ((T)v > cur && !__atomic_compare_exchange_n(p, &cur, (T)v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {}
  return cur;
}
// device math the generators use (not bit-identical to the device library)
inline double cospi(double x) { return __builtin_cos(3.141592653589793238462643383279502884 * x); }
inline double sinpi(double x) { return __builtin_sin(3.141592653589793238462643383279502884 * x); }
template <class A, class B>
inline std::common_type_t<A, B> min(A a, B b) { return a < b ? a : b; }
template <class A, class B>
inline std::common_type_t<A, B> max(A a, B b) { return a < b ? b : a; }

// ---------------------------------------------------------------- runtime
typedef int hipError_t;
enum : int { hipSuccess = 0, hipErrorInvalidValue = 1, hipErrorOutOfMemory = 2, hipErrorNotReady = 600 };
struct emu_stream;
struct emu_event;
struct emu_graph;
typedef emu_stream *hipStream_t;
typedef emu_event *hipEvent_t;
typedef emu_graph *hipGraph_t;
typedef emu_graph *hipGraphExec_t;
typedef struct emu_graph_node *hipGraphNode_t;
enum hipMemcpyKind { hipMemcpyHostToHost = 0, hipMemcpyHostToDevice = 1, hipMemcpyDeviceToHost = 2,
                     hipMemcpyDeviceToDevice = 3, hipMemcpyDefault = 4 };
enum hipStreamCaptureMode { hipStreamCaptureModeGlobal = 0, hipStreamCaptureModeThreadLocal = 1,
                            hipStreamCaptureModeRelaxed = 2 };
enum hipFuncAttribute { hipFuncAttributeMaxDynamicSharedMemorySize = 8 };
enum hipDeviceAttribute_t { hipDeviceAttributeMaxSharedMemoryPerBlock = 74, hipDeviceAttributeMultiprocessorCount = 63 };
#define hipStreamDefault 0u
#define hipStreamNonBlocking 1u
#define hipEventDefault 0u
#define hipEventDisableTiming 2u
#define hipHostMallocDefault 0u
struct hipDeviceProp_t {
  char name[256];
  char gcnArchName[256];
  size_t totalGlobalMem;
  size_t sharedMemPerBlock;
  int multiProcessorCount;
  int l2CacheSize;
  int warpSize;
  int maxThreadsPerBlock;
};

hipError_t hipMalloc(void **p, size_t n);
template <class T> inline hipError_t hipMalloc(T **p, size_t n) { return hipMalloc((void **)p, n); }
hipError_t hipFree(void *p);
hipError_t hipHostMalloc(void **p, size_t n, unsigned flags = 0);
template <class T> inline hipError_t hipHostMalloc(T **p, size_t n, unsigned flags = 0) {
  return hipHostMalloc((void **)p, n, flags);
}
hipError_t hipHostFree(void *p);
hipError_t hipMemcpyAsync(void *dst, const void *src, size_t n, hipMemcpyKind k, hipStream_t s = nullptr);
hipError_t hipMemcpy(void *dst, const void *src, size_t n, hipMemcpyKind k);
hipError_t hipMemGetInfo(size_t *fr, size_t *tot);
// file-scope __device__ arrays are host arrays here: their address is the symbol's
#define HIP_SYMBOL(x) (&(x))
template <class T> inline hipError_t hipGetSymbolAddress(void **p, T *sym) { *p = (void *)sym; return hipSuccess; }
hipError_t hipGetDeviceProperties(hipDeviceProp_t *p, int dev);
hipError_t hipDeviceGetAttribute(int *v, hipDeviceAttribute_t a, int dev);
hipError_t hipSetDevice(int d);
hipError_t hipGetDevice(int *d);
hipError_t hipGetDeviceCount(int *n);
hipError_t hipDeviceSynchronize();
hipError_t hipGetLastError();
const char *hipGetErrorString(hipError_t e);
hipError_t hipFuncSetAttribute(const void *f, hipFuncAttribute a, int v);
hipError_t hipStreamCreate(hipStream_t *s);
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned flags);
hipError_t hipStreamDestroy(hipStream_t s);
hipError_t hipStreamSynchronize(hipStream_t s);
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode m);
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t *g);
hipError_t hipGraphInstantiate(hipGraphExec_t *e, hipGraph_t g, hipGraphNode_t *err, char *log, size_t loglen);
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t s);
hipError_t hipGraphDestroy(hipGraph_t g);
hipError_t hipGraphExecDestroy(hipGraphExec_t e);
hipError_t hipEventCreate(hipEvent_t *e);
hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned flags);
hipError_t hipEventDestroy(hipEvent_t e);
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s = nullptr);
hipError_t hipEventSynchronize(hipEvent_t e);
hipError_t hipEventQuery(hipEvent_t e);
hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b);
// the 100 MHz device clock
uint64_t emu_realtime();
#define __builtin_amdgcn_s_memrealtime() ::emu_realtime()
// streams run their work in submission order on the host: a wait is a no-op
inline hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return hipSuccess; }
enum hipMemoryType { hipMemoryTypeHost = 0, hipMemoryTypeDevice = 1 };
struct hipPointerAttribute_t { hipMemoryType type; int device; void *devicePointer; void *hostPointer; };
// every host pointer reads as pageable here
inline hipError_t hipPointerGetAttributes(hipPointerAttribute_t *, const void *) { return hipErrorInvalidValue; }

namespace emu {
// Runs `grid` workgroups of `block` work-items, each calling body(); on a
// capturing stream the launch is recorded instead.
void submit(hipStream_t s, dim3 grid, dim3 block, size_t lds, std::function<void()> body);
using Invoker = std::function<std::function<void()>(void **)>;
void register_kernel(const void *f, Invoker inv);
// Called by the watchdog before it aborts a stuck launch.
void set_hang_hook(void (*f)());
hipError_t launch_registered(const void *f, dim3 grid, dim3 block, void **args, size_t lds, hipStream_t s);

template <class... P, size_t... I>
std::function<void()> bind_args(void (*k)(P...), void **args, std::index_sequence<I...>) {
  std::tuple<std::decay_t<P>...> t(*(std::decay_t<P> *)args[I]...);
  return [k, t]() { std::apply(k, t); };
}
// Makes a kernel launchable through hipLaunchKernel (type-erased arguments).
template <class... P>
bool register_typed(void (*k)(P...)) {
  register_kernel((const void *)k, [k](void **args) { return bind_args(k, args, std::index_sequence_for<P...>{}); });
  return true;
}
template <class... P, class G, class B, class L, class... A>
void launch_ggl(void (*k)(P...), G grid, B block, L lds, hipStream_t s, A &&...a) {
  std::tuple<std::decay_t<P>...> t(std::forward<A>(a)...);
  submit(s, dim3(grid), dim3(block), (size_t)lds, [k, t]() { std::apply(k, t); });
}
}  // namespace emu

#define hipLaunchKernelGGL(...) ::emu::launch_ggl(__VA_ARGS__)
inline hipError_t hipLaunchKernel(const void *f, dim3 grid, dim3 block, void **args, size_t lds, hipStream_t s) {
  return ::emu::launch_registered(f, grid, block, args, lds, s);
}
