// hip_emu.h — CPU emulation of the subset of HIP the tree library uses, for a
// debugging build of the library's own sources with g++ (bounds-checked with
// AddressSanitizer on the host).  Test infrastructure only: nothing in the
// product loads it.
//
// Kernels run on host threads: one OS thread per GPU thread of a workgroup,
// one workgroup at a time, launches in issue order (every stream is
// synchronous).  __syncthreads is a barrier over the workgroup's threads;
// wavefront votes and shuffles go through per-workgroup arrays.  __shared__
// variables become function-local statics (one workgroup runs at a time, and
// a global lock serializes launches from several host threads).
#pragma once
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <functional>

// ---- qualifiers -----------------------------------------------------------
#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __noinline__ __attribute__((noinline))
#define __launch_bounds__(...)
#define __shared__ static
#define __restrict__ __restrict
#define __align__(n) alignas(n)
#define HIP_SYMBOL(x) (&(x))

// ---- vector types ------------------------------------------------------------
struct dim3 {
    unsigned x, y, z;
    constexpr dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
struct float2 { float x, y; };
struct float4 { float x, y, z, w; };
struct int2 { int x, y; };
struct int4 { int x, y, z, w; };
struct uint2 { unsigned x, y; };
inline float2 make_float2(float x, float y) { return {x, y}; }
inline float4 make_float4(float x, float y, float z, float w) { return {x, y, z, w}; }
inline int2 make_int2(int x, int y) { return {x, y}; }
inline int4 make_int4(int x, int y, int z, int w) { return {x, y, z, w}; }

// ---- thread geometry ---------------------------------------------------------
extern thread_local dim3 threadIdx;
extern thread_local dim3 blockIdx;
extern dim3 blockDim;
extern dim3 gridDim;

// ---- block / wave collectives -------------------------------------------------
void __syncthreads();
unsigned long long __ballot(int pred);
int emu_shfl_i(int v, int src_lane);
float emu_shfl_f(float v, int src_lane);
inline int __shfl(int v, int src, int width = 64) {
    int lane = (int)(threadIdx.x % 64);
    return emu_shfl_i(v, (lane / width) * width + (src & (width - 1)));
}
inline int __shfl_xor(int v, int mask, int width = 64) {
    int lane = (int)(threadIdx.x % 64);
    return emu_shfl_i(v, (lane / width) * width + ((lane % width) ^ mask));
}
inline unsigned __shfl_xor(unsigned v, int mask, int width = 64) {
    return (unsigned)__shfl_xor((int)v, mask, width);
}
inline float __shfl_xor(float v, int mask, int width = 64) {
    int lane = (int)(threadIdx.x % 64);
    return emu_shfl_f(v, (lane / width) * width + ((lane % width) ^ mask));
}
inline int __popcll(unsigned long long v) { return __builtin_popcountll(v); }
inline int __ffsll(long long v) { return __builtin_ffsll(v); }
inline void __threadfence() { std::atomic_thread_fence(std::memory_order_seq_cst); }
#define __builtin_amdgcn_wave_barrier() __syncthreads()
#define __builtin_amdgcn_fence(...) std::atomic_thread_fence(std::memory_order_seq_cst)
inline float __double2float_rn(double d) { return (float)d; }
#include <math.h>
template <class T> inline T min(T a, T b) { return b < a ? b : a; }
template <class T> inline T max(T a, T b) { return a < b ? b : a; }
inline long long min(long long a, int b) { return b < a ? b : a; }
inline long long max(long long a, int b) { return a < b ? b : a; }

// ---- atomics -------------------------------------------------------------------
template <class T> inline T atomicAdd(T* p, T v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
template <class T> inline T atomicSub(T* p, T v) { return __atomic_fetch_sub(p, v, __ATOMIC_SEQ_CST); }
template <class T> inline T atomicOr(T* p, T v) { return __atomic_fetch_or(p, v, __ATOMIC_SEQ_CST); }
template <class T> inline T atomicAnd(T* p, T v) { return __atomic_fetch_and(p, v, __ATOMIC_SEQ_CST); }
template <class T> inline T atomicMax(T* p, T v) {
    T old = __atomic_load_n(p, __ATOMIC_SEQ_CST);
    while (old < v && !__atomic_compare_exchange_n(p, &old, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {}
    return old;
}
template <class T> inline T atomicMin(T* p, T v) {
    T old = __atomic_load_n(p, __ATOMIC_SEQ_CST);
    while (old > v && !__atomic_compare_exchange_n(p, &old, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {}
    return old;
}
inline unsigned atomicOr(unsigned* p, int v) { return atomicOr(p, (unsigned)v); }
inline unsigned atomicAnd(unsigned* p, int v) { return atomicAnd(p, (unsigned)v); }

// ---- runtime -------------------------------------------------------------------
typedef int hipError_t;
constexpr hipError_t hipSuccess = 0;
constexpr hipError_t hipErrorInvalidValue = 1;
constexpr hipError_t hipErrorOutOfMemory = 2;
struct emu_stream;
struct emu_event;
typedef emu_stream* hipStream_t;
typedef emu_event* hipEvent_t;
enum hipMemcpyKind { hipMemcpyHostToHost, hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice,
                     hipMemcpyDefault };
constexpr unsigned hipStreamNonBlocking = 1;
constexpr unsigned hipEventDisableTiming = 2;
constexpr unsigned hipHostMallocDefault = 0;
hipError_t hipMalloc(void** p, size_t n);
template <class T> hipError_t hipMalloc(T** p, size_t n) { return hipMalloc((void**)p, n); }
hipError_t hipFree(void* p);
hipError_t hipHostMalloc(void** p, size_t n, unsigned flags);
template <class T> hipError_t hipHostMalloc(T** p, size_t n, unsigned flags) { return hipHostMalloc((void**)p, n, flags); }
hipError_t hipHostFree(void* p);
// stream-ordered pools: streams run synchronously here, so these are malloc/free
struct emu_pool;
typedef emu_pool* hipMemPool_t;
enum hipMemAllocationType { hipMemAllocationTypePinned = 1 };
enum hipMemAllocationHandleType { hipMemHandleTypeNone = 0 };
enum hipMemLocationType { hipMemLocationTypeDevice = 1 };
struct hipMemLocation { hipMemLocationType type; int id; };
struct hipMemPoolProps {
    hipMemAllocationType allocType;
    hipMemAllocationHandleType handleTypes;
    hipMemLocation location;
    void* win32SecurityAttributes;
    size_t maxSize;
    unsigned char reserved[56];
};
enum hipMemPoolAttr { hipMemPoolAttrReleaseThreshold = 4 };
hipError_t hipMemPoolCreate(hipMemPool_t* mp, const hipMemPoolProps* props);
hipError_t hipMemPoolDestroy(hipMemPool_t mp);
hipError_t hipMemPoolSetAttribute(hipMemPool_t mp, hipMemPoolAttr attr, void* value);
hipError_t hipMallocFromPoolAsync(void** p, size_t n, hipMemPool_t mp, hipStream_t st);
hipError_t hipFreeAsync(void* p, hipStream_t st);
hipError_t hipMallocAsync(void** p, size_t n, hipStream_t st);
hipError_t hipGetDevice(int* dev);
enum hipDeviceAttribute_t { hipDeviceAttributeMultiprocessorCount = 1 };
hipError_t hipDeviceGetAttribute(int* v, hipDeviceAttribute_t attr, int dev);
hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind k);
hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind k, hipStream_t st);
hipError_t hipMemsetAsync(void* d, int v, size_t n, hipStream_t st);
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned flags);
hipError_t hipStreamDestroy(hipStream_t s);
hipError_t hipStreamSynchronize(hipStream_t s);
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned flags);
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned flags);
hipError_t hipEventDestroy(hipEvent_t e);
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s);
hipError_t hipEventSynchronize(hipEvent_t e);
hipError_t hipGetSymbolAddress(void** p, const void* symbol);
hipError_t hipGetLastError();
hipError_t hipDeviceSynchronize();
const char* hipGetErrorString(hipError_t e);

// graphs: a capture on the calling thread records launches (with their
// arguments) instead of running them; a replay runs them in order
struct emu_graph;
struct emu_graph_exec;
typedef emu_graph* hipGraph_t;
typedef emu_graph_exec* hipGraphExec_t;
enum hipStreamCaptureMode { hipStreamCaptureModeGlobal = 0, hipStreamCaptureModeThreadLocal = 1,
                            hipStreamCaptureModeRelaxed = 2 };
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode mode);
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t* g);
hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t g, void* error_node, char* log, size_t log_size);
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t s);
hipError_t hipGraphExecDestroy(hipGraphExec_t e);
hipError_t hipGraphDestroy(hipGraph_t g);

void emu_launch(dim3 grid, dim3 block, size_t shmem, const std::function<void()>& body);
// device operations by name (kernels, copies, fills, primitive calls), for
// counting what one library call issues: emu_dispatch_tally / _reset
void emu_tally(const char* name);
#define hipLaunchKernelGGL(F, G, B, S, ST, ...) \
    (emu_tally(#F), emu_launch(dim3(G), dim3(B), (S), [=]() { F(__VA_ARGS__); }))
