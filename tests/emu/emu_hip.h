// CPU emulation of the HIP subset used by the engine's kernels -- test
// infrastructure only.  Every GPU lane runs as one OS thread; blocks run one
// after another; __syncthreads and wave intrinsics are barriers over the block
// or over the lane group named by the intrinsic's width.  Built with
// AddressSanitizer it turns an out-of-bounds access into a host report.
#pragma once

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __restrict__ __restrict
#define __shared__ static
#define __align__(n) alignas(n)

struct emu_dim3 { uint32_t x = 0, y = 1, z = 1; };
struct double2 { double x, y; };
struct ulonglong2 { unsigned long long x, y; };
struct uint4 { uint32_t x, y, z, w; };
struct uint2 { uint32_t x, y; };
inline uint2 make_uint2(uint32_t x, uint32_t y) { return uint2{x, y}; }
extern thread_local emu_dim3 threadIdx, blockIdx, blockDim, gridDim;

namespace emu {

struct Barrier {
    std::mutex m;
    std::condition_variable cv;
    int count = 0, waiting = 0;
    uint64_t gen = 0;
    void wait() {
        std::unique_lock<std::mutex> l(m);
        const uint64_t g = gen;
        if (++waiting == count) { waiting = 0; ++gen; cv.notify_all(); }
        else cv.wait(l, [&] { return gen != g; });
    }
};

void syncthreads();
Barrier& group_barrier(int width, int group);
extern uint64_t g_slots[1024];

template <class T> inline uint64_t tobits(T v) { uint64_t b = 0; std::memcpy(&b, &v, sizeof(T)); return b; }
template <class T> inline T frombits(uint64_t b) { T v; std::memcpy(&v, &b, sizeof(T)); return v; }

inline uint64_t exch(uint64_t v, int width, int src_lane) {
    const int t = (int)threadIdx.x;
    const int g = t / width;
    g_slots[t] = v;
    Barrier& b = group_barrier(width, g);
    b.wait();
    const uint64_t r = g_slots[g * width + src_lane];
    b.wait();
    return r;
}

void launch_impl(uint32_t grid, uint32_t block, const std::function<void()>& fn, const char* name = "");

}  // namespace emu

inline void __syncthreads() { emu::syncthreads(); }

template <class T> inline T __shfl_xor(T v, int mask, int width = 64) {
    const int lane = (int)threadIdx.x % width;
    return emu::frombits<T>(emu::exch(emu::tobits(v), width, (lane ^ mask) % width));
}
template <class T> inline T __shfl(T v, int src, int width = 64) {
    return emu::frombits<T>(emu::exch(emu::tobits(v), width, ((src % width) + width) % width));
}
template <class T> inline T __shfl_up(T v, unsigned d, int width = 64) {
    const int lane = (int)threadIdx.x % width;
    const int s = lane >= (int)d ? lane - (int)d : lane;
    return emu::frombits<T>(emu::exch(emu::tobits(v), width, s));
}
template <class T> inline T __shfl_down(T v, unsigned d, int width = 64) {
    const int lane = (int)threadIdx.x % width;
    const int s = lane + (int)d < width ? lane + (int)d : lane;
    return emu::frombits<T>(emu::exch(emu::tobits(v), width, s));
}
inline unsigned long long __ballot(int pred) {
    const int t = (int)threadIdx.x, g = t / 64;
    emu::g_slots[t] = pred ? 1 : 0;
    emu::Barrier& b = emu::group_barrier(64, g);
    b.wait();
    unsigned long long m = 0;
    for (int i = 0; i < 64; ++i) m |= (emu::g_slots[g * 64 + i] ? 1ull : 0ull) << i;
    b.wait();
    return m;
}
inline int __all(int pred) { return __ballot(pred) == ~0ull; }
inline int __any(int pred) { return __ballot(pred) != 0ull; }

inline int __popcll(unsigned long long v) { return __builtin_popcountll(v); }
inline int __clzll(long long v) { return v ? __builtin_clzll((unsigned long long)v) : 64; }
inline int __clz(int v) { return v ? __builtin_clz((unsigned)v) : 32; }
inline long long __double_as_longlong(double d) { long long r; std::memcpy(&r, &d, 8); return r; }
inline uint32_t __float_as_uint(float f) { uint32_t r; std::memcpy(&r, &f, 4); return r; }
inline float __uint_as_float(uint32_t u) { float r; std::memcpy(&r, &u, 4); return r; }

template <class T> inline T min(T a, T b) { return b < a ? b : a; }
template <class T> inline T max(T a, T b) { return a < b ? b : a; }

float erfcinvf(float y);

template <class T> inline T atomicAdd(T* p, T v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
template <class T> inline T atomicOr(T* p, T v) { return __atomic_fetch_or(p, v, __ATOMIC_SEQ_CST); }
template <class T> inline T atomicExch(T* p, T v) { return __atomic_exchange_n(p, v, __ATOMIC_SEQ_CST); }
inline void __builtin_amdgcn_s_sleep(int) { std::this_thread::yield(); }
template <class T> inline T atomicAnd(T* p, T v) { return __atomic_fetch_and(p, v, __ATOMIC_SEQ_CST); }
template <class T> inline T atomicMin(T* p, T v) {
    T old = __atomic_load_n(p, __ATOMIC_SEQ_CST);
    while (v < old && !__atomic_compare_exchange_n(p, &old, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {}
    return old;
}
#define __HIP_MEMORY_SCOPE_AGENT 3
template <class T> inline T __hip_atomic_load(T* p, int order, int) { (void)order; return __atomic_load_n(p, __ATOMIC_SEQ_CST); }
template <class T> inline void __hip_atomic_store(T* p, T v, int order, int) { (void)order; __atomic_store_n(p, v, __ATOMIC_SEQ_CST); }
template <class T> inline T atomicCAS(T* p, T cmp, T v) {
    __atomic_compare_exchange_n(p, &cmp, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST);
    return cmp;                       // the old value (cmp when the exchange happened)
}
template <class T> inline T atomicMax(T* p, T v) {
    T old = __atomic_load_n(p, __ATOMIC_SEQ_CST);
    while (old < v && !__atomic_compare_exchange_n(p, &old, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {}
    return old;
}

// ------------------------------------------------------------------ host runtime subset
typedef int hipError_t;
typedef void* hipStream_t;
typedef struct emu_event* hipEvent_t;
enum { hipSuccess = 0, hipErrorInvalidValue = 1 };
enum hipMemcpyKind { hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice, hipMemcpyHostToHost };
#define hipStreamNonBlocking 1

hipError_t hipMalloc(void** p, size_t bytes);
template <class T> inline hipError_t hipMalloc(T** p, size_t bytes) { return hipMalloc((void**)p, bytes); }
hipError_t hipFree(void* p);
enum { hipHostMallocDefault = 0 };
inline hipError_t hipHostMalloc(void** p, size_t bytes, unsigned) { return hipMalloc(p, bytes); }
inline hipError_t hipHostFree(void* p) { return hipFree(p); }
inline hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind) { std::memmove(d, s, n); return 0; }
inline hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind, hipStream_t) { std::memmove(d, s, n); return 0; }
inline hipError_t hipMemsetAsync(void* d, int v, size_t n, hipStream_t) { std::memset(d, v, n); return 0; }
inline hipError_t hipStreamSynchronize(hipStream_t) { return 0; }
inline hipError_t hipDeviceSynchronize() { return 0; }
inline hipError_t hipGetLastError() { return 0; }
inline const char* hipGetErrorString(hipError_t) { return "emulated"; }
inline hipError_t hipSetDevice(int) { return 0; }
inline hipError_t hipGetDevice(int* d) { *d = 0; return 0; }
#define HIP_SYMBOL(x) (&(x))
inline hipError_t hipGetSymbolAddress(void** p, const void* sym) { *p = const_cast<void*>(sym); return 0; }
inline hipError_t hipGetDeviceCount(int* n) { *n = 1; return 0; }
inline hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) { *s = (void*)1; return 0; }
inline hipError_t hipStreamDestroy(hipStream_t) { return 0; }
hipError_t hipEventCreate(hipEvent_t* e);
hipError_t hipEventDestroy(hipEvent_t e);
hipError_t hipEventRecord(hipEvent_t e, hipStream_t);
inline hipError_t hipEventSynchronize(hipEvent_t) { return 0; }
hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b);
