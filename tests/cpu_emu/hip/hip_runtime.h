// CPU emulation of the part of HIP that lcodepy_amd's kernels use, for
// checking them on the host under AddressSanitizer / UBSan before they run on
// a GPU.  Each GPU thread of a block is a fiber (ucontext) on one OS thread;
// blocks are spread over a small pool of OS threads; __syncthreads and the
// wave-level exchange switch fibers and report a deadlock (a barrier that not
// every thread reaches) instead of hanging.  Static LDS arrays become
// thread_local function statics and dynamic LDS an exactly sized heap block,
// so every out-of-range global or LDS access is an ASan report.  Device
// allocations are host mallocs of the exact size.
#pragma once
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <thread>
#include <vector>
#include <atomic>
#include <algorithm>
#include <memory>
#include <cstring>

struct dim3 { unsigned x = 1, y = 1, z = 1; dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {} };
struct double2 { double x, y; };
struct float2 { float x, y; };
struct alignas(16) float4 { float x, y, z, w; };
inline float4 make_float4(float x, float y, float z, float w) { return float4{x, y, z, w}; }
struct alignas(16) int4 { int x, y, z, w; };
typedef int hipError_t;
typedef void *hipStream_t;
#define hipSuccess 0
#define hipErrorInvalidValue 1
// Index checks the kernels compile in for this build only.
#define RF_CHECK(cond)                                                                  \
    do {                                                                                \
        if (!(cond)) {                                                                  \
            std::fprintf(stderr, "RF_CHECK failed: %s (%s:%d)\n", #cond, __FILE__, __LINE__); \
            std::abort();                                                               \
        }                                                                               \
    } while (0)
inline hipError_t hipGetLastError() { return 0; }
inline const char *hipGetErrorString(hipError_t) { return "emu"; }

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__ static thread_local
#define __align__(n) alignas(n)
#define __restrict__

extern thread_local dim3 threadIdx, blockIdx;
extern dim3 gridDim, blockDim;
extern thread_local double emu_shfl_d[1024];
extern thread_local float emu_shfl_f[1024];
extern thread_local int emu_shfl_i[1024];
void emu_sync_block();
void emu_sync_wave();
void emu_sync_group(unsigned width);
inline void __syncthreads() { emu_sync_block(); }
inline void __threadfence() { std::atomic_thread_fence(std::memory_order_seq_cst); }
inline double __shfl_xor(double v, int off) {
    int t = threadIdx.x, w = t / 64;
    emu_shfl_d[t] = v;
    emu_sync_wave();
    double r = emu_shfl_d[(t % 64 ^ off) + w * 64];
    emu_sync_wave();
    return r;
}
inline int __shfl_up(int v, unsigned d) {
    int t = threadIdx.x;
    emu_shfl_i[t] = v;
    emu_sync_wave();
    int r = (unsigned)(t % 64) >= d ? emu_shfl_i[t - d] : v;
    emu_sync_wave();
    return r;
}
inline int __shfl(int v, int src) {
    int t = threadIdx.x, w = t / 64;
    emu_shfl_i[t] = v;
    emu_sync_wave();
    int r = emu_shfl_i[(src & 63) + w * 64];
    emu_sync_wave();
    return r;
}
inline unsigned long long __ballot(int pred) {
    int t = threadIdx.x, w = t / 64;
    emu_shfl_i[t] = pred != 0;
    emu_sync_wave();
    unsigned long long m = 0;
    for (int l = 0; l < 64 && w * 64 + l < (int)(blockDim.x * blockDim.y * blockDim.z); ++l)
        if (emu_shfl_i[w * 64 + l]) m |= 1ull << l;
    emu_sync_wave();
    return m;
}
// Width-limited form: only the `width` lanes of the caller's group take part
// (and must all reach it); off < width keeps the partner in the group.
inline double __shfl_xor(double v, int off, int width) {
    int t = threadIdx.x, base = t / width * width;
    emu_shfl_d[t] = v;
    emu_sync_group(width);
    double r = emu_shfl_d[base + ((t - base) ^ off)];
    emu_sync_group(width);
    return r;
}
inline float __shfl_xor(float v, int off) {
    int t = threadIdx.x, w = t / 64;
    emu_shfl_f[t] = v;
    emu_sync_wave();
    float r = emu_shfl_f[(t % 64 ^ off) + w * 64];
    emu_sync_wave();
    return r;
}
inline int atomicAdd(int *p, int v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
inline unsigned long long atomicAdd(unsigned long long *p, unsigned long long v) {
    return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST);
}
inline int atomicSub(int *p, int v) { return __atomic_fetch_sub(p, v, __ATOMIC_SEQ_CST); }
inline unsigned atomicOr(unsigned *p, unsigned v) { return __atomic_fetch_or(p, v, __ATOMIC_SEQ_CST); }
inline int atomicMax(int *p, int v) {
    int old = __atomic_load_n(p, __ATOMIC_SEQ_CST);
    while (old < v && !__atomic_compare_exchange_n(p, &old, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {}
    return old;
}
inline double atomicAdd(double *p, double v) {
    double old = *p, nw;
    do { nw = old + v; } while (!__atomic_compare_exchange(p, &old, &nw, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST));
    return old;
}
using std::isfinite; using std::floor; using std::sqrt; using std::fmin; using std::fmax;
using std::min; using std::max;

// Runs body for every thread of every block of the grid.
void emu_launch(dim3 grid, dim3 block, size_t lds, const std::function<void()> &body);
template <typename... A> inline void emu_noop(A...) {}
// Runtime stand-ins: host memory with exact sizes so ASan sees overruns.
enum hipMemcpyKind { hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice };
template <typename T> inline hipError_t hipMalloc(T **p, size_t n) { *p = (T *)std::malloc(n ? n : 1); return 0; }
inline hipError_t hipFree(void *p) { std::free(p); return 0; }
inline hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind) { std::memcpy(d, s, n); return 0; }
inline hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind, hipStream_t) { std::memcpy(d, s, n); return 0; }
inline hipError_t hipMemset(void *d, int v, size_t n) { std::memset(d, v, n); return 0; }
inline hipError_t hipMemsetAsync(void *d, int v, size_t n, hipStream_t) { std::memset(d, v, n); return 0; }
inline hipError_t hipDeviceSynchronize() { return 0; }
inline hipError_t hipStreamSynchronize(hipStream_t) { return 0; }
inline hipError_t hipGetDevice(int *d) { *d = 0; return 0; }
// Events: launches run to completion when issued, so an event is the host
// time at which it was recorded.
typedef double *hipEvent_t;
inline hipError_t hipEventCreate(hipEvent_t *e) { *e = new double(0); return 0; }
inline hipError_t hipEventDestroy(hipEvent_t e) { delete e; return 0; }
inline hipError_t hipEventRecord(hipEvent_t e, hipStream_t) {
    *e = std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
    return 0;
}
inline hipError_t hipEventSynchronize(hipEvent_t) { return 0; }
inline hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b) { *ms = (float)(1e3 * (*b - *a)); return 0; }

// Device symbols: the emulation's "device" variables are ordinary host
// objects, so a symbol's device address is its own address.
#define HIP_SYMBOL(x) (&(x))
#define hipErrorInvalidDevice 101
inline hipError_t hipGetSymbolAddress(void **p, const void *sym) { *p = const_cast<void *>(sym); return 0; }
