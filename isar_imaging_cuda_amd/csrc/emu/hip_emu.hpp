// hip_emu.hpp — CPU execution of the HIP kernel subset used by this library.
//
// Used only by the CPU test build (ISAR_EMU defined, compiled by g++ with AddressSanitizer) so
// that kernel indexing, LDS usage and barrier structure can be checked without a GPU. Each
// workgroup runs its work-items as std::threads; __syncthreads is a std::barrier, wave shuffles
// exchange through a per-workgroup buffer. The product library never includes this file.
#pragma once

#include <algorithm>
#include <atomic>
#include <barrier>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <thread>
#include <vector>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __restrict__
#define __align__(n) alignas(n)
#define HIP_SYMBOL(x) (&(x))

struct float2 { float x, y; };
struct float4 { float x, y, z, w; };
struct double2 { double x, y; };
struct int2 { int x, y; };
inline float2 make_float2(float x, float y) { return float2{x, y}; }
inline double2 make_double2(double x, double y) { return double2{x, y}; }

struct dim3 {
    unsigned x = 1, y = 1, z = 1;
    dim3() = default;
    dim3(unsigned a, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};

namespace isar_emu {
struct BlockCtx {
    std::barrier<>* bar;
    unsigned char* smem;
    std::vector<uint64_t>* xbuf;
};
inline thread_local dim3 t_threadIdx, t_blockIdx, t_blockDim, t_gridDim;
inline thread_local BlockCtx t_ctx;
}  // namespace isar_emu

#define threadIdx (isar_emu::t_threadIdx)
#define blockIdx (isar_emu::t_blockIdx)
#define blockDim (isar_emu::t_blockDim)
#define gridDim (isar_emu::t_gridDim)

#define HIP_DYNAMIC_SHARED(type, var) type* var = reinterpret_cast<type*>(isar_emu::t_ctx.smem)

inline void __syncthreads() { isar_emu::t_ctx.bar->arrive_and_wait(); }
inline void __threadfence() { std::atomic_thread_fence(std::memory_order_seq_cst); }
inline int atomicAdd(int* p, int v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
inline int atomicOr(int* p, int v) { return __atomic_fetch_or(p, v, __ATOMIC_SEQ_CST); }

namespace isar_emu {
template <typename T>
inline T shfl_generic(T v, int src_lane_fn_mode, int arg, int width) {
    static_assert(sizeof(T) <= 8, "shuffle of at most 8 bytes");
    auto& xb = *t_ctx.xbuf;
    const unsigned tid = t_threadIdx.x + t_blockDim.x * (t_threadIdx.y + t_blockDim.y * t_threadIdx.z);
    uint64_t raw = 0;
    std::memcpy(&raw, &v, sizeof(T));
    xb[tid] = raw;
    t_ctx.bar->arrive_and_wait();
    const int lane = static_cast<int>(tid % width);
    const unsigned base = tid - lane;
    int src;
    if (src_lane_fn_mode == 0) src = lane ^ arg;                  // xor
    else if (src_lane_fn_mode == 1) src = lane - arg < 0 ? lane : lane - arg;   // up
    else if (src_lane_fn_mode == 2) src = lane + arg >= width ? lane : lane + arg;  // down
    else src = ((arg % width) + width) % width;                   // idx
    uint64_t got = xb[base + src];
    t_ctx.bar->arrive_and_wait();
    T out;
    std::memcpy(&out, &got, sizeof(T));
    return out;
}
}  // namespace isar_emu

template <typename T> inline T __shfl_xor(T v, int m, int w = 64) { return isar_emu::shfl_generic(v, 0, m, w); }
template <typename T> inline T __shfl_up(T v, unsigned d, int w = 64) { return isar_emu::shfl_generic(v, 1, static_cast<int>(d), w); }
template <typename T> inline T __shfl_down(T v, unsigned d, int w = 64) { return isar_emu::shfl_generic(v, 2, static_cast<int>(d), w); }
template <typename T> inline T __shfl(T v, int l, int w = 64) { return isar_emu::shfl_generic(v, 3, l, w); }

inline void sincospif(float x, float* s, float* c) {
    const double a = 3.14159265358979323846 * static_cast<double>(x);
    *s = static_cast<float>(std::sin(a));
    *c = static_cast<float>(std::cos(a));
}
inline void sincospi(double x, double* s, double* c) {
    *s = std::sin(3.14159265358979323846 * x);
    *c = std::cos(3.14159265358979323846 * x);
}
inline float cospif(float x) { return static_cast<float>(std::cos(3.14159265358979323846 * static_cast<double>(x))); }
inline float sinpif(float x) { return static_cast<float>(std::sin(3.14159265358979323846 * static_cast<double>(x))); }
using std::min;
using std::max;
using std::sqrt;
using std::floor;
using std::fabs;
inline float sqrtf(float x) { return std::sqrt(x); }
inline float fabsf(float x) { return std::fabs(x); }
inline float rintf(float x) { return std::rint(x); }
inline float fmaxf(float a, float b) { return std::fmax(a, b); }
inline float fminf(float a, float b) { return std::fmin(a, b); }
inline float logf(float x) { return std::log(x); }
inline float atan2f(float y, float x) { return std::atan2(y, x); }
inline float rsqrtf(float x) { return 1.0f / std::sqrt(x); }
inline float __int_as_float(int i) { float f; std::memcpy(&f, &i, 4); return f; }
inline int __float_as_int(float f) { int i; std::memcpy(&i, &f, 4); return i; }
inline unsigned __float_as_uint(float f) { unsigned i; std::memcpy(&i, &f, 4); return i; }
inline float __uint_as_float(unsigned i) { float f; std::memcpy(&f, &i, 4); return f; }

// ---- runtime subset ----
typedef int hipError_t;
typedef void* hipStream_t;
typedef void* hipEvent_t;
enum { hipSuccess = 0, hipErrorInvalidValue = 1, hipErrorOutOfMemory = 2 };
enum hipMemcpyKind { hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice, hipMemcpyDefault };

// device memory starts as all-ones bytes (a NaN in every float), so that a kernel reading
// workspace it has not written shows up in the results
inline hipError_t hipMalloc(void** p, size_t n) {
    *p = std::malloc(n ? n : 1);
    if (*p) std::memset(*p, 0xFF, n ? n : 1);
    return *p ? hipSuccess : hipErrorOutOfMemory;
}
template <typename T> inline hipError_t hipMalloc(T** p, size_t n) { return hipMalloc(reinterpret_cast<void**>(p), n); }
inline hipError_t hipFree(void* p) { std::free(p); return hipSuccess; }
inline hipError_t hipGetSymbolAddress(void** p, const void* sym) { *p = const_cast<void*>(sym); return hipSuccess; }
inline hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind, hipStream_t = nullptr) { std::memmove(d, s, n); return hipSuccess; }
inline hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind) { std::memmove(d, s, n); return hipSuccess; }
inline hipError_t hipMemsetAsync(void* d, int v, size_t n, hipStream_t = nullptr) { std::memset(d, v, n); return hipSuccess; }
inline hipError_t hipMemset(void* d, int v, size_t n) { std::memset(d, v, n); return hipSuccess; }
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
inline hipError_t hipDeviceSynchronize() { return hipSuccess; }
inline hipError_t hipGetLastError() { return hipSuccess; }
inline hipError_t hipPeekAtLastError() { return hipSuccess; }
inline const char* hipGetErrorString(hipError_t) { return "emulated"; }
inline hipError_t hipSetDevice(int) { return hipSuccess; }
inline hipError_t hipGetDevice(int* d) { *d = 0; return hipSuccess; }
inline hipError_t hipGetDeviceCount(int* n) { *n = 1; return hipSuccess; }
inline hipError_t hipStreamCreate(hipStream_t* s) { *s = nullptr; return hipSuccess; }
inline hipError_t hipStreamDestroy(hipStream_t) { return hipSuccess; }
// graphs: not emulated; every call reports an error so that graph mode cannot be used here
typedef void* hipGraph_t;
typedef void* hipGraphExec_t;
enum hipStreamCaptureMode { hipStreamCaptureModeGlobal, hipStreamCaptureModeThreadLocal, hipStreamCaptureModeRelaxed };
enum { hipStreamNonBlocking = 1 };
inline hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) { *s = nullptr; return hipSuccess; }
inline hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) { return hipErrorInvalidValue; }
inline hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t* g) { *g = nullptr; return hipErrorInvalidValue; }
inline hipError_t hipGraphInstantiate(hipGraphExec_t*, hipGraph_t, void*, void*, size_t) { return hipErrorInvalidValue; }
inline hipError_t hipGraphLaunch(hipGraphExec_t, hipStream_t) { return hipErrorInvalidValue; }
inline hipError_t hipGraphExecDestroy(hipGraphExec_t) { return hipSuccess; }
inline hipError_t hipGraphDestroy(hipGraph_t) { return hipSuccess; }
inline hipError_t hipEventCreate(hipEvent_t* e) { *e = nullptr; return hipSuccess; }
inline hipError_t hipEventDestroy(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventRecord(hipEvent_t, hipStream_t = nullptr) { return hipSuccess; }
inline hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventElapsedTime(float* ms, hipEvent_t, hipEvent_t) { *ms = 0.0f; return hipSuccess; }

namespace isar_emu {
// Runs grid x block work-items; blocks sequentially, work-items of one block concurrently.
template <typename K, typename... A>
inline void launch(K kernel, dim3 grid, dim3 block, size_t smem, A... args) {
    const unsigned nt = block.x * block.y * block.z;
    std::vector<unsigned char> lds(smem ? smem : 1, 0xCD);   // exact size, poisoned like fresh LDS
    std::vector<uint64_t> xbuf(nt);
    for (unsigned bz = 0; bz < grid.z; ++bz)
        for (unsigned by = 0; by < grid.y; ++by)
            for (unsigned bx = 0; bx < grid.x; ++bx) {
                std::barrier<> bar(static_cast<std::ptrdiff_t>(nt));
                std::vector<std::thread> th;
                th.reserve(nt);
                for (unsigned t = 0; t < nt; ++t) {
                    th.emplace_back([&, t]() {
                        t_threadIdx = dim3(t % block.x, (t / block.x) % block.y, t / (block.x * block.y));
                        t_blockIdx = dim3(bx, by, bz);
                        t_blockDim = block;
                        t_gridDim = grid;
                        t_ctx = BlockCtx{&bar, lds.data(), &xbuf};
                        kernel(args...);
                    });
                }
                for (auto& x : th) x.join();
            }
}
// Runs all blocks of the grid concurrently, each with its own LDS and barrier.
template <typename K, typename... A>
inline void launch_all(K kernel, dim3 grid, dim3 block, size_t smem, A... args) {
    const unsigned nt = block.x * block.y * block.z;
    const unsigned nb = grid.x * grid.y * grid.z;
    std::vector<std::vector<unsigned char>> lds(nb, std::vector<unsigned char>(smem ? smem : 1, 0xCD));
    std::vector<std::vector<uint64_t>> xbuf(nb, std::vector<uint64_t>(nt));
    std::vector<std::unique_ptr<std::barrier<>>> bars;
    for (unsigned b = 0; b < nb; ++b) bars.emplace_back(std::make_unique<std::barrier<>>(static_cast<std::ptrdiff_t>(nt)));
    std::vector<std::thread> th;
    th.reserve(static_cast<size_t>(nb) * nt);
    for (unsigned b = 0; b < nb; ++b) {
        const unsigned bx = b % grid.x, by = (b / grid.x) % grid.y, bz = b / (grid.x * grid.y);
        for (unsigned t = 0; t < nt; ++t) {
            th.emplace_back([&, b, bx, by, bz, t]() {
                t_threadIdx = dim3(t % block.x, (t / block.x) % block.y, t / (block.x * block.y));
                t_blockIdx = dim3(bx, by, bz);
                t_blockDim = block;
                t_gridDim = grid;
                t_ctx = BlockCtx{bars[b].get(), lds[b].data(), &xbuf[b]};
                kernel(args...);
            });
        }
    }
    for (auto& x : th) x.join();
}
}  // namespace isar_emu

#define ISAR_LAUNCH(kernel, grid, block, smem, stream, ...) \
    isar_emu::launch(kernel, dim3(grid), dim3(block), (smem), __VA_ARGS__)
// every workgroup at once (each its own LDS), for kernels whose workgroups wait for each other
#define ISAR_LAUNCH_RESIDENT(kernel, grid, block, smem, stream, ...) \
    isar_emu::launch_all(kernel, dim3(grid), dim3(block), (smem), __VA_ARGS__)
