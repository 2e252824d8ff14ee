// CPU execution of the library's HIP sources, for tests and debugging only.
//
// TEST INFRASTRUCTURE ONLY.  With tests/wave_emu/include first on the include
// path, <hip/hip_runtime.h> resolves to this file, and the kernels and host
// code of marvel_amd/csrc compile with the host C++ compiler unchanged:
//   * device memory is host memory (malloc), so AddressSanitizer sees every
//     kernel access against the allocation the host code made for it;
//   * hipLaunchKernelGGL runs the grid's blocks one after another, each block
//     with one host thread per work-item; __syncthreads is a barrier over the
//     block, __shared__ arrays are function statics (blocks never overlap);
//   * the wave64 cross-lane operations of wave_emu.hpp need blocks of 64
//     work-items (one wave), the only shape the kernels use them in.
// Streams are synchronous and every API call succeeds unless memory runs out.
#pragma once
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <thread>
#include <memory>
#include <vector>

#include "wave_emu.hpp"

#define __launch_bounds__(...)
#define __shared__ static
#define HIP_KERNEL_NAME(...) __VA_ARGS__

struct dim3 {
    unsigned x, y, z;
    dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};

namespace emu {
extern thread_local dim3 t_block, t_thread;
extern dim3 g_bdim, g_gdim;
extern YieldBarrier* g_blockbar;

template <class F>
void launch(dim3 grid, dim3 block, F&& body)
{
    const unsigned nb = grid.x * grid.y * grid.z, nt = block.x * block.y * block.z;
    if (nb == 0 || nt == 0) return;
    YieldBarrier bar((int)nt);
    // wave operations exist for blocks of whole waves: one exchange context per wave
    const unsigned nw = (nt % 64 == 0) ? nt / 64 : 0;
    std::unique_ptr<Ctx[]> ctx(new Ctx[nw ? nw : 1]);
    Ctx* saved = g;
    g_blockbar = &bar;
    g_bdim = block;
    g_gdim = grid;
    std::vector<std::thread> th;
    th.reserve(nt);
    for (unsigned t = 0; t < nt; t++)
        th.emplace_back([&, t] {
            lane = (int)(t & 63);
            g = nw ? &ctx[t / 64] : nullptr;
            t_thread = dim3(t, 0, 0);
            for (unsigned b = 0; b < nb; b++) {
                t_block = dim3(b, 0, 0);
                body();
                bar.arrive_and_wait();
            }
        });
    for (auto& x : th) x.join();
    g = saved;
}
}  // namespace emu

#define blockIdx emu::t_block
#define threadIdx emu::t_thread
#define blockDim emu::g_bdim
#define gridDim emu::g_gdim
#define __syncthreads() emu::g_blockbar->arrive_and_wait()
namespace emu {
extern thread_local const void* t_kernarg;
// a kernel called with copies of its arguments; the first one's address
// stands for the kernarg segment (kernels that read their arguments there)
template <class F, class A0, class... As>
inline void call_kernel(F&& f, A0 a0, As... as)
{
    t_kernarg = &a0;
    f(a0, as...);
}
inline const void* kernarg() { return t_kernarg; }
}  // namespace emu
#define hipLaunchKernelGGL(K, G, B, S, ST, ...) emu::launch((G), (B), [&]() { emu::call_kernel(K, __VA_ARGS__); })

inline int atomicAdd(int* p, int v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
inline unsigned atomicAdd(unsigned* p, unsigned v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
inline unsigned long long atomicAdd(unsigned long long* p, unsigned long long v)
{
    return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST);
}
inline int atomicMax(int* p, int v)
{
    int o = __atomic_load_n(p, __ATOMIC_SEQ_CST);
    while (o < v && !__atomic_compare_exchange_n(p, &o, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {
    }
    return o;
}
inline unsigned atomicOr(unsigned* p, unsigned v) { return __atomic_fetch_or(p, v, __ATOMIC_SEQ_CST); }
inline unsigned long long atomicOr(unsigned long long* p, unsigned long long v)
{
    return __atomic_fetch_or(p, v, __ATOMIC_SEQ_CST);
}

// ---- runtime API --------------------------------------------------------------------

typedef int hipError_t;
constexpr hipError_t hipSuccess = 0;
constexpr hipError_t hipErrorNotReady = 600;
constexpr hipError_t hipErrorOutOfMemory = 2;
typedef void* hipStream_t;
enum hipMemcpyKind { hipMemcpyHostToHost, hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice };
constexpr unsigned hipHostMallocMapped = 2;

struct hipDeviceProp_t {
    int multiProcessorCount;
    char gcnArchName[256];
};

inline const char* hipGetErrorString(hipError_t) { return "cpu emulation error"; }
inline hipError_t hipGetLastError() { return hipSuccess; }
template <class T>
inline hipError_t hipMalloc(T** p, size_t n)
{
    *p = (T*)malloc(n ? n : 1);
    return *p ? hipSuccess : hipErrorOutOfMemory;
}
inline hipError_t hipFree(void* p)
{
    free(p);
    return hipSuccess;
}
template <class T>
inline hipError_t hipHostMalloc(T** p, size_t n, unsigned = 0)
{
    *p = (T*)calloc(1, n ? n : 1);
    return *p ? hipSuccess : hipErrorOutOfMemory;
}
inline hipError_t hipHostFree(void* p)
{
    free(p);
    return hipSuccess;
}
template <class T>
inline hipError_t hipHostGetDevicePointer(T** d, void* h, unsigned)
{
    *d = (T*)h;
    return hipSuccess;
}
inline hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind)
{
    if (n) memmove(d, s, n);
    return hipSuccess;
}
inline hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind k, hipStream_t = nullptr)
{
    return hipMemcpy(d, s, n, k);
}
inline hipError_t hipMemset(void* d, int v, size_t n)
{
    if (n) memset(d, v, n);
    return hipSuccess;
}
inline hipError_t hipMemsetAsync(void* d, int v, size_t n, hipStream_t = nullptr) { return hipMemset(d, v, n); }
inline hipError_t hipStreamCreate(hipStream_t* s)
{
    *s = (hipStream_t)0x1;
    return hipSuccess;
}
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
inline hipError_t hipStreamQuery(hipStream_t) { return hipSuccess; }
inline hipError_t hipDeviceSynchronize() { return hipSuccess; }
inline hipError_t hipGetDevice(int* d)
{
    *d = 0;
    return hipSuccess;
}
inline hipError_t hipGetDeviceCount(int* n)
{
    *n = 1;
    return hipSuccess;
}
// HIP graphs: launches run as they are made here, so there is nothing to
// record; the library's graph cache is switched off (MI_GRAPHS=0, set by
// emu_runtime.cpp before main) and these only have to compile.
typedef struct EmuGraph* hipGraph_t;
typedef struct EmuGraphExec* hipGraphExec_t;
constexpr hipError_t hipErrorNotSupported = 801;
enum hipStreamCaptureMode { hipStreamCaptureModeGlobal, hipStreamCaptureModeThreadLocal, hipStreamCaptureModeRelaxed };
inline hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) { return hipErrorNotSupported; }
inline hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t*) { return hipErrorNotSupported; }
inline hipError_t hipGraphInstantiate(hipGraphExec_t*, hipGraph_t, void*, void*, size_t) { return hipErrorNotSupported; }
inline hipError_t hipGraphLaunch(hipGraphExec_t, hipStream_t) { return hipErrorNotSupported; }
inline hipError_t hipGraphDestroy(hipGraph_t) { return hipSuccess; }
inline hipError_t hipGraphExecDestroy(hipGraphExec_t) { return hipSuccess; }
typedef struct EmuEvent* hipEvent_t;
inline hipError_t hipEventCreate(hipEvent_t* e)
{
    *e = (hipEvent_t)0x1;
    return hipSuccess;
}
inline hipError_t hipEventRecord(hipEvent_t, hipStream_t = nullptr) { return hipSuccess; }
inline hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventElapsedTime(float* ms, hipEvent_t, hipEvent_t)
{
    *ms = 0;
    return hipSuccess;
}
inline hipError_t hipGetDeviceProperties(hipDeviceProp_t* p, int)
{
    const char* e = getenv("MI_EMU_CUS");
    p->multiProcessorCount = e ? atoi(e) : 2;
    strcpy(p->gcnArchName, "cpu-emulation");
    return hipSuccess;
}
