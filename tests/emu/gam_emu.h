// gam_emu.h -- minimal HIP runtime + device builtins so the kernel sources of
// kaiju_amd/csrc compile as plain C++ for the CPU.
//
// TEST INFRASTRUCTURE ONLY.  Used by tests/test_emu_*.py to run every kernel's
// indexing under AddressSanitizer and to compare the emulated step with the
// oracle before anything reaches a GPU.  Each launch runs its blocks in order
// and the threads of a block in order 0..blockDim-1 (block reductions in the
// kernels have a sequential form under GAM_EMU).
#pragma once
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <functional>
#include <vector>

using std::fmax;
using std::fmin;
using std::sqrt;
using std::erfc;
using std::exp;
using std::copysign;
using std::fabs;

#define __global__
#define __device__
#define __host__
#define __constant__
#define __shared__ static
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __align__(n) alignas(n)

struct dim3 {
    unsigned x, y, z;
    dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
struct emu_idx { unsigned x, y, z; };
inline thread_local emu_idx threadIdx{}, blockIdx{}, blockDim{}, gridDim{};

inline void __syncthreads() {}
template <class T> inline T __shfl_xor(T v, int, int = 64) { return v; }

typedef int hipError_t;
typedef void *hipStream_t;
typedef void *hipEvent_t;
enum { hipSuccess = 0, hipErrorOutOfMemory = 2, hipErrorInvalidValue = 1 };
enum hipMemcpyKind { hipMemcpyHostToHost = 0, hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice,
                     hipMemcpyDefault };
#define hipStreamNonBlocking 1
#define hipHostMallocDefault 0

inline const char *hipGetErrorString(hipError_t e) { return e ? "emulation error" : "no error"; }
inline hipError_t hipGetLastError() { return hipSuccess; }
template <class T> inline hipError_t hipMalloc(T **p, size_t n) {
    *p = (T *)std::malloc(n ? n : 1);
    return *p ? hipSuccess : hipErrorOutOfMemory;
}
template <class T> inline hipError_t hipHostMalloc(T **p, size_t n, int) { return hipMalloc(p, n); }
inline hipError_t hipFree(void *p) { std::free(p); return hipSuccess; }
inline hipError_t hipHostFree(void *p) { std::free(p); return hipSuccess; }
inline hipError_t hipMemset(void *p, int v, size_t n) { std::memset(p, v, n); return hipSuccess; }
inline hipError_t hipMemsetAsync(void *p, int v, size_t n, hipStream_t) { std::memset(p, v, n); return hipSuccess; }
inline hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind) { std::memmove(d, s, n); return hipSuccess; }
inline hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind, hipStream_t);
inline hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned) { *s = (void *)0x1; return hipSuccess; }
inline hipError_t hipGetDevice(int *d) { *d = 0; return hipSuccess; }
#define HIP_SYMBOL(x) (&(x))
inline hipError_t hipGetSymbolAddress(void **p, const void *sym) { *p = const_cast<void *>(sym); return hipSuccess; }
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
inline hipError_t hipStreamDestroy(hipStream_t) { return hipSuccess; }
inline hipError_t hipEventCreate(hipEvent_t *e) { *e = (void *)0x1; return hipSuccess; }
inline hipError_t hipEventDestroy(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
inline hipError_t hipEventElapsedTime(float *t, hipEvent_t, hipEvent_t) { *t = 0.f; return hipSuccess; }

// Stream capture: while a capture is open, launches and copies are recorded
// (arguments by value) instead of run; a graph launch replays them in the
// recorded order, one valid order of the captured dependencies.
typedef std::vector<std::function<void()>> emu_graph;
typedef emu_graph *hipGraph_t;
typedef emu_graph *hipGraphExec_t;
inline thread_local emu_graph *emu_capture = nullptr;
#define hipStreamCaptureModeThreadLocal 1
#define hipEventDisableTiming 2
inline hipError_t hipStreamBeginCapture(hipStream_t, int) {
    if (emu_capture) return (hipError_t)1;
    emu_capture = new emu_graph();
    return hipSuccess;
}
inline hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t *g) {
    *g = emu_capture;
    emu_capture = nullptr;
    return *g ? hipSuccess : (hipError_t)1;
}
inline hipError_t hipGraphInstantiate(hipGraphExec_t *e, hipGraph_t g, void *, void *, size_t) {
    *e = new emu_graph(*g);
    return hipSuccess;
}
inline hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t) {
    for (auto &f : *e) f();
    return hipSuccess;
}
inline hipError_t hipGraphDestroy(hipGraph_t g) { delete g; return hipSuccess; }
inline hipError_t hipGraphGetNodes(hipGraph_t g, void *, size_t *n) { *n = g->size(); return hipSuccess; }
inline hipError_t hipGraphExecDestroy(hipGraphExec_t e) { delete e; return hipSuccess; }
inline hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned) { *e = (void *)0x1; return hipSuccess; }
inline hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return hipSuccess; }

template <class F> inline void emu_launch_now(dim3 g, dim3 b, const F &f) {
    blockDim = {b.x, b.y, b.z};
    gridDim = {g.x, g.y, g.z};
    for (unsigned bz = 0; bz < g.z; bz++)
        for (unsigned by = 0; by < g.y; by++)
            for (unsigned bx = 0; bx < g.x; bx++) {
                blockIdx = {bx, by, bz};
                for (unsigned tz = 0; tz < b.z; tz++)
                    for (unsigned ty = 0; ty < b.y; ty++)
                        for (unsigned tx = 0; tx < b.x; tx++) {
                            threadIdx = {tx, ty, tz};
                            f();
                        }
            }
}

// The kernel arguments are evaluated once, at the launch, as HIP does; a
// recorded launch keeps those values.
template <class F, class... A> inline void emu_launch(dim3 g, dim3 b, F f, A... args) {
    auto run = [g, b, f, args...]() { emu_launch_now(g, b, [&]() { f(args...); }); };
    if (emu_capture) emu_capture->push_back(run);
    else run();
}

#define HIP_KERNEL_NAME(...) __VA_ARGS__
#define hipLaunchKernelGGL(K, G, B, SH, S, ...) emu_launch(dim3(G), dim3(B), [](auto... a) { K(a...); }, __VA_ARGS__)

inline hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind, hipStream_t) {
    if (emu_capture) emu_capture->push_back([d, s, n]() { std::memmove(d, s, n); });
    else std::memmove(d, s, n);
    return hipSuccess;
}
