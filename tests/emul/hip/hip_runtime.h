// CPU emulation of the small part of the HIP runtime that libddafft_amd uses.
//
// Test tool only: tests/emul/Makefile compiles dda-fft_amd/csrc/ddafft_lib.hip
// against this header into tests/emul/libddafft_emul.so, with AddressSanitizer,
// so that every kernel's indexing and every host-side launch sequence can be
// exercised on the CPU, where an out-of-bounds or null access is reported with
// a stack trace instead of faulting a GPU.  The product binding
// (dda-fft_amd/ddafft.py) never loads this library.
//
// Execution model: a kernel launch runs its blocks one after another; the
// threads of one block run as fibers (ucontext) on the calling host thread and
// meet at __syncthreads(), so `__shared__` arrays can simply be static.
#pragma once

#include <setjmp.h>
#include <ucontext.h>

#include <algorithm>
#include <cstdio>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#if defined(__SANITIZE_ADDRESS__)
#include <sanitizer/common_interface_defs.h>
#define EMUL_ASAN 1
#endif

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(x)
#define __shared__ static
// instruction-scheduling fence and occupancy hint: no effect on results
inline void __builtin_amdgcn_sched_barrier(int) {}
#define DDA_WAVES_PER_EU(...)

struct float2 {
    float x, y;
};
struct float4 {
    float x, y, z, w;
};
inline float2 make_float2(float a, float b) { return float2{a, b}; }
struct int2 {
    int x, y;
};
inline int2 make_int2(int a, int b) { return int2{a, b}; }

struct dim3 {
    unsigned x, y, z;
    dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
struct uint3e {
    unsigned x, y, z;
};

inline uint3e threadIdx;
inline uint3e blockIdx;
inline dim3 blockDim;
inline dim3 gridDim;

using std::max;
using std::min;

namespace emul {

// Each block runs its threads as fibers on the calling host thread: a fiber
// runs until it reaches __syncthreads() (or returns), then the next one runs;
// when every fiber waits at the barrier they are all released.  A block whose
// threads disagree about a barrier (some wait, some returned) aborts.
constexpr size_t kStack = 512 * 1024;
constexpr int kMaxThreads = 1024;

struct Fiber {
    ucontext_t ctx;     // used once, to enter the fiber
    jmp_buf jb;         // later switches: _setjmp/_longjmp (no signal-mask syscall)
    char* stack = nullptr;
    int state = 0;      // 0 runnable, 1 waiting at the barrier, 2 returned
    bool started = false;
};

inline ucontext_t g_main;
inline jmp_buf g_main_jb;
inline Fiber g_fib[kMaxThreads];
inline int g_cur = -1;
inline std::function<void()>* g_body = nullptr;
inline const void* g_main_bottom = nullptr;
inline size_t g_main_size = 0;
inline double g_shfl[kMaxThreads];

inline void to_main() {
    Fiber& f = g_fib[g_cur];
    if (!_setjmp(f.jb)) {
#ifdef EMUL_ASAN
        __sanitizer_start_switch_fiber(nullptr, g_main_bottom, g_main_size);
#endif
        _longjmp(g_main_jb, 1);
    }
#ifdef EMUL_ASAN
    __sanitizer_finish_switch_fiber(nullptr, nullptr, nullptr);
#endif
}

inline void trampoline(int t) {
#ifdef EMUL_ASAN
    __sanitizer_finish_switch_fiber(nullptr, &g_main_bottom, &g_main_size);
#endif
    (*g_body)();
    g_fib[t].state = 2;
#ifdef EMUL_ASAN
    __sanitizer_start_switch_fiber(nullptr, g_main_bottom, g_main_size);
#endif
    _longjmp(g_main_jb, 1);
}

inline void resume(int t) {
    g_cur = t;
    Fiber& f = g_fib[t];
#ifdef EMUL_ASAN
    void* fake = nullptr;
    __sanitizer_start_switch_fiber(&fake, f.stack, kStack);
#endif
    if (!_setjmp(g_main_jb)) {
        if (!f.started) {
            f.started = true;
            swapcontext(&g_main, &f.ctx);
        } else {
            _longjmp(f.jb, 1);
        }
    }
#ifdef EMUL_ASAN
    __sanitizer_finish_switch_fiber(fake, nullptr, nullptr);
#endif
}

// Stream capture: while a capture is open, launches are recorded (by value)
// instead of run; a graph launch runs them in order.
inline std::vector<std::function<void()>>* g_capture = nullptr;

template <class F, class... A>
void launch_now(F f, dim3 g, dim3 b, A... args);

template <class F, class... A>
void launch(F f, dim3 g, dim3 b, A... args) {
    if (g_capture) {
        g_capture->push_back([=]() { launch_now(f, g, b, args...); });
        return;
    }
    launch_now(f, g, b, args...);
}

template <class F, class... A>
void launch_now(F f, dim3 g, dim3 b, A... args) {
    static std::mutex serial;   // one kernel at a time, blocks in order
    std::lock_guard<std::mutex> lk(serial);
    const int nt = (int)(b.x * b.y * b.z);
    if (nt > kMaxThreads) std::abort();
    gridDim = g;
    blockDim = b;
    std::function<void()> body = [&]() { f(args...); };
    g_body = &body;
    for (unsigned bz = 0; bz < g.z; ++bz)
        for (unsigned by = 0; by < g.y; ++by)
            for (unsigned bx = 0; bx < g.x; ++bx) {
                blockIdx = uint3e{bx, by, bz};
                for (int t = 0; t < nt; ++t) {
                    Fiber& fb = g_fib[t];
                    if (!fb.stack) fb.stack = static_cast<char*>(std::malloc(kStack));
                    getcontext(&fb.ctx);
                    fb.ctx.uc_stack.ss_sp = fb.stack;
                    fb.ctx.uc_stack.ss_size = kStack;
                    fb.ctx.uc_link = nullptr;
                    makecontext(&fb.ctx, (void (*)())trampoline, 1, t);
                    fb.state = 0;
                    fb.started = false;
                }
                for (;;) {
                    for (int t = 0; t < nt; ++t)
                        if (g_fib[t].state == 0) {
                            threadIdx = uint3e{(unsigned)(t % b.x), (unsigned)((t / b.x) % b.y),
                                               (unsigned)(t / (b.x * b.y))};
                            resume(t);
                        }
                    int waiting = 0, done = 0;
                    for (int t = 0; t < nt; ++t) {
                        waiting += g_fib[t].state == 1;
                        done += g_fib[t].state == 2;
                    }
                    if (done == nt) break;
                    if (waiting != nt) {
                        std::fprintf(stderr, "emul: divergent __syncthreads in block (%u,%u,%u)\n", bx, by, bz);
                        std::abort();
                    }
                    for (int t = 0; t < nt; ++t) g_fib[t].state = 0;
                }
            }
}

}  // namespace emul

inline void __threadfence_system() {}   // blocks run in turn: every store is already visible

inline void __syncthreads() {
    emul::g_fib[emul::g_cur].state = 1;
    emul::to_main();
}

inline double __shfl_xor(double v, int off, int width = 64) {
    (void)width;
    emul::g_shfl[threadIdx.x] = v;
    __syncthreads();
    const double r = emul::g_shfl[threadIdx.x ^ (unsigned)off];
    __syncthreads();
    return r;
}

// 16x16x4 FP32 matrix product of one wave (v_mfma_f32_16x16x4f32): lane l
// gives A[l % 16][l / 16] and B[l / 16][l % 16] and gets back
// D[4 (l / 16) + v][l % 16] (+ its accumulator), the layout the GPU has.  The
// block's threads meet at a barrier to exchange their operands, so every
// thread of the block must call it the same number of times.
struct emul_f4 {
    float v[4];
    float& operator[](int i) { return v[i]; }
    float operator[](int i) const { return v[i]; }
};
#define DDA_FLOAT4_MFMA emul_f4
inline emul_f4 __builtin_amdgcn_mfma_f32_16x16x4f32(float a, float b, emul_f4 c, int, int, int) {
    static float A[emul::kMaxThreads], B[emul::kMaxThreads];
    const int t = (int)threadIdx.x;
    A[t] = a;
    B[t] = b;
    __syncthreads();
    const int w = t / 64, l = t % 64;
    emul_f4 d = c;
    for (int v = 0; v < 4; ++v) {
        const int row = 4 * (l / 16) + v, col = l % 16;
        float acc = d.v[v];
        for (int k = 0; k < 4; ++k) acc += A[w * 64 + 16 * k + row] * B[w * 64 + 16 * k + col];
        d.v[v] = acc;
    }
    __syncthreads();
    return d;
}

inline unsigned __brev(unsigned x) {
    unsigned r = 0;
    for (int i = 0; i < 32; ++i) r |= ((x >> i) & 1u) << (31 - i);
    return r;
}

// ---------------------------------------------------------------------------
// runtime API
// ---------------------------------------------------------------------------

typedef int hipError_t;
enum { hipSuccess = 0, hipErrorOutOfMemory = 2, hipErrorNotReady = 600 };
enum hipMemcpyKind { hipMemcpyHostToDevice = 1, hipMemcpyDeviceToHost = 2, hipMemcpyDeviceToDevice = 3,
                     hipMemcpyDefault = 4 };
#define __align__(n) alignas(n)
typedef struct emul_stream* hipStream_t;
struct emul_event {
    std::chrono::steady_clock::time_point t;
};
typedef emul_event* hipEvent_t;
constexpr unsigned hipStreamNonBlocking = 1;
constexpr unsigned hipEventDisableTiming = 2;
constexpr unsigned hipHostMallocDefault = 0;

inline const char* hipGetErrorString(hipError_t) { return "emulated HIP error"; }
inline hipError_t hipGetLastError() { return hipSuccess; }
inline hipError_t hipGetDeviceCount(int* n) {
    *n = 1;
    return hipSuccess;
}
inline hipError_t hipGetDevice(int* d) {
    *d = 0;
    return hipSuccess;
}
inline hipError_t hipMalloc(void** p, size_t bytes) {
    *p = std::malloc(bytes ? bytes : 1);
    return *p ? hipSuccess : hipErrorOutOfMemory;
}
inline hipError_t hipFree(void* p) {
    std::free(p);
    return hipSuccess;
}
inline hipError_t hipHostMalloc(void** p, size_t bytes, unsigned) { return hipMalloc(p, bytes); }
#define HIP_SYMBOL(x) (x)
inline hipError_t hipGetSymbolAddress(void** p, const void* sym) {
    *p = const_cast<void*>(sym);   // device variables are host arrays here
    return hipSuccess;
}
inline hipError_t hipHostGetDevicePointer(void** d, void* h, unsigned) {
    *d = h;   // one address space here, as with page-locked memory mapped for the GPU
    return hipSuccess;
}
inline hipError_t hipHostFree(void* p) { return hipFree(p); }
inline hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind) {
    std::memcpy(d, s, n);
    return hipSuccess;
}
inline hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind k, hipStream_t) {
    if (emul::g_capture) {
        emul::g_capture->push_back([=]() { hipMemcpy(d, s, n, k); });
        return hipSuccess;
    }
    return hipMemcpy(d, s, n, k);
}
struct emul_graph {
    std::vector<std::function<void()>> ops;
};
typedef emul_graph* hipGraph_t;
typedef emul_graph* hipGraphExec_t;
enum hipStreamCaptureMode { hipStreamCaptureModeGlobal = 0, hipStreamCaptureModeThreadLocal = 1 };
inline hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) {
    emul::g_capture = new std::vector<std::function<void()>>();
    return hipSuccess;
}
inline hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t* g) {
    *g = new emul_graph{std::move(*emul::g_capture)};
    delete emul::g_capture;
    emul::g_capture = nullptr;
    return hipSuccess;
}
inline hipError_t hipGraphInstantiate(hipGraphExec_t* x, hipGraph_t g, void*, void*, size_t) {
    *x = g;
    return hipSuccess;
}
inline hipError_t hipGraphLaunch(hipGraphExec_t x, hipStream_t) {
    for (auto& op : x->ops) op();
    return hipSuccess;
}
inline hipError_t hipGraphExecDestroy(hipGraphExec_t) { return hipSuccess; }
inline hipError_t hipGraphDestroy(hipGraph_t g) {
    delete g;
    return hipSuccess;
}
inline hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) {
    *s = reinterpret_cast<hipStream_t>(new char[1]);
    return hipSuccess;
}
inline hipError_t hipStreamDestroy(hipStream_t s) {
    delete[] reinterpret_cast<char*>(s);
    return hipSuccess;
}
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
inline hipError_t hipMemset(void* p, int v, size_t n) {
    std::memset(p, v, n);
    return hipSuccess;
}
inline hipError_t hipDeviceSynchronize() { return hipSuccess; }
inline hipError_t hipEventCreate(hipEvent_t* e) {
    *e = new emul_event();
    return hipSuccess;
}
inline hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { return hipEventCreate(e); }
inline hipError_t hipEventDestroy(hipEvent_t e) {
    delete e;
    return hipSuccess;
}
inline hipError_t hipEventRecord(hipEvent_t e, hipStream_t) {
    e->t = std::chrono::steady_clock::now();
    return hipSuccess;
}
inline hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventQuery(hipEvent_t) { return hipSuccess; }   // work runs synchronously here
inline hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b) {
    *ms = std::chrono::duration<float, std::milli>(b->t - a->t).count();
    return hipSuccess;
}

#define hipLaunchKernelGGL(F, G, B, SH, ST, ...) emul::launch(F, dim3(G), dim3(B), __VA_ARGS__)
