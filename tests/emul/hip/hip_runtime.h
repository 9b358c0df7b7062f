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
// threads of one block run as real host threads that meet at __syncthreads()
// through a std::barrier, so `__shared__` arrays can simply be static.
#pragma once

#include <algorithm>
#include <barrier>
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

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(x)
#define __shared__ static

struct float2 {
    float x, y;
};
struct float4 {
    float x, y, z, w;
};
inline float2 make_float2(float a, float b) { return float2{a, b}; }

struct dim3 {
    unsigned x, y, z;
    dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
struct uint3e {
    unsigned x, y, z;
};

inline thread_local uint3e threadIdx;
inline uint3e blockIdx;
inline dim3 blockDim;
inline dim3 gridDim;

using std::max;
using std::min;

namespace emul {

inline std::barrier<>* g_bar = nullptr;
inline double g_shfl[1024];

struct Pool {
    int n;
    std::vector<std::thread> th;
    std::barrier<> start, done, block;
    std::function<void(int)> task;
    bool quit = false;
    explicit Pool(int n_) : n(n_), start(n_ + 1), done(n_ + 1), block(n_) {
        for (int i = 0; i < n; ++i)
            th.emplace_back([this, i] {
                for (;;) {
                    start.arrive_and_wait();
                    if (quit) return;
                    task(i);
                    done.arrive_and_wait();
                }
            });
    }
    void run(std::function<void(int)> f) {
        task = std::move(f);
        start.arrive_and_wait();
        done.arrive_and_wait();
    }
    ~Pool() {
        quit = true;
        start.arrive_and_wait();
        for (auto& t : th) t.join();
    }
};

inline Pool* pool_for(int n) {
    static std::mutex m;
    static std::map<int, std::unique_ptr<Pool>> pools;
    std::lock_guard<std::mutex> g(m);
    auto& p = pools[n];
    if (!p) p = std::make_unique<Pool>(n);
    return p.get();
}

template <class F, class... A>
void launch(F f, dim3 g, dim3 b, A... args) {
    static std::mutex serial;   // one kernel at a time, blocks in order
    std::lock_guard<std::mutex> lk(serial);
    const int nt = (int)(b.x * b.y * b.z);
    Pool* p = pool_for(nt);
    gridDim = g;
    blockDim = b;
    for (unsigned bz = 0; bz < g.z; ++bz)
        for (unsigned by = 0; by < g.y; ++by)
            for (unsigned bx = 0; bx < g.x; ++bx) {
                blockIdx = uint3e{bx, by, bz};
                g_bar = &p->block;
                p->run([&](int t) {
                    threadIdx = uint3e{(unsigned)(t % b.x), (unsigned)((t / b.x) % b.y), (unsigned)(t / (b.x * b.y))};
                    f(args...);
                });
            }
}

}  // namespace emul

inline void __syncthreads() { emul::g_bar->arrive_and_wait(); }

inline double __shfl_xor(double v, int off, int width = 64) {
    (void)width;
    emul::g_shfl[threadIdx.x] = v;
    __syncthreads();
    const double r = emul::g_shfl[threadIdx.x ^ (unsigned)off];
    __syncthreads();
    return r;
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
enum { hipSuccess = 0, hipErrorOutOfMemory = 2 };
enum hipMemcpyKind { hipMemcpyHostToDevice = 1, hipMemcpyDeviceToHost = 2, hipMemcpyDeviceToDevice = 3 };
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
inline hipError_t hipHostFree(void* p) { return hipFree(p); }
inline hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind) {
    std::memcpy(d, s, n);
    return hipSuccess;
}
inline hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind k, hipStream_t) {
    return hipMemcpy(d, s, n, k);
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
inline hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b) {
    *ms = std::chrono::duration<float, std::milli>(b->t - a->t).count();
    return hipSuccess;
}

#define hipLaunchKernelGGL(F, G, B, SH, ST, ...) emul::launch(F, dim3(G), dim3(B), __VA_ARGS__)
