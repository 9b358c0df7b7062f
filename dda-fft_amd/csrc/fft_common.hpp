// Device building blocks shared by every kernel of the DDA-FFT solve path.
//
// Floating-point contract: every float expression below is written in the
// same association order as the reference's x86-64 build (SSE, no FMA), and
// the library is compiled with -ffp-contract=off, so each butterfly, twiddle
// product and 3x3 product rounds exactly as the reference's does.  The FFT
// stages reproduce the reference's mixed radix-4/8 decimation-in-time
// algorithm (fft_kernel.hpp:35-322): the same radix plan (choose_radices,
// fft_kernel.hpp:193-208), the same digit-reversed input order
// (fft_kernel.hpp:222-236), the same exact float twiddles W_N^q
// (fft_kernel.hpp:238-243, built on the host) and the same butterflies
// (bf2c/bf4c/bf8c, fft_kernel.hpp:47-106).  A line therefore transforms to the
// same bits on the GPU as in the reference.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#pragma clang fp contract(off)

namespace ddafft {

constexpr int kBlock = 256;   // threads per block for every pass kernel
#ifndef DDA_FLOAT4_MFMA   // the CPU emulation of the kernels (tests/emul) defines its own
#define DDA_FLOAT4_MFMA float __attribute__((ext_vector_type(4)))
#endif
constexpr int kLpad = 1;      // LDS line padding (float2) against bank conflicts

// ---------------------------------------------------------------------------
// Complex helpers (complex.cpp:61-64, 11-14, 32-35, 67-71)
// ---------------------------------------------------------------------------

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cneg(float2 a) { return make_float2(-a.x, -a.y); }
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ float2 rmul(float r, float2 a) { return make_float2(r * a.x, r * a.y); }

// ---------------------------------------------------------------------------
// Radix plan: log2 N = 3a + b with the radix-8 stages first
// ---------------------------------------------------------------------------

__host__ __device__ constexpr int ilog2c(int n) {
    int k = 0;
    while ((1 << k) < n) ++k;
    return k;
}

__host__ __device__ constexpr int plan_nst(int k) {
    int m = 0;
    while (k > 6) { ++m; k -= 3; }
    return m + (k == 0 ? 0 : (k == 4 || k == 5 || k == 6) ? 2 : 1);
}

__host__ __device__ constexpr int plan_radix(int k, int t) {
    int m = 0;
    while (k > 6) {
        if (m == t) return 8;
        ++m;
        k -= 3;
    }
    int r0 = 0, r1 = 0;
    switch (k) {
        case 1: r0 = 2; break;
        case 2: r0 = 4; break;
        case 3: r0 = 8; break;
        case 4: r0 = 4; r1 = 4; break;
        case 5: r0 = 8; r1 = 4; break;
        case 6: r0 = 8; r1 = 8; break;
        default: break;
    }
    return (t == m) ? r0 : r1;
}

__host__ __device__ constexpr int pow2floor(int v) {
    int p = 1;
    while (p * 2 <= v) p *= 2;
    return p;
}
__host__ __device__ constexpr int clamp16(int v) { return v < 1 ? 1 : (v > 16 ? 16 : v); }

// product of the radices of the stages before stage t
__host__ __device__ constexpr int span_before(int n, int t) {
    int h = 1;
    for (int u = 0; u < t; ++u) h *= plan_radix(ilog2c(n), u);
    return h;
}

template <int N>
struct Plan {
    static constexpr int K = ilog2c(N);
    static constexpr int NST = plan_nst(K);
    static_assert((1 << K) == N, "power-of-two line length");
};

// ---------------------------------------------------------------------------
// Butterflies (fft_kernel.hpp:47-106)
// ---------------------------------------------------------------------------

__device__ __forceinline__ void bf2(float2& a, float2& b) {
    float tr = a.x + b.x, ti = a.y + b.y;
    b.x = a.x - b.x;
    b.y = a.y - b.y;
    a.x = tr;
    a.y = ti;
}

__device__ __forceinline__ void bf4(float2& x0, float2& x1, float2& x2, float2& x3) {
    float t0r = x0.x + x2.x, t0i = x0.y + x2.y;
    float t1r = x0.x - x2.x, t1i = x0.y - x2.y;
    float t2r = x1.x + x3.x, t2i = x1.y + x3.y;
    float t3r = x1.x - x3.x, t3i = x1.y - x3.y;
    x0.x = t0r + t2r; x0.y = t0i + t2i;
    x2.x = t0r - t2r; x2.y = t0i - t2i;
    x1.x = t1r + t3i; x1.y = t1i - t3r;
    x3.x = t1r - t3i; x3.y = t1i + t3r;
}

__device__ __forceinline__ void bf8(float2* z) {
    const float s = 0.70710678118654752440f;
    float2 a0, a1, a2, a3, b0, b1, b2, b3;
    a0 = make_float2(z[0].x + z[4].x, z[0].y + z[4].y);
    a1 = make_float2(z[1].x + z[5].x, z[1].y + z[5].y);
    a2 = make_float2(z[2].x + z[6].x, z[2].y + z[6].y);
    a3 = make_float2(z[3].x + z[7].x, z[3].y + z[7].y);
    b0 = make_float2(z[0].x - z[4].x, z[0].y - z[4].y);
    {
        float dr = z[1].x - z[5].x, di = z[1].y - z[5].y;
        b1 = make_float2((dr + di) * s, (di - dr) * s);
    }
    {
        float dr = z[2].x - z[6].x, di = z[2].y - z[6].y;
        b2 = make_float2(di, -dr);
    }
    {
        float dr = z[3].x - z[7].x, di = z[3].y - z[7].y;
        b3 = make_float2((di - dr) * s, -(dr + di) * s);
    }
    bf4(a0, a1, a2, a3);
    bf4(b0, b1, b2, b3);
    z[0] = a0; z[1] = b0; z[2] = a1; z[3] = b1;
    z[4] = a2; z[5] = b2; z[6] = a3; z[7] = b3;
}

template <int R>
__device__ __forceinline__ void butterfly(float2* z);
template <>
__device__ __forceinline__ void butterfly<2>(float2* z) { bf2(z[0], z[1]); }
template <>
__device__ __forceinline__ void butterfly<4>(float2* z) { bf4(z[0], z[1], z[2], z[3]); }
template <>
__device__ __forceinline__ void butterfly<8>(float2* z) { bf8(z); }

// ---------------------------------------------------------------------------
// Line FFTs in LDS.  `nlines` lines of N points at stride `ls`, input in
// natural order, output in natural order.  Stage 0 gathers its inputs through
// the digit-reversed order (fft_kernel.hpp:222-236) into registers, so loads
// into LDS stay contiguous; every later stage t combines r_t sub-transforms of
// length h into transforms of length span in place (fft_kernel.hpp:270-285,
// stage kernels :119-185).
//
// Thread mapping: consecutive lanes take the same butterfly of consecutive
// lines, so with the padded line stride their LDS addresses fall in
// different banks.  The flat index f = b * nlines + line advances by the
// block size without divisions (LineWalk).
// ---------------------------------------------------------------------------

struct LineWalk {
    int nlines, dq, dr, b0, l0;
    __device__ __forceinline__ explicit LineWalk(int nl) : nlines(nl) {
        dq = kBlock / nl;
        dr = kBlock - dq * nl;
        b0 = (int)threadIdx.x / nl;
        l0 = (int)threadIdx.x - b0 * nl;
    }
    __device__ __forceinline__ void step(int& b, int& line) const {
        b += dq;
        line += dr;
        if (line >= nlines) {
            line -= nlines;
            ++b;
        }
    }
};

// Natural index of the element that digit-reversed slot g * r_0 + q holds,
// less q * (N / r_0): the digit reversal of g over the radices r_1, r_2, ...
template <int N>
__device__ __forceinline__ int rest_natural(int g) {
    constexpr int K = Plan<N>::K;
    unsigned j = 0, rem = (unsigned)g, L = plan_radix(K, 0);
#pragma unroll
    for (int t = 1; t < Plan<N>::NST; ++t) {
        const unsigned r = plan_radix(K, t);
        L *= r;
        const unsigned d = rem % r;
        rem /= r;
        j += d * (N / L);
    }
    return (int)j;
}

template <int N, int T>
struct StageRunner {
    static __device__ __forceinline__ void run(float2* s, const LineWalk& w, int ls, const float2* __restrict__ tw,
                                               int span_prev) {
        if constexpr (T < Plan<N>::NST) {
            constexpr int K = Plan<N>::K;
            constexpr int R = plan_radix(K, T);
            const int span = span_prev * R;
            const int H = span_prev;
            const int step = N / span;
            constexpr int NB = N / R;
            int b = w.b0, line = w.l0;
            while (b < NB) {
                const unsigned g = (unsigned)b / (unsigned)H;   // H, span, step: powers of two
                const int i = b - (int)g * H;
                float2* v = s + line * ls + (int)g * span + i;
                float2 z[R];
#pragma unroll
                for (int q = 0; q < R; ++q) z[q] = v[q * H];
                if (R > 2 && i > 0) {
                    const int ts = i * step;
                    int tq = 0;
#pragma unroll
                    for (int q = 1; q < R; ++q) {
                        tq += ts;   // q * i * step
                        const float2 tv = tw[tq];
                        const float tr = z[q].x * tv.x - z[q].y * tv.y;
                        z[q].y = z[q].x * tv.y + z[q].y * tv.x;
                        z[q].x = tr;
                    }
                }
                butterfly<R>(z);
#pragma unroll
                for (int q = 0; q < R; ++q) v[q * H] = z[q];
                w.step(b, line);
            }
            __syncthreads();
            StageRunner<N, T + 1>::run(s, w, ls, tw, span);
        }
    }
};

// The same transforms for a line count fixed at compile time: every lane's
// (line, butterfly) pairs are constants of its thread index, the loops unroll
// completely and the index arithmetic reduces to shifts, masks and constant
// multiplies.  Same operations in the same order.
//
// Lane grouping: in stage T, consecutive lanes take G consecutive butterflies
// of one line before moving to the next line (G = 1: one butterfly per line).
// Which lane runs which butterfly does not change any result; G only decides
// which LDS banks the lanes of a half-wave meet.  LaneGroup holds the choice
// per (N, lines, line stride, stage).
template <int N, int NL, int LS, int T>
struct LaneGroup {
    static constexpr int G = 1;
};

template <int G, int NL>
__device__ __forceinline__ void lane_pos(unsigned f, unsigned& line, unsigned& b) {
    if constexpr (G == 1) {
        line = f % NL;
        b = f / NL;
    } else {
        line = (f / G) % NL;
        b = f % G + G * (f / (G * NL));
    }
}

// TEND: the stage to stop before (the caller runs the stages from TEND on).
template <int N, int NL, int LS, int T, int BT = kBlock, int TEND = Plan<N>::NST>
struct FixedStage {
    static __device__ __forceinline__ void run(float2* s, const float2* __restrict__ tw) {
        if constexpr (T < TEND) {
            constexpr int K = Plan<N>::K;
            constexpr int R = plan_radix(K, T);
            constexpr int H = span_before(N, T);
            constexpr int span = H * R;
            constexpr int step = N / span;
            constexpr int W = NL * (N / R);
            constexpr int IT = (W + BT - 1) / BT;
            constexpr int G = LaneGroup<N, NL, LS, T>::G;
            static_assert((N / R) % G == 0, "lane group divides the butterflies of a line");
#pragma unroll
            for (int it = 0; it < IT; ++it) {
                const unsigned f = threadIdx.x + it * BT;
                if (W % BT == 0 || f < (unsigned)W) {
                    unsigned line, b;
                    lane_pos<G, NL>(f, line, b);
                    const unsigned g = b / H, i = b % H;
                    float2* v = s + line * LS + g * span + i;
                    float2 z[R];
#pragma unroll
                    for (int q = 0; q < R; ++q) z[q] = v[q * H];
                    if (R > 2 && i > 0) {
                        const unsigned ts = i * step;
                        unsigned tq = 0;
#pragma unroll
                        for (int q = 1; q < R; ++q) {
                            tq += ts;   // q * i * step
                            const float2 tv = tw[tq];
                            const float tr = z[q].x * tv.x - z[q].y * tv.y;
                            z[q].y = z[q].x * tv.y + z[q].y * tv.x;
                            z[q].x = tr;
                        }
                    }
                    butterfly<R>(z);
#pragma unroll
                    for (int q = 0; q < R; ++q) v[q * H] = z[q];
                }
            }
            __syncthreads();
            FixedStage<N, NL, LS, T + 1, BT, TEND>::run(s, tw);
        }
    }
};

// BT: threads per block; TEND: the stage to stop before (Plan<N>::NST: all)
template <int N, int NL, int LS, int BT = kBlock, int TEND = Plan<N>::NST>
__device__ __forceinline__ void fft_lds_fixed(float2* s, const float2* __restrict__ tw) {
    constexpr int K = Plan<N>::K;
    constexpr int R0 = plan_radix(K, 0);
    constexpr int W = NL * (N / R0);
    constexpr int IT = (W + BT - 1) / BT;
    constexpr int G = LaneGroup<N, NL, LS, 0>::G;
    static_assert((N / R0) % G == 0, "lane group divides the butterflies of a line");
    float2 z[IT][R0];
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const unsigned f = threadIdx.x + it * BT;
        if (W % BT == 0 || f < (unsigned)W) {
            unsigned line, b;
            lane_pos<G, NL>(f, line, b);
            const float2* v = s + line * LS + rest_natural<N>(b);
#pragma unroll
            for (int q = 0; q < R0; ++q) z[it][q] = v[q * (N / R0)];
            butterfly<R0>(z[it]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const unsigned f = threadIdx.x + it * BT;
        if (W % BT == 0 || f < (unsigned)W) {
            unsigned line, b;
            lane_pos<G, NL>(f, line, b);
            float2* v = s + line * LS + b * R0;
#pragma unroll
            for (int q = 0; q < R0; ++q) v[q] = z[it][q];
        }
    }
    __syncthreads();
    FixedStage<N, NL, LS, 1, BT, TEND>::run(s, tw);
}

// The last stage of NLINES line transforms of N points in LDS (stride LS),
// in registers: every stage before it has run (fft_lds_fixed with TEND =
// NST - 1), and each output (line, position p, value) with p < KEEP (a
// multiple of N / r_last: the test folds at compile time) goes to `out`
// instead of back to LDS.  The butterflies, twiddles and order of the stage
// in FixedStage: the same bits.  No barrier at the end (nothing is written to
// LDS).
template <int N, int NLINES, int LS, int KEEP, class Out>
__device__ __forceinline__ void fft_last_stage_out(const float2* s, const float2* __restrict__ tw, Out out) {
    constexpr int NST = Plan<N>::NST;
    static_assert(NST >= 2, "a transform of at least two stages");
    constexpr int RL = plan_radix(Plan<N>::K, NST - 1), HL = N / RL;   // span N, step 1
    constexpr int WL = NLINES * HL, ITL = (WL + kBlock - 1) / kBlock;
    static_assert(KEEP % HL == 0 && KEEP <= N, "kept outputs: whole runs of the last stage");
#pragma unroll
    for (int it = 0; it < ITL; ++it) {
        const int f = threadIdx.x + it * kBlock;
        if (WL % kBlock == 0 || f < WL) {
            const int line = f % NLINES, bf = f / NLINES;   // consecutive lanes: consecutive lines
            const float2* v = s + line * LS + bf;
            float2 z[RL];
#pragma unroll
            for (int q = 0; q < RL; ++q) z[q] = v[q * HL];
            if (RL > 2 && bf > 0) {
                int tq = 0;
#pragma unroll
                for (int q = 1; q < RL; ++q) {
                    tq += bf;   // q * bf * step
                    const float2 tv = tw[tq];
                    const float tr = z[q].x * tv.x - z[q].y * tv.y;
                    z[q].y = z[q].x * tv.y + z[q].y * tv.x;
                    z[q].x = tr;
                }
            }
            butterfly<RL>(z);
#pragma unroll
            for (int q = 0; q < RL; ++q)
                if (q * HL < KEEP) out(line, bf + q * HL, z[q]);
        }
    }
}

// MAXL: the most lines any call of this instance passes (sizes the registers
// that carry stage 0 across its barrier); LS: the line stride in float2.
// Ends with a __syncthreads.
template <int N, int MAXL, int LS>
__device__ __forceinline__ void fft_lds(float2* s, int nlines, const float2* __restrict__ tw) {
    constexpr int ls = LS;
    if (nlines == MAXL) {   // the usual case: full tiles
        fft_lds_fixed<N, MAXL, LS>(s, tw);
        return;
    }
    constexpr int K = Plan<N>::K;
    constexpr int R0 = plan_radix(K, 0);
    constexpr int NB0 = N / R0;
    constexpr int ITEMS = (MAXL * NB0 + kBlock - 1) / kBlock;
    const LineWalk w(nlines);
    // stage 0: gather in digit-reversed order (no twiddles: i = 0 throughout)
    float2 z[ITEMS][R0];
    int bb[ITEMS], ll[ITEMS];
    {
        int b = w.b0, line = w.l0;
#pragma unroll
        for (int it = 0; it < ITEMS; ++it) {
            bb[it] = b;
            ll[it] = line;
            if (b < NB0) {
                const float2* v = s + line * ls + rest_natural<N>(b);
#pragma unroll
                for (int q = 0; q < R0; ++q) z[it][q] = v[q * (N / R0)];
                butterfly<R0>(z[it]);
            }
            w.step(b, line);
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < ITEMS; ++it) {
        if (bb[it] < NB0) {
            float2* v = s + ll[it] * ls + bb[it] * R0;
#pragma unroll
            for (int q = 0; q < R0; ++q) v[q] = z[it][q];
        }
    }
    __syncthreads();
    StageRunner<N, 1>::run(s, w, ls, tw, R0);
}

// ---------------------------------------------------------------------------
// XCD-aware tile order.  The dispatcher hands consecutive workgroups to the
// chip's 8 XCDs in turn, each with its own L2.  Renumbering them so that XCD x
// works through the x-th contiguous range of tiles keeps the neighbouring
// tiles whose short transposed runs share 128-byte lines in one L2.
// ---------------------------------------------------------------------------

constexpr int kXcds = 8;

__device__ __forceinline__ int xcd_tile(int b, int nb) {
    if (nb % kXcds != 0) return b;
    return (b % kXcds) * (nb / kXcds) + b / kXcds;
}

// ---------------------------------------------------------------------------
// Block reductions in FP64 (the only FP64 accumulation on the path)
// ---------------------------------------------------------------------------

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Sums NV values over a block of kBlock threads; the result is valid in
// thread 0.  The wave count is the compile-time block size: reading blockDim
// would make the compiler wait for every memory operation in flight, the
// caller's stores included, before the sum.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* scratch /* >= NV*16 doubles */) {
    constexpr int nw = kBlock >> 6;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NV; ++q) v[q] = wave_sum(v[q]);
    __syncthreads();
    if (lane == 0)
#pragma unroll
        for (int q = 0; q < NV; ++q) scratch[q * 16 + wid] = v[q];
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NV; ++q) {
            double t = 0;
            for (int w = 0; w < nw; ++w) t += scratch[q * 16 + w];
            v[q] = t;
        }
    }
}

}  // namespace ddafft
