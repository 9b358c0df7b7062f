// Â: the transform of the Green tensor on the zero-padded grid, built once
// per (lattice, kd) exactly as main.cpp:466-476 builds it: CreateAMatrix
// (dda_func.cpp:85-148) followed by the full legacy radix-2 transform of the
// six components along l, then j, then i, scaled by 1/(8N)
// (fft_func.cpp:317-416 with the kernel of fft_func.cpp:94-186).
//
// A(r) is computed on the host for one octant r = (a d, b d, c d),
// 0 <= a <= ni etc.; every other padded cell is that value with the odd
// components' signs flipped (the reference's (real)(i - 2 i_max) * d0 is the
// exact negation of (real)(2 i_max - i) * d0, so the mirrored cells are
// bit-identical up to sign).  The device expands the octant, then runs the
// legacy transform with the same float twiddle recurrence tables.
#pragma once

#include "fft_common.hpp"

namespace ddafft {

// Octant table oct[a][b][c][6] with a in [0, ni], b in [0, nj], c in [0, nl]
// -> full grid g[I][J][L][6], I < 2ni, J < 2nj, L < 2nl.  zero_nyquist: the
// cells on the planes I = ni, J = nj, L = nl (r = -n d along an axis, where
// the padded grid has no mirror cell) are set to zero; A is then exactly even
// or odd along every axis, and so is its transform.  Those cells never pair a
// source with a target inside the lattice (|r_x| <= (ni - 1) d there), so
// the convolution is mathematically unchanged.
__global__ __launch_bounds__(kBlock) void k_ahat_expand(int ni, int nj, int nl, const float2* __restrict__ oct,
                                                        float2* __restrict__ g, int zero_nyquist) {
    const long long NI2 = 2 * ni, NJ2 = 2 * nj, NL2 = 2 * nl;
    const long long total = NI2 * NJ2 * NL2;
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < total;
         k += (long long)gridDim.x * blockDim.x) {
        const int L = (int)(k % NL2);
        const long long ij = k / NL2;
        const int J = (int)(ij % NJ2), I = (int)(ij / NJ2);
        const int a = I < ni ? I : (int)NI2 - I, b = J < nj ? J : (int)NJ2 - J, c = L < nl ? L : (int)NL2 - L;
        const bool sx = I >= ni, sy = J >= nj, sz = L >= nl;
        const float2* src = oct + (((long long)a * (nj + 1) + b) * (nl + 1) + c) * 6;
        float2 v[6];
#pragma unroll
        for (int e = 0; e < 6; ++e) v[e] = src[e];
        if (sx != sy) v[1] = cneg(v[1]);
        if (sx != sz) v[2] = cneg(v[2]);
        if (sy != sz) v[4] = cneg(v[4]);
        if (zero_nyquist && (I == ni || J == nj || L == nl))
#pragma unroll
            for (int e = 0; e < 6; ++e) v[e] = make_float2(0.f, 0.f);
#pragma unroll
        for (int e = 0; e < 6; ++e) g[k * 6 + e] = v[e];
    }
}

// The octant J <= nj, L <= nl, I <= ni of the transformed grid full[J][L][I][6]
// (2nj x 2nl x 2ni) -> oct[J][L][I][6] ((nj+1) x (nl+1) x (ni+1)).
__global__ __launch_bounds__(kBlock) void k_ahat_octant(int ni, int nj, int nl, const float2* __restrict__ full,
                                                        float2* __restrict__ oct) {
    const long long total = (long long)(nj + 1) * (nl + 1) * (ni + 1);
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < total;
         k += (long long)gridDim.x * blockDim.x) {
        const int I = (int)(k % (ni + 1));
        const long long jl = k / (ni + 1);
        const int L = (int)(jl % (nl + 1)), J = (int)(jl / (nl + 1));
        const float2* src = full + (((long long)J * (2 * nl) + L) * (2 * ni) + I) * 6;
#pragma unroll
        for (int e = 0; e < 6; ++e) oct[k * 6 + e] = src[e];
    }
}

__device__ __forceinline__ int bitrev(int x, int bits) {
    return (int)(__brev((unsigned)x) >> (32 - bits));
}

// Legacy radix-2 transform (fft_func.cpp:94-186) of lines of NA points along
// one axis of g[I][J][L][6] (6 complex components per point).
//   axis 0: along L, tile over J;  axis 1: along J, tile over L;
//   axis 2: along I, tile over L, written to out[J][L][I][6] scaled by norma.
template <int NA>
__global__ __launch_bounds__(kBlock) void k_legacy_axis(int ni, int nj, int nl, int axis, float2* __restrict__ g,
                                                        float2* __restrict__ out, const float2* __restrict__ w,
                                                        float norma) {
    constexpr int LS = NA + kLpad;
    constexpr int TT = clamp16(pow2floor(8192 / (6 * LS)));
    constexpr int NB = ilog2c(NA);
    __shared__ float2 s[TT * 6 * LS];
    const long long NJ2 = 2 * nj, NL2 = 2 * nl;
    // tile axis extent and per-element / per-tile strides, in cmat units
    long long es, ts, base;
    int ext;
    if (axis == 0) {
        ext = (int)NJ2;
        es = 1;
        ts = NL2;
    } else {
        ext = (int)NL2;
        es = (axis == 1) ? NL2 : NJ2 * NL2;
        ts = 1;
    }
    const int tt = min(TT, ext), ntile = ext / tt;
    const int outer = blockIdx.x / ntile, t0 = (blockIdx.x - outer * ntile) * tt;
    if (axis == 0)
        base = ((long long)outer * NJ2 + t0) * NL2;   // outer = I
    else if (axis == 1)
        base = (long long)outer * NJ2 * NL2 + t0;     // outer = I
    else
        base = (long long)outer * NL2 + t0;           // outer = J
    // load, bit-reversed (fft_func.cpp:196-203)
    const int tot = tt * NA * 6;
    for (int f = threadIdx.x; f < tot; f += blockDim.x) {
        int comp, k, t;
        if (axis == 0) {
            comp = f % 6;
            k = (f / 6) % NA;
            t = f / (6 * NA);
        } else {
            comp = f % 6;
            t = (f / 6) % tt;
            k = f / (6 * tt);
        }
        const float2 v = g[(base + (long long)t * ts + (long long)k * es) * 6 + comp];
        s[(t * 6 + comp) * LS + bitrev(k, NB)] = v;
    }
    __syncthreads();
    const int nlines = tt * 6;
    // first stage (fft_func.cpp:208-223)
    for (int f = threadIdx.x; f < nlines * (NA / 2); f += blockDim.x) {
        const int line = f / (NA / 2), p = f - line * (NA / 2);
        float2* v = s + line * LS + 2 * p;
        const float2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
    __syncthreads();
    // remaining stages (fft_func.cpp:227-255)
    for (int l = 1; l < NB; ++l) {
        const int step = 2 << l, half = step >> 1;
        for (int f = threadIdx.x; f < nlines * (NA / 2); f += blockDim.x) {
            const int line = f / (NA / 2), b = f - line * (NA / 2);
            const int grp = b / half, i = b - grp * half;
            const int in = grp * step + i, m = in + half, kk = i * (NA / step);
            float2* v = s + line * LS;
            const float2 t = cmul(w[kk], v[m]);
            const float2 a = v[in];
            v[in] = cadd(t, a);
            v[m] = csub(a, t);
        }
        __syncthreads();
    }
    if (axis < 2) {
        for (int f = threadIdx.x; f < tot; f += blockDim.x) {
            int comp, k, t;
            if (axis == 0) {
                comp = f % 6;
                k = (f / 6) % NA;
                t = f / (6 * NA);
            } else {
                comp = f % 6;
                t = (f / 6) % tt;
                k = f / (6 * tt);
            }
            g[(base + (long long)t * ts + (long long)k * es) * 6 + comp] = s[(t * 6 + comp) * LS + k];
        }
    } else {
        // out[J][L][I][6]: J = outer, L = t0 + t, I = k  (fft_func.cpp:410-413 scaling)
        for (int f = threadIdx.x; f < tot; f += blockDim.x) {
            const int comp = f % 6, k = (f / 6) % NA, t = f / (6 * NA);
            out[((((long long)outer * NL2 + t0 + t) * NA) + k) * 6 + comp] = rmul(norma, s[(t * 6 + comp) * LS + k]);
        }
    }
}

}  // namespace ddafft
