// Elementwise and scalar kernels of the DDA_FFT solvers (dda_func.cpp:1075-1570)
// plus the observables computed from the converged dipoles: CabsCext
// (dda_func.cpp:1679-1727) and FarField (dda_func.cpp:1576-1625).
#pragma once

#include "device_types.hpp"
#include "fft_common.hpp"
#include "finalize.hpp"
#include "matvec_kernels.hpp"

namespace ddafft {

// exp(i x) of a float phase.  The reference rounds cosf/sinf of the float
// argument (complex.cpp:176-181); evaluating in FP64 and rounding once gives
// the correctly rounded float, which is what glibc returns for all but a
// handful of arguments.
__device__ __forceinline__ float2 exp_i(float x) {
    double s, c;
    sincos((double)x, &s, &c);
    return make_float2((float)c, (float)s);
}

// ---------------------------------------------------------------------------
// Geometry passed to the observables and to the first approximation.
// ---------------------------------------------------------------------------
struct Geometry {
    float d0;          // kd
    float kx, ky, kz;  // incident direction
    float cx, cy, cz;  // centre used by FirstApproximation / FarField: (n-1) d / 2
    float ex, ey, ez;  // centre used by CabsCext: n d / 2 (dda_func.cpp:1685)
};

// Incident-wave phases exp(i k.(r - c)) of an axis-aligned incidence (the
// reference's fixed orientation, kinc along a lattice axis) depend on one
// coordinate only; the host tabulates them with the same float argument and
// the C library's cosf/sinf, as the reference evaluates them
// (complex.cpp:176-181), so they match it bit for bit.  axis[r] = -1 marks a
// general incidence, whose phases are FP64 sincos rounded once to float.
// tab: [nrhs][2][tn], tn = max(ni, nj, nl): FirstApproximation phases, then
// CabsCext phases.  axis[r] = -2 marks a right-hand side whose phases the
// host evaluated per cell: cell [nrhs][2][n].
struct IncPhase {
    const int* axis;
    const float2* tab;
    int tn;
    const float2* cell;
};

__device__ __forceinline__ int axis_coord(int ax, int i, int j, int l) { return ax == 0 ? i : (ax == 1 ? j : l); }

// P0 = alpha_m exp(i k.(r - c)) E on dipoles (dda_func.cpp:654-689).
// kinc: [nrhs][3]; einc: [nrhs][3] complex; alpha: [nrhs][nmat].
__global__ __launch_bounds__(kBlock) void k_first_approx(Dims d, Geometry g, const float* __restrict__ kinc,
                                                         IncPhase ip, const int8_t* __restrict__ occ,
                                                         const int* __restrict__ mv, const float2* __restrict__ alpha,
                                                         int nmat, const float2* __restrict__ einc,
                                                         float2* __restrict__ p) {
    const int rhs = blockIdx.y;
    const long long rb = (long long)rhs * 3 * d.n;
    const float kx = kinc[3 * rhs], ky = kinc[3 * rhs + 1], kz = kinc[3 * rhs + 2];
    const float2 E0 = einc[rhs * 3 + 0], E1 = einc[rhs * 3 + 1], E2 = einc[rhs * 3 + 2];
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < d.n;
         k += (long long)gridDim.x * blockDim.x) {
        // lattice dimensions are powers of two: cell index -> (i, j, l) by shifts
        const int l = (int)(k & (d.nl - 1));
        const long long ij = k >> __builtin_ctz(d.nl);
        const int j = (int)(ij & (d.nj - 1)), i = (int)(ij >> __builtin_ctz(d.nj));
        const int m = occ[k];
        float2 o0 = make_float2(0.f, 0.f), o1 = o0, o2 = o0;
        if (mv[m] != 0) {
            const float rx = (float)i * g.d0, ry = (float)j * g.d0, rz = (float)l * g.d0;
            const float dx = rx - g.cx, dy = ry - g.cy, dz = rz - g.cz;
            const float ph = kx * dx + ky * dy + kz * dz;
            const int ax = ip.axis[rhs];
            const float2 c0 = ax >= 0    ? ip.tab[(size_t)rhs * 2 * ip.tn + axis_coord(ax, i, j, l)]
                              : ax == -2 ? ip.cell[(size_t)rhs * 2 * d.n + k]
                                         : exp_i(ph);
            const float2 c1 = cmul(c0, alpha[(long long)rhs * nmat + m]);
            o0 = cmul(c1, E0);
            o1 = cmul(c1, E1);
            o2 = cmul(c1, E2);
        }
        p[rb + k] = o0;
        p[rb + d.n + k] = o1;
        p[rb + 2 * d.n + k] = o2;
    }
}

// ---------------------------------------------------------------------------
// Elementwise solver updates with fused FP64 partial sums.
// ---------------------------------------------------------------------------

enum UpdKind : int {
    UK_BICG = 0,  // P = -cW R + (P - cA S); R += -cW W; dot1 = (R,G); err(R)   (:1461-1466, 1523)
    UK_CG = 1,    // P += -cA S; R += -cA W; dot1 = (R,W); err(R)              (:1364-1368, 1523)
    UK_AA = 2,    // R += rA W; P += rA S; real1 = |R|^2; err(R)              (:1375-1380, 1523)
};

struct UpdArgs {
    int kind;
    float2* P;
    float2* R;
    const float2* S;
    const float2* W;
    const float2* G;
    double* partials;  // [nrhs][gridDim.x][kSlots]
};

__global__ __launch_bounds__(kBlock) void k_update(Dims d, SolveState* __restrict__ st, int iter, UpdArgs ua,
                                                   const int8_t* __restrict__ dipf) {
    __shared__ double red[kSlots * 16];
    const int rhs = blockIdx.y;
    if (iter >= st[rhs].stop_at) return;
    const float2 ncA = cneg(st[rhs].cA), ncW = cneg(st[rhs].cW);
    const float rA = st[rhs].rA;
    const long long rb = (long long)rhs * 3 * d.n;
    double acc[kSlots] = {0, 0, 0, 0, 0, 0};
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < d.n;
         k += (long long)gridDim.x * blockDim.x) {
        const bool dip = dipf[k] != 0;
        double nrm = 0;
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const long long ck = rb + q * d.n + k;
            float2 p = ua.P[ck], r = ua.R[ck];
            if (ua.kind == UK_BICG) {
                if (dip) {
                    const float2 t2 = cadd(p, cmul(ncA, ua.S[ck]));
                    p = cadd(cmul(ncW, r), t2);
                    r = cadd(r, cmul(ncW, ua.W[ck]));
                } else {
                    p = make_float2(0.f, 0.f);
                }
                if (dip) dot_acc(r, ua.G[ck], acc[0], acc[1]);
            } else if (ua.kind == UK_CG) {
                if (dip) {
                    const float2 w = ua.W[ck];
                    p = cadd(p, cmul(ncA, ua.S[ck]));
                    r = cadd(r, cmul(ncA, w));
                    dot_acc(r, w, acc[0], acc[1]);
                }
            } else {
                if (dip) {
                    r = cadd(r, rmul(rA, ua.W[ck]));
                    p = cadd(p, rmul(rA, ua.S[ck]));
                    acc[4] += norm_term(r);
                }
            }
            ua.P[ck] = p;
            ua.R[ck] = r;
            if (dip) nrm += norm_term(r);
        }
        if (dip) acc[5] += sqrt(nrm);
    }
    block_sum<kSlots>(acc, red);
    if (threadIdx.x == 0) {
        double* p = ua.partials + ((long long)rhs * gridDim.x + blockIdx.x) * kSlots;
#pragma unroll
        for (int q = 0; q < kSlots; ++q) p[q] = acc[q];
    }
}

// Scalar steps: the fixed-order sum of the block partials and the solver
// state update of one right-hand side (one block per right-hand side).
__global__ __launch_bounds__(kBlock) void k_finalize(SolveState* __restrict__ st, int iter, FinArgs fa) {
    __shared__ double red[kSlots * 16];
    const int rhs = blockIdx.x;
    if (iter >= st[rhs].stop_at) return;
    finalize_block(st, iter, fa, rhs, red);
}

// Best-iterate fallback (dda_func.cpp:1537-1548).
__global__ __launch_bounds__(kBlock) void k_restore(Dims d, SolveState* __restrict__ st, float2* __restrict__ P,
                                                    const float2* __restrict__ Pmin) {
    const int rhs = blockIdx.y;
    const SolveState s = st[rhs];
    if (!(s.err > s.err_min)) return;
    if (blockIdx.x == 0 && threadIdx.x == 0) st[rhs].restored = 1;
    const long long rb = (long long)rhs * 3 * d.n;
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < 3 * d.n;
         k += (long long)gridDim.x * blockDim.x)
        P[rb + k] = Pmin[rb + k];
}

// ---------------------------------------------------------------------------
// Cross sections (dda_func.cpp:1679-1727): per-dipole float terms, FP64 sums.
// partials: [nrhs][gridDim.x][2]
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(kBlock) void k_cabs_cext(Dims d, Geometry g, const float* __restrict__ kinc,
                                                      IncPhase ip, const int8_t* __restrict__ occ,
                                                      const float2* __restrict__ alpha, int nmat,
                                                      const float2* __restrict__ einc, const float2* __restrict__ P,
                                                      double* __restrict__ partials) {
    __shared__ double red[2 * 16];
    const int rhs = blockIdx.y;
    const float kx = kinc[3 * rhs], ky = kinc[3 * rhs + 1], kz = kinc[3 * rhs + 2];
    const long long rb = (long long)rhs * 3 * d.n;
    const float2 e0 = cconj(einc[rhs * 3]), e1 = cconj(einc[rhs * 3 + 1]), e2 = cconj(einc[rhs * 3 + 2]);
    const float r4 = (float)(2. / 3.);
    double acc[2] = {0, 0};
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < d.n;
         k += (long long)gridDim.x * blockDim.x) {
        const int m = occ[k];
        const float2 a = alpha[(long long)rhs * nmat + m];
        if (a.x == 0.f && a.y == 0.f) continue;
        // lattice dimensions are powers of two: cell index -> (i, j, l) by shifts
        const int l = (int)(k & (d.nl - 1));
        const long long ij = k >> __builtin_ctz(d.nl);
        const int j = (int)(ij & (d.nj - 1)), i = (int)(ij >> __builtin_ctz(d.nj));
        const float rx = (float)i * g.d0, ry = (float)j * g.d0, rz = (float)l * g.d0;
        const float r3 = -(kx * (rx - g.ex) + ky * (ry - g.ey) + kz * (rz - g.ez));
        const int ax = ip.axis[rhs];
        const float2 c0 = ax >= 0    ? ip.tab[((size_t)rhs * 2 + 1) * ip.tn + axis_coord(ax, i, j, l)]
                          : ax == -2 ? ip.cell[((size_t)rhs * 2 + 1) * d.n + k]
                                     : exp_i(r3);
        const float2 px = P[rb + k], py = P[rb + d.n + k], pz = P[rb + 2 * d.n + k];
        const float2 c5 = cadd(cmul(pz, e2), cadd(cmul(px, e0), cmul(py, e1)));
        acc[0] += (double)cmul(c0, c5).y;
        const float nx = px.x * px.x + px.y * px.y, ny = py.x * py.x + py.y * py.y, nz = pz.x * pz.x + pz.y * pz.y;
        const float ix = px.y * px.x - px.x * px.y, iy = py.y * py.x - py.x * py.y, iz = pz.y * pz.x - pz.x * pz.y;
        const float2 q = make_float2(nz + (nx + ny), iz + (ix + iy));
        const float2 ca = make_float2(a.x, -a.y);
        const float nn = ca.x * ca.x + ca.y * ca.y;
        const float qi = (-q.x * ca.y + q.y * ca.x) / nn;
        acc[1] += (double)(qi - r4 * q.x);
    }
    block_sum<2>(acc, red);
    if (threadIdx.x == 0) {
        double* p = partials + ((long long)rhs * gridDim.x + blockIdx.x) * 2;
        p[0] = acc[0];
        p[1] = acc[1];
    }
}

// ---------------------------------------------------------------------------
// Far field (dda_func.cpp:1576-1625) over a compact list of dipole cells:
//   Esc(n) = S(n) - n (n . S(n)),   S(n) = sum_j exp(i n.(r_j - c)) P_j.
// The transverse projection is linear, so it is applied once per direction
// after the sum (k_far_field_sum).  The phase factorizes per axis,
// exp(i n_x (i d - c_x)) exp(i n_y (j d - c_y)) exp(i n_z (l d - c_z)).
// k_ff_phase tabulates the three factors of every direction once per
// direction set (FP64 sincos rounded to float, arguments formed in float as
// the reference forms its coordinates); each block of k_far_field copies the
// rows of its kFD directions into LDS, so no transcendental is evaluated per
// dipole, and one phase serves all R right-hand sides.
// partials: [nrhs][nchunk][ndirs][6] (x, y, z as re, im)
// ---------------------------------------------------------------------------

constexpr int kFD = 4;            // directions per block
constexpr int kFFAxis = 3 * 256;  // largest ni + nj + nl

// complex product and multiply-accumulate with fused multiply-adds
__device__ __forceinline__ float2 cmul_fma(float2 a, float2 b) {
    return make_float2(fmaf(a.x, b.x, -(a.y * b.y)), fmaf(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ void cmac_fma(float2& acc, float2 a, float2 b) {
    acc.x = fmaf(a.x, b.x, acc.x);
    acc.x = fmaf(-a.y, b.y, acc.x);
    acc.y = fmaf(a.x, b.y, acc.y);
    acc.y = fmaf(a.y, b.x, acc.y);
}

// The far field does not reproduce the reference's sequential float sum in
// any case (its block sums are FP64), so its products use fused
// multiply-adds: a complex multiply-accumulate is four FMAs.
// ptab: [ndirs][ni + nj + nl]
__global__ __launch_bounds__(kBlock) void k_ff_phase(Dims d, Geometry g, const float* __restrict__ dirs, int ndirs,
                                                     float2* __restrict__ ptab) {
    const int naxis = d.ni + d.nj + d.nl;
    const int total = ndirs * naxis;
    for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < total; f += gridDim.x * blockDim.x) {
        const int dd = f / naxis, a = f - dd * naxis;
        float arg;
        if (a < d.ni)
            arg = dirs[3 * dd] * ((float)a * g.d0 - g.cx);
        else if (a < d.ni + d.nj)
            arg = dirs[3 * dd + 1] * ((float)(a - d.ni) * g.d0 - g.cy);
        else
            arg = dirs[3 * dd + 2] * ((float)(a - d.ni - d.nj) * g.d0 - g.cz);
        ptab[f] = exp_i(arg);
    }
}

template <int R>
__global__ __launch_bounds__(kBlock) void k_far_field(Dims d, const float2* __restrict__ ptab,
                                                      const int* __restrict__ cells, int ncells, int ndirs, int rhs0,
                                                      const float2* __restrict__ P, double* __restrict__ partials) {
    constexpr int NV = R * 6;
    __shared__ float2 tab[kFD * kFFAxis];
    __shared__ double red[NV * 16];
    const int nchunk = gridDim.x;
    const int d0 = blockIdx.y * kFD;
    const int naxis = d.ni + d.nj + d.nl;
    for (int f = threadIdx.x; f < kFD * naxis; f += blockDim.x) {
        const int q = f / naxis, a = f - q * naxis;
        const int dd = min(d0 + q, ndirs - 1);
        tab[q * kFFAxis + a] = ptab[dd * naxis + a];
    }
    __syncthreads();
    float2 acc[kFD][R][3];
#pragma unroll
    for (int q = 0; q < kFD; ++q)
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) acc[q][r][c] = make_float2(0.f, 0.f);
    // lattice dimensions are powers of two: cell index -> (i, j, l) by shifts
    const int sl = __builtin_ctz(d.nl), sj = __builtin_ctz(d.nj);
    const int per = (ncells + nchunk - 1) / nchunk;
    const int lo = blockIdx.x * per, hi = min(ncells, lo + per);
    for (int e = lo + threadIdx.x; e < hi; e += blockDim.x) {
        const int k = cells[e];
        const int l = k & (d.nl - 1), ij = k >> sl;
        const int j = ij & (d.nj - 1), i = ij >> sj;
        float2 p[R][3];
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) p[r][c] = P[(long long)(rhs0 + r) * 3 * d.n + c * d.n + k];
#pragma unroll
        for (int q = 0; q < kFD; ++q) {
            const float2* t = tab + q * kFFAxis;
            const float2 ph = cmul_fma(cmul_fma(t[i], t[d.ni + j]), t[d.ni + d.nj + l]);
#pragma unroll
            for (int r = 0; r < R; ++r)
#pragma unroll
                for (int c = 0; c < 3; ++c) cmac_fma(acc[q][r][c], ph, p[r][c]);
        }
    }
    // FP64 block sums, one direction at a time, through the phase table's LDS:
    // value k of thread t goes to sums[k * kBlock + t]; thread (k, g) adds the
    // copies of threads g, g + 16, ..., then thread k adds the 16 group sums,
    // both in a fixed order.
    static_assert(NV * kBlock * sizeof(double) <= sizeof(tab), "block sums fit the phase table");
    static_assert(NV * 16 * sizeof(double) <= sizeof(red), "group sums fit");
    double* sums = reinterpret_cast<double*>(tab);
    const int tid = threadIdx.x;
    __syncthreads();   // every phase read of the loop above is done
#pragma unroll
    for (int q = 0; q < kFD; ++q) {
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                sums[((r * 3 + c) * 2) * kBlock + tid] = acc[q][r][c].x;
                sums[((r * 3 + c) * 2 + 1) * kBlock + tid] = acc[q][r][c].y;
            }
        __syncthreads();
        if (tid < NV * 16) {
            const int k = tid >> 4, g = tid & 15;
            double t = 0.0;
#pragma unroll
            for (int u = 0; u < kBlock / 16; ++u) t += sums[k * kBlock + u * 16 + g];
            red[k * 16 + g] = t;
        }
        __syncthreads();
        if (tid < NV && d0 + q < ndirs) {
            double t = 0.0;
#pragma unroll
            for (int g = 0; g < 16; ++g) t += red[tid * 16 + g];
            const int r = tid / 6, e = tid - r * 6;
            partials[(((long long)(rhs0 + r) * nchunk + blockIdx.x) * ndirs + d0 + q) * 6 + e] = t;
        }
        __syncthreads();   // sums and red are reused by the next direction
    }
}

// Sums the chunks and applies the transverse projection:
// out[nrhs][ndirs][3] complex float.  Right-hand side r uses the directions
// dirs + r * dir_stride (dir_stride 0: one set shared by all).
__global__ __launch_bounds__(kBlock) void k_far_field_sum(const double* __restrict__ partials, int nchunk, int ndirs,
                                                          int nrhs, const float* __restrict__ dirs,
                                                          long long dir_stride, float2* __restrict__ out) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nrhs * ndirs) return;
    const int dd = idx % ndirs, rhs = idx / ndirs;
    double s[6] = {0, 0, 0, 0, 0, 0};
    for (int c = 0; c < nchunk; ++c) {
        const double* p = partials + (((long long)rhs * nchunk + c) * ndirs + dd) * 6;
#pragma unroll
        for (int e = 0; e < 6; ++e) s[e] += p[e];
    }
    const float* dv = dirs + rhs * dir_stride + 3 * dd;
    const double n[3] = {dv[0], dv[1], dv[2]};
    const double dr = n[0] * s[0] + n[1] * s[2] + n[2] * s[4];
    const double di = n[0] * s[1] + n[1] * s[3] + n[2] * s[5];
#pragma unroll
    for (int c = 0; c < 3; ++c)
        out[(long long)idx * 3 + c] = make_float2((float)(s[2 * c] - n[c] * dr), (float)(s[2 * c + 1] - n[c] * di));
}

// ---------------------------------------------------------------------------
// Layout conversion between the reference's cvector AoS ([n][3] complex) and
// the component planes used on the device ([3][n]).
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(kBlock) void k_aos_to_soa(long long n, const float2* __restrict__ aos,
                                                       float2* __restrict__ soa) {
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long long)gridDim.x * blockDim.x) {
        soa[k] = aos[3 * k];
        soa[n + k] = aos[3 * k + 1];
        soa[2 * n + k] = aos[3 * k + 2];
    }
}

__global__ __launch_bounds__(kBlock) void k_soa_to_aos(long long n, const float2* __restrict__ soa,
                                                       float2* __restrict__ aos) {
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long long)gridDim.x * blockDim.x) {
        aos[3 * k] = soa[k];
        aos[3 * k + 1] = soa[n + k];
        aos[3 * k + 2] = soa[2 * n + k];
    }
}

}  // namespace ddafft
