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

__global__ __launch_bounds__(kBlock) void k_update(Dims d, SolveState* __restrict__ st, UpdArgs ua,
                                                   const int8_t* __restrict__ dipf) {
    __shared__ double red[kSlots * 16];
    const int rhs = blockIdx.y;
    if (st[rhs].it >= st[rhs].stop_at) return;
    const float2 ncA = cneg(st[rhs].cA), ncW = cneg(st[rhs].cW);
    const float rA = st[rhs].rA;
    const long long rb = (long long)rhs * 3 * d.n;
    double acc[kSlots] = {0, 0, 0, 0, 0, 0};
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < d.n;
         k += (long long)gridDim.x * blockDim.x) {
        if (dipf[k] == 0) {
            // off the dipoles R keeps its value, and so does P except in
            // BiCGStab, which sets it to zero: nothing is read there
            if (ua.kind == UK_BICG) {
#pragma unroll
                for (int q = 0; q < 3; ++q) ua.P[rb + q * d.n + k] = make_float2(0.f, 0.f);
            }
            continue;
        }
        double nrm = 0;
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const long long ck = rb + q * d.n + k;
            float2 p = ua.P[ck], r = ua.R[ck];
            if (ua.kind == UK_BICG) {
                const float2 t2 = cadd(p, cmul(ncA, ua.S[ck]));
                p = cadd(cmul(ncW, r), t2);
                r = cadd(r, cmul(ncW, ua.W[ck]));
                dot_acc(r, ua.G[ck], acc[0], acc[1]);
            } else if (ua.kind == UK_CG) {
                const float2 w = ua.W[ck];
                p = cadd(p, cmul(ncA, ua.S[ck]));
                r = cadd(r, cmul(ncA, w));
                dot_acc(r, w, acc[0], acc[1]);
            } else {
                r = cadd(r, rmul(rA, ua.W[ck]));
                p = cadd(p, rmul(rA, ua.S[ck]));
                acc[4] += norm_term(r);
            }
            ua.P[ck] = p;
            ua.R[ck] = r;
            nrm += norm_term(r);
        }
        acc[5] += sqrt(nrm);
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
__global__ __launch_bounds__(kBlock) void k_finalize(SolveState* __restrict__ st, FinArgs fa) {
    __shared__ double red[kSlots * 16];
    const int rhs = blockIdx.x;
    const int iter = st[rhs].it;
    if (iter >= st[rhs].stop_at) {   // stopped earlier: its iteration still advances, its state still goes out
        if (threadIdx.x == 0 && ends_step(fa.kind)) {
            st[rhs].it = iter + 1;
            publish_state(fa, rhs, iter, st[rhs]);
        }
        return;
    }
    finalize_block(st, iter, fa, rhs, red);
}

// Best-iterate fallback (dda_func.cpp:1537-1548); block 0 also writes the
// final states to `out` (host-visible, may be nullptr).
__global__ __launch_bounds__(kBlock) void k_restore(Dims d, SolveState* __restrict__ st, float2* __restrict__ P,
                                                    const float2* __restrict__ Pmin, SolveState* __restrict__ out) {
    const int rhs = blockIdx.y;
    SolveState s = st[rhs];
    const bool restore = s.err > s.err_min;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (restore) {
            st[rhs].restored = 1;
            s.restored = 1;
        }
        if (out) out[rhs] = s;   // read by the host after the solve's synchronization
    }
    if (!restore) return;
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
constexpr int kFFAxis = 3 * 256;  // ni + nj + nl of lattices up to 256 per axis (AX of k_far_field)
constexpr int kFFAxisMax = 1088;  // largest ni + nj + nl (axes up to 512, at most 2^24 cells)

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

// AX: the phase rows held in LDS per direction (>= ni + nj + nl)
template <int R, int AX>
__global__ __launch_bounds__(kBlock) void k_far_field(Dims d, const float2* __restrict__ ptab,
                                                      const int* __restrict__ cells, int ncells, int ndirs, int rhs0,
                                                      const float2* __restrict__ P, double* __restrict__ partials) {
    constexpr int NV = R * 6;
    __shared__ float2 tab[kFD * AX];
    __shared__ double red[NV * 16];
    const int nchunk = gridDim.x;
    const int d0 = blockIdx.y * kFD;
    const int naxis = d.ni + d.nj + d.nl;
    for (int f = threadIdx.x; f < kFD * naxis; f += blockDim.x) {
        const int q = f / naxis, a = f - q * naxis;
        const int dd = min(d0 + q, ndirs - 1);
        tab[q * AX + a] = ptab[dd * naxis + a];
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
            const float2* t = tab + q * AX;
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

// The same sums as k_far_field on the matrix cores.  A block takes 16
// directions and a chunk of the dipoles; each wave, 4 dipoles per step: lane
// (m = lane % 16, q = lane / 16) forms the phase of direction d0 + m at dipole
// q (the same two complex products as k_far_field) and loads that dipole's
// moment for column n = lane % 16 (complex column n / 2: right-hand side and
// component; real or imaginary part n % 2).  Two 16x16x4 FP32 MFMAs per step,
//   D += Re(phase) [Pr | Pi]  and  D += Im(phase) [-Pi | Pr],
// accumulate the real and imaginary parts of sum phase * P (MFMA lane layout:
// A[l % 16][l / 16], B[l / 16][l % 16], D[4 (l / 16) + v][l % 16]).  Eight
// steps run per iteration, their cell and moment loads issued first; after
// each, the FP32 accumulator is added into FP64 per lane; the waves' sums
// meet in LDS in a fixed order.  Same partials layout as k_far_field, with
// 16 directions per block row.  AXM: LDS row stride of the phase table
// (ni + nj + nl <= AXM).
constexpr int kFDM = 16;   // directions per block of the MFMA form
template <int R, int AXM>
__global__ __launch_bounds__(kBlock) void k_far_field_mfma(Dims d, const float2* __restrict__ ptab,
                                                           const int* __restrict__ cells, int ncells, int ndirs,
                                                           int rhs0, const float2* __restrict__ P,
                                                           double* __restrict__ partials) {
    typedef DDA_FLOAT4_MFMA f4;
    constexpr int NC = 6 * R;   // real columns: (right-hand side, component) x (re, im)
    static_assert(NC <= 16, "columns fit one MFMA tile");
    constexpr int NW = kBlock / 64;
    constexpr int TS = AXM + 1;   // odd row stride: the 16 directions' rows in different banks
    constexpr int U = 8;          // steps per iteration, their loads issued first
    __shared__ float2 tab[kFDM * TS];
    __shared__ double red[NW * kFDM * 16];
    const int nchunk = gridDim.x;
    const int d0 = blockIdx.y * kFDM;
    const int naxis = d.ni + d.nj + d.nl;
    for (int f = threadIdx.x; f < kFDM * naxis; f += kBlock) {
        const int q = f / naxis, a = f - q * naxis;
        const int dd = min(d0 + q, ndirs - 1);
        tab[q * TS + a] = ptab[dd * naxis + a];
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int m = lane & 15, q = lane >> 4, n = lane & 15;
    const int cpx = n >> 1, rr = cpx / 3, comp = cpx - rr * 3;
    const bool col = n < NC;
    const float2* t = tab + m * TS;
    const long long pbase = (long long)(rhs0 + (col ? rr : 0)) * 3 * d.n + (long long)comp * d.n;
    const int sl = __builtin_ctz(d.nl), sj = __builtin_ctz(d.nj);
    const int per = (ncells + nchunk - 1) / nchunk;
    const int lo = blockIdx.x * per, hi = min(ncells, lo + per);
    f4 acc = {0.f, 0.f, 0.f, 0.f};
    double dacc[4] = {0.0, 0.0, 0.0, 0.0};
    // every wave runs the same number of steps (lanes past the chunk add zero)
    const int nstep = (max(hi - lo, 0) + 4 * NW - 1) / (4 * NW);
    for (int st = 0; st < nstep; st += U) {
        int k[U];
        bool ok[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int e = lo + 4 * (wave + NW * (st + u)) + q;
            ok[u] = e < hi;   // st + u < nstep too: e grows with the step
            k[u] = ok[u] ? cells[e] : 0;
        }
        float2 pv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) pv[u] = (ok[u] && col) ? P[pbase + k[u]] : make_float2(0.f, 0.f);
#pragma unroll
        for (int u = 0; u < U; ++u) {
            float a1 = 0.f, a2 = 0.f;
            if (ok[u]) {
                const int l = k[u] & (d.nl - 1), ij = k[u] >> sl;
                const int j = ij & (d.nj - 1), i = ij >> sj;
                const float2 ph = cmul_fma(cmul_fma(t[i], t[d.ni + j]), t[d.ni + d.nj + l]);
                a1 = ph.x;
                a2 = ph.y;
            }
            const float b1 = (n & 1) ? pv[u].y : pv[u].x, b2 = (n & 1) ? pv[u].x : -pv[u].y;
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b1, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a2, b2, acc, 0, 0, 0);
        }
        // FP32 runs of U steps (4 U dipoles per wave) go into FP64
#pragma unroll
        for (int v = 0; v < 4; ++v) dacc[v] += (double)acc[v];
        acc = f4{0.f, 0.f, 0.f, 0.f};
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) dacc[v] += (double)acc[v];
    // lane holds D[4 (lane / 16) + v][lane % 16]: direction row, real column
#pragma unroll
    for (int v = 0; v < 4; ++v) red[(wave * kFDM + 4 * q + v) * 16 + n] = dacc[v];
    __syncthreads();
    const int tid = threadIdx.x;
    if (tid < kFDM * 16) {
        const int row = tid >> 4, cn = tid & 15;
        double sum = 0.0;
#pragma unroll
        for (int w = 0; w < NW; ++w) sum += red[(w * kFDM + row) * 16 + cn];
        if (cn < NC && d0 + row < ndirs) {
            const int r = (cn >> 1) / 3, e6 = cn - r * 6;   // (component, re / im) of right-hand side r
            partials[(((long long)(rhs0 + r) * nchunk + blockIdx.x) * ndirs + d0 + row) * 6 + e6] = sum;
        }
    }
}

// Far field of a direction set made of groups that share their z component
// (the scattering planes of a fixed orientation are rotations of one plane
// about z, so every polar angle gives one group with one direction per
// plane).  With X, Y, Z the per-axis phase factors,
//   S(n) = sum_(i,j) X_n[i] Y_n[j] T_g(i,j),   T_g(i,j) = sum_l Z_g[l] P(i,j,l),
// so the l-sums are formed once per group instead of once per direction.
// Group g holds the directions g, g + ngroups, ..., g + (S - 1) ngroups.
// The dipole list is cut into segments of at most kFFSeg consecutive dipoles
// of one (i, j) column (seg: {first dipole, (i nj + j) << 6 | count}).
// Stage 1: one thread per segment sums its T for the block's G groups, from
// the right-hand sides' dipole moments gathered in list order (pc, by
// k_ff_gather).  Stage 2: the T of 256 segments go through LDS to 16 threads
// per direction, which add X Y T.  Partials: those of k_far_field.
// ---------------------------------------------------------------------------

constexpr int kFFSeg = 16;    // dipoles per segment at most (below 64)
constexpr int kFFDirs = 16;   // directions per block at most (G * S)
constexpr int kFFStep = 4;    // dipoles of a segment whose loads are in flight together

// pc[e][r * 3 + c] = P of right-hand side rhs0 + r, component c, at dipole e
template <int R>
__global__ __launch_bounds__(kBlock) void k_ff_gather(Dims d, const int* __restrict__ cells, int ncells, int rhs0,
                                                      const float2* __restrict__ P, float2* __restrict__ pc) {
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < ncells; e += gridDim.x * blockDim.x) {
        const int k = cells[e];
        float2 v[3 * R];
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) v[r * 3 + c] = P[(long long)(rhs0 + r) * 3 * d.n + c * d.n + k];
#pragma unroll
        for (int q = 0; q < 3 * R; ++q) pc[(long long)e * 3 * R + q] = v[q];
    }
}

// AXY >= ni + nj, AZ >= nl: the phase rows held in LDS
template <int R, int G, int AXY, int AZ>
__global__ __launch_bounds__(kBlock) void k_far_field_grp(Dims d, const float2* __restrict__ ptab,
                                                          const float2* __restrict__ pc, const int* __restrict__ cells,
                                                          const int2* __restrict__ seg, int nseg, int ngroups, int S,
                                                          int ndirs, int rhs0, double* __restrict__ partials) {
    constexpr int NV = 3 * R;
    constexpr int GT = G > 2 ? G : 2;   // the T rows also hold the FP64 sums at the end
    constexpr int XS = AXY + 1;         // row stride: the direction slots of a wave in different banks
    __shared__ float2 sxy[kFFDirs * XS];   // X and Y rows of the block's directions
    __shared__ float2 sz[G * AZ];           // Z rows of its groups
    __shared__ float2 sT[GT * NV * kBlock];
    __shared__ int sij[kBlock];
    static_assert(kFFDirs * 2 * NV * 16 * sizeof(double) <= sizeof(sT), "direction sums fit the T rows");
    const int naxis = d.ni + d.nj + d.nl, nxy = d.ni + d.nj;
    const int g0 = blockIdx.y * G, ndt = G * S;
    for (int f = threadIdx.x; f < ndt * nxy; f += kBlock) {
        const int q = f / nxy, a = f - q * nxy;
        const int g = min(g0 + q / S, ngroups - 1), sd = q - (q / S) * S;
        sxy[q * XS + a] = ptab[(long long)(g + sd * ngroups) * naxis + a];
    }
    for (int f = threadIdx.x; f < G * d.nl; f += kBlock) {
        const int q = f / d.nl, a = f - q * d.nl;
        const int g = min(g0 + q, ngroups - 1);
        sz[q * AZ + a] = ptab[(long long)g * naxis + nxy + a];   // the same in every direction of the group
    }
    // stage-2 role: direction slot q (group qg, member qs), 16 threads per slot
    const int q = threadIdx.x >> 4, sub = threadIdx.x & 15;
    const int qg = q / S;
    const int lnj = __builtin_ctz(d.nj);
    float2 acc[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) acc[v] = make_float2(0.f, 0.f);
    __syncthreads();
    for (int b = blockIdx.x; b * kBlock < nseg; b += gridDim.x) {
        const int s = b * kBlock + threadIdx.x;
        float2 T[G][NV];
#pragma unroll
        for (int g = 0; g < G; ++g)
#pragma unroll
            for (int v = 0; v < NV; ++v) T[g][v] = make_float2(0.f, 0.f);
        int ij = -1;
        if (s < nseg) {
            const int2 sd = seg[s];
            const int cnt = sd.y & 63;
            ij = sd.y >> 6;
            const float2* p = pc + (long long)sd.x * NV;
            const int* ce = cells + sd.x;
            // kFFStep dipoles per step, their loads issued together
            for (int t0 = 0; t0 < cnt; t0 += kFFStep) {
                float2 pv[kFFStep][NV];
                int lv[kFFStep];
#pragma unroll
                for (int u = 0; u < kFFStep; ++u) {
                    lv[u] = 0;
#pragma unroll
                    for (int v = 0; v < NV; ++v) pv[u][v] = make_float2(0.f, 0.f);
                    if (t0 + u < cnt) {
                        lv[u] = ce[t0 + u] & (d.nl - 1);
#pragma unroll
                        for (int v = 0; v < NV; ++v) pv[u][v] = p[(t0 + u) * NV + v];
                    }
                }
#pragma unroll
                for (int u = 0; u < kFFStep; ++u)
                    if (t0 + u < cnt)
#pragma unroll
                        for (int g = 0; g < G; ++g) {
                            const float2 z = sz[g * AZ + lv[u]];
#pragma unroll
                            for (int v = 0; v < NV; ++v) cmac_fma(T[g][v], z, pv[u][v]);
                        }
            }
        }
#pragma unroll
        for (int g = 0; g < G; ++g)
#pragma unroll
            for (int v = 0; v < NV; ++v) sT[(g * NV + v) * kBlock + threadIdx.x] = T[g][v];
        sij[threadIdx.x] = ij;
        __syncthreads();
        if (q < ndt) {
            const float2* xr = sxy + q * XS;
            const int nv = min(kBlock, nseg - b * kBlock);   // the batch's segments fill its first slots
#pragma unroll 4
            for (int m = sub; m < nv; m += 16) {
                const int ijm = sij[m];
                const int i = ijm >> lnj, j = ijm & (d.nj - 1);
                const float2 ph = cmul_fma(xr[i], xr[d.ni + j]);
#pragma unroll
                for (int v = 0; v < NV; ++v) cmac_fma(acc[v], ph, sT[(qg * NV + v) * kBlock + m]);
            }
        }
        __syncthreads();   // sT and sij are refilled by the next batch
    }
    // FP64 sums of the 16 threads of each direction, in a fixed order
    double* sums = reinterpret_cast<double*>(sT);   // [slot][2 NV][16]
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        sums[(q * 2 * NV + 2 * v) * 16 + sub] = acc[v].x;
        sums[(q * 2 * NV + 2 * v + 1) * 16 + sub] = acc[v].y;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < ndt * 2 * NV; t += kBlock) {
        const int qq = t / (2 * NV), w = t - qq * 2 * NV;
        const int g = g0 + qq / S, sd = qq - (qq / S) * S;
        if (g >= ngroups) continue;
        double a = 0.0;
#pragma unroll
        for (int u = 0; u < 16; ++u) a += sums[t * 16 + u];
        const int r = w / 6, e = w - r * 6;
        partials[(((long long)(rhs0 + r) * gridDim.x + blockIdx.x) * ndirs + g + sd * ngroups) * 6 + e] = a;
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
    const double* p0 = partials + ((long long)rhs * nchunk * ndirs + dd) * 6;
    const long long cs = (long long)ndirs * 6;   // chunk stride
    // chunks in order, the loads of four in flight together
    int c = 0;
    for (; c + 4 <= nchunk; c += 4) {
        double t[4][6];
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int e = 0; e < 6; ++e) t[u][e] = p0[(c + u) * cs + e];
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int e = 0; e < 6; ++e) s[e] += t[u][e];
    }
    for (; c < nchunk; ++c)
#pragma unroll
        for (int e = 0; e < 6; ++e) s[e] += p0[c * cs + e];
    const float* dv = dirs + rhs * dir_stride + 3 * dd;
    const double n[3] = {dv[0], dv[1], dv[2]};
    const double dr = n[0] * s[0] + n[1] * s[2] + n[2] * s[4];
    const double di = n[0] * s[1] + n[1] * s[3] + n[2] * s[5];
#pragma unroll
    for (int c = 0; c < 3; ++c)
        out[(long long)idx * 3 + c] = make_float2((float)(s[2 * c] - n[c] * dr), (float)(s[2 * c + 1] - n[c] * di));
}

// ---------------------------------------------------------------------------
// Reference-order float sums (opt-in, ddafft_set_sum_mode): Cext, Cabs and the
// far field accumulated sequentially in float over the dipoles in the
// reference's storage order (i, then j, then l: dda_func.cpp:1684-1725,
// 1582-1618), from the same float terms.  A float sum is not associative, so
// each sum stays one sequential chain, run by one thread; the terms of the
// chains are formed in parallel beforehand.  Dipole cells are the compact
// list `cells` (ascending cell index, mv != 0); every polarizability model
// gives alpha = 0 exactly for m = 1 (the only material with mv = 0) and a
// non-zero alpha otherwise, so the list is the set of cells CabsCext visits
// (alpha != 0) as well.
// ---------------------------------------------------------------------------

// Cross-section terms of the dipole list: terms [nrhs][2][ncells], the Cext
// term Im(c0 (P . conj E)) and the Cabs term Im(|P|^2 / conj(alpha)) - (2/3)|P|^2
// exactly as k_cabs_cext forms them (dda_func.cpp:1694-1722).
__global__ __launch_bounds__(kBlock) void k_cc_terms(Dims d, Geometry g, const float* __restrict__ kinc, IncPhase ip,
                                                     const int* __restrict__ cells, int ncells,
                                                     const int8_t* __restrict__ occ, const float2* __restrict__ alpha,
                                                     int nmat, const float2* __restrict__ einc,
                                                     const float2* __restrict__ P, float* __restrict__ terms) {
    const int rhs = blockIdx.y;
    const float kx = kinc[3 * rhs], ky = kinc[3 * rhs + 1], kz = kinc[3 * rhs + 2];
    const long long rb = (long long)rhs * 3 * d.n;
    const float2 e0 = cconj(einc[rhs * 3]), e1 = cconj(einc[rhs * 3 + 1]), e2 = cconj(einc[rhs * 3 + 2]);
    const float r4 = (float)(2. / 3.);
    const int sl = __builtin_ctz(d.nl), sj = __builtin_ctz(d.nj);
    float* t1 = terms + (long long)rhs * 2 * ncells;
    float* t2 = t1 + ncells;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < ncells; e += gridDim.x * blockDim.x) {
        const int k = cells[e];
        const int l = k & (d.nl - 1), ij = k >> sl;
        const int j = ij & (d.nj - 1), i = ij >> sj;
        const float2 a = alpha[(long long)rhs * nmat + occ[k]];
        const float rx = (float)i * g.d0, ry = (float)j * g.d0, rz = (float)l * g.d0;
        const float r3 = -(kx * (rx - g.ex) + ky * (ry - g.ey) + kz * (rz - g.ez));
        const int ax = ip.axis[rhs];
        const float2 c0 = ax >= 0    ? ip.tab[((size_t)rhs * 2 + 1) * ip.tn + axis_coord(ax, i, j, l)]
                          : ax == -2 ? ip.cell[((size_t)rhs * 2 + 1) * d.n + k]
                                     : exp_i(r3);
        const float2 px = P[rb + k], py = P[rb + d.n + k], pz = P[rb + 2 * d.n + k];
        const float2 c5 = cadd(cmul(pz, e2), cadd(cmul(px, e0), cmul(py, e1)));
        t1[e] = cmul(c0, c5).y;
        const float nx = px.x * px.x + px.y * px.y, ny = py.x * py.x + py.y * py.y, nz = pz.x * pz.x + pz.y * pz.y;
        const float ix = px.y * px.x - px.x * px.y, iy = py.y * py.x - py.x * py.y, iz = pz.y * pz.x - pz.x * pz.y;
        const float2 q = make_float2(nz + (nx + ny), iz + (ix + iy));
        const float2 ca = make_float2(a.x, -a.y);
        const float nn = ca.x * ca.x + ca.y * ca.y;
        const float qi = (-q.x * ca.y + q.y * ca.x) / nn;
        t2[e] = qi - r4 * q.x;
    }
}

// Sequential float sums: thread t < nsums adds src[t * len + e] for e = 0, 1,
// ..., len - 1 into a float that starts at 0 (dda_func.cpp:1684-1725: r1, r2
// start at 0 and take one += per dipole) and stores it widened to double
// (exact) in out[t].
__global__ __launch_bounds__(64) void k_seq_sum(const float* __restrict__ src, int len, int nsums,
                                                double* __restrict__ out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nsums) return;
    const float* p = src + (long long)t * len;
    float s = 0.f;
    int e = 0;
    for (; e + 16 <= len; e += 16) {   // loads in flight together, adds in order
        float v[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) v[u] = p[e + u];
#pragma unroll
        for (int u = 0; u < 16; ++u) s += v[u];
    }
    for (; e < len; ++e) s += p[e];
    out[t] = (double)s;
}

// Far-field terms of the dipoles [e0, e0 + cnt) of the list, for every
// right-hand side and direction (FarField, dda_func.cpp:1586-1614):
//   c0 = exp(i n.(r - c)),  u = n . P,  term = c0 (P - n u)
// with the reference's float operations.  The phase is the correctly rounded
// float of exp(i arg) for the reference's float argument (glibc's cosf/sinf
// differ from it by one ulp for about 1% of arguments).
// T [cnt][nrhs][ndirs][6] floats (x, y, z as re, im); right-hand side r uses
// the directions dirs + r * dir_stride.
__global__ __launch_bounds__(kBlock) void k_ff_terms(Dims d, Geometry g, const int* __restrict__ cells, int e0, int cnt,
                                                     int ndirs, int nrhs, const float* __restrict__ dirs,
                                                     long long dir_stride, const float2* __restrict__ P,
                                                     float* __restrict__ T) {
    const long long total = (long long)cnt * nrhs * ndirs;
    const int sl = __builtin_ctz(d.nl), sj = __builtin_ctz(d.nj);
    for (long long f = (long long)blockIdx.x * blockDim.x + threadIdx.x; f < total;
         f += (long long)gridDim.x * blockDim.x) {
        const int dd = (int)(f % ndirs);
        const long long er = f / ndirs;
        const int rhs = (int)(er % nrhs), e = (int)(er / nrhs);
        const int k = cells[e0 + e];
        const int l = k & (d.nl - 1), ij = k >> sl;
        const int j = ij & (d.nj - 1), i = ij >> sj;
        const float* nv = dirs + rhs * dir_stride + 3 * dd;
        const float kx = nv[0], ky = nv[1], kz = nv[2];
        const float rx = (float)i * g.d0 - g.cx, ry = (float)j * g.d0 - g.cy, rz = (float)l * g.d0 - g.cz;
        const float r0 = kx * rx + ky * ry + kz * rz;
        const float2 c0 = exp_i(r0);
        const long long rb = (long long)rhs * 3 * d.n;
        const float2 px = P[rb + k], py = P[rb + d.n + k], pz = P[rb + 2 * d.n + k];
        float2 c1 = rmul(kx, px), c2 = rmul(ky, py), c3 = rmul(kz, pz);
        const float2 c4 = cadd(c1, c2);
        c1 = cadd(c3, c4);
        c2 = rmul(kx, c1);
        c3 = rmul(ky, c1);
        const float2 c4b = rmul(kz, c1);
        const float2 o0 = cmul(c0, csub(px, c2)), o1 = cmul(c0, csub(py, c3)), o2 = cmul(c0, csub(pz, c4b));
        float* o = T + f * 6;
        o[0] = o0.x;
        o[1] = o0.y;
        o[2] = o1.x;
        o[3] = o1.y;
        o[4] = o2.x;
        o[5] = o2.y;
    }
}

// Continues the sequential far-field sums over one chunk of terms: thread
// t < nsums (= nrhs * ndirs * 6) adds T[e * nsums + t], e = 0 .. cnt - 1, to
// acc[t] (first chunk: to 0), in dipole order.  acc and out (last chunk
// only, else nullptr) hold the far field [nrhs][ndirs][3] complex float.
__global__ __launch_bounds__(kBlock) void k_ff_seq_sum(const float* __restrict__ T, int cnt, int nsums, int first,
                                                       float* __restrict__ acc, float* __restrict__ out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nsums) return;
    float s = first ? 0.f : acc[t];
    int e = 0;
    for (; e + 16 <= cnt; e += 16) {   // loads in flight together, adds in order
        float v[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) v[u] = T[(long long)(e + u) * nsums + t];
#pragma unroll
        for (int u = 0; u < 16; ++u) s += v[u];
    }
    for (; e < cnt; ++e) s += T[(long long)e * nsums + t];
    acc[t] = s;
    if (out) out[t] = s;   // last chunk: the host-visible result
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
