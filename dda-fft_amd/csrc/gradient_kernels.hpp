// Adjoint gradients of Cext, Cabs and S11 with respect to the polarizabilities.
//
// The DDA system of one polarization p is K P = e with K = diag(alpha)^-1 - A
// on the occupied cells (FirstApproximation, dda_func.cpp:654-689, and the
// matvec x - alpha (A x), dda_func.cpp:1114-1135).  A(r) is a symmetric 3x3
// tensor, even in r, so K^T = K and the adjoint system K lambda = g is the
// forward system with the source diag(alpha) g: the same FFT matvec and solver
// serve it.  For an objective f(P, alpha) with Wirtinger derivative g = df/dP,
//   df/dalpha_j = (lambda_j . P_j + explicit_j) / alpha_j^2,
// where explicit_j = (w_abs / npol) r0 |P_j|^2 / (2i) comes from the 1/conj(alpha)
// of CabsCext (dda_func.cpp:1716-1722).  Only the converged forward dipoles
// are kept; no iterate is stored.
//
// Right-hand side r of the adjoint batch is objective o = r / npol and
// polarization p = r % npol.
#pragma once

#include "device_types.hpp"
#include "fft_common.hpp"
#include "solver_kernels.hpp"

namespace ddafft {

constexpr int kAdjLds = 6144;   // float2 slots of LDS for phase and weight tiles

// Per-axis phase factors exp(i n_a,x (i d - c_x)) ... of every direction:
// tab[a][ni + nj + nl], the factors FarField multiplies (dda_func.cpp:1594-1595).
__global__ __launch_bounds__(kBlock) void k_axis_phase(Dims d, Geometry g, const float* __restrict__ dirs, int ndirs,
                                                       float2* __restrict__ tab) {
    const int naxis = d.ni + d.nj + d.nl;
    const long long tot = (long long)ndirs * naxis;
    for (long long f = (long long)blockIdx.x * blockDim.x + threadIdx.x; f < tot;
         f += (long long)gridDim.x * blockDim.x) {
        const int dd = (int)(f / naxis), a = (int)(f - (long long)dd * naxis);
        float arg;
        if (a < d.ni)
            arg = dirs[3 * dd] * ((float)a * g.d0 - g.cx);
        else if (a < d.ni + d.nj)
            arg = dirs[3 * dd + 1] * ((float)(a - d.ni) * g.d0 - g.cy);
        else
            arg = dirs[3 * dd + 2] * ((float)(a - d.ni - d.nj) * g.d0 - g.cz);
        tab[f] = exp_i(arg);
    }
}

struct AdjArgs {
    const float2* Pf;       // forward dipoles [npol][3][n]
    const float2* alpha;    // [nadj][nmat]
    const float2* einc;     // [npol][3]
    const float2* cext_c;   // [nadj] (w_ext / npol) r0 / (2i)
    const float* cabs_c;    // [nadj] (w_abs / npol) r0
    const float2* u;        // [nadj][ndirs][3]: w_a (I - n n^T) dM00(a)/dEsc_p
    const float2* tab;      // [ndirs][ni + nj + nl]
    int ndirs, npol, nmat, tdirs;
    float2* b;              // [nadj][3][n] adjoint source diag(alpha) g
    double* partials;       // [nadj][gridDim.x] sum over dipoles of |b_j|
};

// g_j = cext_c psi_j conj(E) + cabs_c w_j conj(P_j) + sum_a phi_j(a) u_a, with
// psi_j = exp(-i k.(r_j - c2)) (the CabsCext phase, dda_func.cpp:1685-1706),
// w_j = Im(alpha)/|alpha|^2 - 2/3 and phi_j(a) the FarField phase; b_j = alpha_j g_j.
// Grid (blocks, nadj); every cell is written (0 off the dipoles).
__global__ __launch_bounds__(kBlock) void k_adj_source(Dims d, Geometry g, const int8_t* __restrict__ occ,
                                                       const int* __restrict__ mv, AdjArgs aa) {
    __shared__ float2 lds[kAdjLds];
    __shared__ double red[16];
    const int r = blockIdx.y, p = r % aa.npol;
    const int naxis = d.ni + d.nj + d.nl;
    const int td = aa.tdirs;
    float2* tab = lds;                      // [td][naxis]
    float2* uu = lds + td * naxis;          // [td][3]
    const float2 ce = aa.cext_c[r];
    const float ca = aa.cabs_c[r];
    const float2 E0 = cconj(aa.einc[p * 3]), E1 = cconj(aa.einc[p * 3 + 1]), E2 = cconj(aa.einc[p * 3 + 2]);
    const long long pb = (long long)p * 3 * d.n, bb = (long long)r * 3 * d.n;
    double acc[1] = {0};
    for (long long base = (long long)blockIdx.x * blockDim.x; base < d.n; base += (long long)gridDim.x * blockDim.x) {
        const long long k = base + threadIdx.x;
        const bool in = k < d.n;
        int i = 0, j = 0, l = 0, m = 0;
        bool dip = false;
        if (in) {
            // lattice dimensions are powers of two: (i, j, l) by shifts
            l = (int)(k & (d.nl - 1));
            const long long ij = k >> __builtin_ctz(d.nl);
            j = (int)(ij & (d.nj - 1));
            i = (int)(ij >> __builtin_ctz(d.nj));
            m = occ[k];
            dip = mv[m] != 0;
        }
        float2 gx = make_float2(0.f, 0.f), gy = gx, gz = gx;
        for (int t0 = 0; t0 < aa.ndirs; t0 += td) {
            const int nt = min(td, aa.ndirs - t0);
            __syncthreads();
            for (int f = threadIdx.x; f < nt * naxis; f += blockDim.x) tab[f] = aa.tab[(long long)t0 * naxis + f];
            for (int f = threadIdx.x; f < nt * 3; f += blockDim.x)
                uu[f] = aa.u[((long long)r * aa.ndirs + t0) * 3 + f];
            __syncthreads();
            if (dip) {
                for (int q = 0; q < nt; ++q) {
                    const float2* t = tab + q * naxis;
                    const float2 ph = cmul_fma(cmul_fma(t[i], t[d.ni + j]), t[d.ni + d.nj + l]);
                    cmac_fma(gx, ph, uu[3 * q]);
                    cmac_fma(gy, ph, uu[3 * q + 1]);
                    cmac_fma(gz, ph, uu[3 * q + 2]);
                }
            }
        }
        if (in) {
            float2 bx = make_float2(0.f, 0.f), by = bx, bz = bx;
            if (dip) {
                const float rx = (float)i * g.d0, ry = (float)j * g.d0, rz = (float)l * g.d0;
                const float ph = -(g.kx * (rx - g.ex) + g.ky * (ry - g.ey) + g.kz * (rz - g.ez));
                const float2 psi = cmul(ce, exp_i(ph));
                const float2 a = aa.alpha[(long long)r * aa.nmat + m];
                const float w = ca * (a.y / (a.x * a.x + a.y * a.y) - 2.f / 3.f);
                gx = cadd(gx, cadd(cmul(psi, E0), rmul(w, cconj(aa.Pf[pb + k]))));
                gy = cadd(gy, cadd(cmul(psi, E1), rmul(w, cconj(aa.Pf[pb + d.n + k]))));
                gz = cadd(gz, cadd(cmul(psi, E2), rmul(w, cconj(aa.Pf[pb + 2 * d.n + k]))));
                bx = cmul(a, gx);
                by = cmul(a, gy);
                bz = cmul(a, gz);
                acc[0] += sqrt((double)bx.x * bx.x + (double)bx.y * bx.y + (double)by.x * by.x +
                               (double)by.y * by.y + (double)bz.x * bz.x + (double)bz.y * bz.y);
            }
            aa.b[bb + k] = bx;
            aa.b[bb + d.n + k] = by;
            aa.b[bb + 2 * d.n + k] = bz;
        }
    }
    block_sum<1>(acc, red);
    if (threadIdx.x == 0) aa.partials[(long long)r * gridDim.x + blockIdx.x] = acc[0];
}

// Scale of each adjoint source: s = sum_b partials[r][b] / na[r] (the sum in
// block order, in FP64), 1 when either is zero; inv_s = 1 / s.
__global__ void k_adj_scale(const double* __restrict__ partials, int nblk, int nadj, const double* __restrict__ na,
                            float* __restrict__ s, float* __restrict__ inv_s) {
    const int r = threadIdx.x;
    if (r >= nadj) return;
    double sb = 0.;
    for (int b = 0; b < nblk; ++b) sb += partials[(long long)r * nblk + b];
    const double sc = (sb > 0. && na[r] > 0.) ? sb / na[r] : 1.;
    s[r] = (float)sc;
    inv_s[r] = (float)(1. / sc);
}

// x[r] *= scale[r] over the 3 n components of right-hand side r.
__global__ __launch_bounds__(kBlock) void k_scale_rhs(Dims d, const float* __restrict__ scale, float2* __restrict__ x) {
    const int r = blockIdx.y;
    const float s = scale[r];
    const long long rb = (long long)r * 3 * d.n;
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < 3 * d.n;
         k += (long long)gridDim.x * blockDim.x)
        x[rb + k] = rmul(s, x[rb + k]);
}

struct GradArgs {
    const float2* lam;      // adjoint solutions of the scaled sources [nadj][3][n]
    const float* s;         // [nadj] source scale (lambda = s * lam)
    const float2* Pf;       // [npol][3][n]
    const float2* alpha;    // [nadj][nmat]
    const float2* dadchi;   // [npol][nmat] d alpha_p / d chi
    const float2* expl;     // [nadj] (w_abs / npol) r0 / (2i)
    int npol, nmat;
    float2* H;              // [nadj][n]  df/dalpha_j per right-hand side
    float2* gchi;           // [nobj][n]  (df/dRe chi_j, df/dIm chi_j)
};

// Grid (blocks, nobj).
__global__ __launch_bounds__(kBlock) void k_grad(Dims d, const int8_t* __restrict__ occ, const int* __restrict__ mv,
                                                 GradArgs ga) {
    const int o = blockIdx.y;
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < d.n;
         k += (long long)gridDim.x * blockDim.x) {
        const int m = occ[k];
        float2 q = make_float2(0.f, 0.f);
        if (mv[m] != 0) {
            for (int p = 0; p < ga.npol; ++p) {
                const int r = o * ga.npol + p;
                const long long lb = (long long)r * 3 * d.n, pb = (long long)p * 3 * d.n;
                float2 lp = make_float2(0.f, 0.f);
                float pp = 0.f;
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const float2 pv = ga.Pf[pb + c * d.n + k];
                    lp = cadd(lp, cmul(ga.lam[lb + c * d.n + k], pv));
                    pp += pv.x * pv.x + pv.y * pv.y;
                }
                const float2 num = cadd(rmul(ga.s[r], lp), rmul(pp, ga.expl[r]));
                const float2 a = ga.alpha[(long long)r * ga.nmat + m];
                const float2 a2 = cmul(a, a);
                const float den = a2.x * a2.x + a2.y * a2.y;
                const float2 h = make_float2((num.x * a2.x + num.y * a2.y) / den, (num.y * a2.x - num.x * a2.y) / den);
                ga.H[(long long)r * d.n + k] = h;
                q = cadd(q, cmul(h, ga.dadchi[p * ga.nmat + m]));
            }
        } else {
            for (int p = 0; p < ga.npol; ++p) ga.H[(long long)(o * ga.npol + p) * d.n + k] = make_float2(0.f, 0.f);
        }
        ga.gchi[(long long)o * d.n + k] = make_float2(2.f * q.x, -2.f * q.y);
    }
}

// Per-material sums of H: partials[r][gridDim.x][nmat][2] in FP64.
// Grid (blocks, nadj).
__global__ __launch_bounds__(kBlock) void k_mat_sum(Dims d, const int8_t* __restrict__ occ, const float2* __restrict__ H,
                                                    int nmat, double* __restrict__ partials) {
    __shared__ double red[2 * 16];
    const int r = blockIdx.y;
    const long long hb = (long long)r * d.n;
    for (int m = 1; m < nmat; ++m) {
        double acc[2] = {0, 0};
        for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < d.n;
             k += (long long)gridDim.x * blockDim.x) {
            if (occ[k] == m) {
                const float2 h = H[hb + k];
                acc[0] += h.x;
                acc[1] += h.y;
            }
        }
        block_sum<2>(acc, red);
        if (threadIdx.x == 0) {
            double* o = partials + (((long long)r * gridDim.x + blockIdx.x) * nmat + m) * 2;
            o[0] = acc[0];
            o[1] = acc[1];
        }
        __syncthreads();
    }
}

}  // namespace ddafft
