// Scalar steps of the DDA_FFT solvers (dda_func.cpp:1075-1570): the
// fixed-order FP64 sums of a reduction's block partials and the coefficients
// and stopping decision computed from them.
#pragma once

#include "device_types.hpp"
#include "fft_common.hpp"

namespace ddafft {

constexpr int kSlots = 6;  // partial sums: dot1 (re,im), dot2 (re,im), real1, err

enum FinKind : int {
    FK_BICG_0 = 0,   // rho_prev = rho = (R,G)                       (:1183)
    FK_BICG_A = 1,   // A = rho/(G,H); pro_c1 = -cA                   (:1218-1222, 1421-1426)
    FK_BICG_W = 2,   // W = (W,R)/(W,W)                               (:1256-1260, 1461-1464)
    FK_BICG_E = 3,   // err; next rho; B; pro_c1 = cB, pro_c2 = -(cB cW)   (:1280, 1315-1323)
    FK_CG_E0 = 4,    // err after step 0                              (:1280)
    FK_CG_A = 5,     // den = (S,W); A = (R,S)/den                    (:1360-1362)
    FK_CG_E = 6,     // B = (R,W)/den; pro_c1 = -cB; err              (:1367-1370, 1523)
    FK_AA_0 = 7,     // RR_1 = |R|^2                                  (:1141)
    FK_AA_E0 = 8,    // err after step 0
    FK_AA_A = 9,     // rA = -RR_1/|S|^2                              (:1374-1375)
    FK_AA_E = 10,    // RR; rB = RR/RR_1; RR_1 = RR; err              (:1382-1383, 1523)
    FK_SIM_E = 11,   // err                                           (:1523)
};

struct FinArgs {
    int kind;
    const double* partials;
    int nblk;              // partial blocks per right-hand side
    double* hist;          // [nrhs][max_step+1]
    int max_step;
    double err_max;
    double inv_ndip;       // 1 / number of dipoles (ldError's n)
    SolveState* ring;      // steps that end an iteration: the host-visible copies of the states,
                           // [4][ring_stride] in page-locked host memory (slot: iteration & 3), else nullptr
    int ring_stride;
};

// Publishes one right-hand side's state after iteration `iter` to the
// host-visible ring.  The host reads it only after the event recorded behind
// the iteration has completed, which orders the kernel's stores to host
// memory before it (no fence here: a system-scope fence writes back the
// whole L2).
__device__ __forceinline__ void publish_state(const FinArgs& fa, int rhs, int iter, const SolveState& s) {
    if (fa.ring) fa.ring[(size_t)(iter & 3) * fa.ring_stride + rhs] = s;
}

// End-of-step kinds: the scalar step that closes an iteration (or step 0).
__host__ __device__ constexpr bool ends_step(int kind) {
    return kind == 3 || kind == 4 || kind == 6 || kind == 8 || kind == 10 || kind == 11;
}

__device__ __forceinline__ void ld_div(double ar, double ai, double br, double bi, double& cr, double& ci) {
    const double n = br * br + bi * bi;
    cr = (ar * br + ai * bi) / n;
    ci = (ai * br - ar * bi) / n;
}

__device__ __forceinline__ float2 to_f(double r, double i) { return make_float2((float)r, (float)i); }

// End-of-step bookkeeping shared by every solver: records the discrepancy of
// step `o` and decides whether iteration o+1 runs (dda_func.cpp:1280-1313).
__device__ void end_step(SolveState& s, const FinArgs& fa, int rhs, int o, double err) {
    s.err = err;
    s.last = o;
    fa.hist[(long long)rhs * (fa.max_step + 1) + o] = err;
    if (o == 0) {
        s.err_min = err;
        s.step_min = 0;
        s.snapshot = 1;
    } else {
        s.snapshot = 0;
    }
    if (o + 1 > fa.max_step || err <= fa.err_max) {
        s.stop_at = o + 1;
    } else if (err < s.err_min) {
        s.err_min = err;
        s.step_min = o;
        s.snapshot = 1;
    }
}

// The fixed-order FP64 sum of one right-hand side's block partials and the
// scalar step of its solver; run by every thread of one block.
__device__ __forceinline__ void finalize_block(SolveState* __restrict__ st, int iter, const FinArgs& fa, int rhs,
                                               double* red) {
    double v[kSlots] = {0, 0, 0, 0, 0, 0};
    const double* p = fa.partials + (long long)rhs * fa.nblk * kSlots;
    // each thread sums partials b = tid, tid + 256, ... in that order; the
    // loads of four of them are in flight together
    constexpr int kBatch = 4;
    for (int b0 = threadIdx.x; b0 < fa.nblk; b0 += kBatch * kBlock) {
        double t[kBatch][kSlots];
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
            const int b = b0 + u * kBlock;
#pragma unroll
            for (int q = 0; q < kSlots; ++q) t[u][q] = b < fa.nblk ? p[(long long)b * kSlots + q] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < kBatch; ++u)
            if (b0 + u * kBlock < fa.nblk)
#pragma unroll
                for (int q = 0; q < kSlots; ++q) v[q] += t[u][q];
    }
    block_sum<kSlots>(v, red);
    if (threadIdx.x != 0) return;
    SolveState& s = st[rhs];
    const double err = v[5] * fa.inv_ndip;
    switch (fa.kind) {
        case FK_BICG_0:
            s.rho_prev = s.rho = v[0];
            s.rho_prev_i = s.rho_i = v[1];
            break;
        case FK_BICG_A: {
            ld_div(s.rho, s.rho_i, v[0], v[1], s.A, s.A_i);
            s.cA = to_f(s.A, s.A_i);
            s.pro_c1 = cneg(s.cA);
            break;
        }
        case FK_BICG_W: {
            ld_div(v[0], v[1], v[2], v[3], s.W, s.W_i);
            s.cW = to_f(s.W, s.W_i);
            break;
        }
        case FK_BICG_E: {
            if (iter > 0) {
                s.rho_prev = s.rho;
                s.rho_prev_i = s.rho_i;
            }
            s.rho = v[0];
            s.rho_i = v[1];
            // ldcB = (rho*A)/(rho_prev*W)
            const double nr = s.rho * s.A - s.rho_i * s.A_i, ni = s.rho * s.A_i + s.rho_i * s.A;
            const double dr = s.rho_prev * s.W - s.rho_prev_i * s.W_i, di = s.rho_prev * s.W_i + s.rho_prev_i * s.W;
            double br, bi;
            ld_div(nr, ni, dr, di, br, bi);
            s.cB = to_f(br, bi);
            s.pro_c1 = s.cB;
            s.pro_c2 = cneg(cmul(s.cB, s.cW));
            end_step(s, fa, rhs, iter, err);
            break;
        }
        case FK_CG_E0:
        case FK_AA_E0:
        case FK_SIM_E:
            end_step(s, fa, rhs, iter, err);
            break;
        case FK_CG_A: {
            s.den = v[0];
            s.den_i = v[1];
            ld_div(v[2], v[3], s.den, s.den_i, s.A, s.A_i);
            s.cA = to_f(s.A, s.A_i);
            break;
        }
        case FK_CG_E: {
            double br, bi;
            ld_div(v[0], v[1], s.den, s.den_i, br, bi);
            s.cB = to_f(br, bi);
            s.pro_c1 = cneg(s.cB);
            end_step(s, fa, rhs, iter, err);
            break;
        }
        case FK_AA_0:
            s.RR_1 = v[4];
            break;
        case FK_AA_A:
            s.rA = (float)(-s.RR_1 / v[4]);
            break;
        case FK_AA_E: {
            const double rr = v[4];
            s.rB = (float)(rr / s.RR_1);
            s.RR_1 = rr;
            end_step(s, fa, rhs, iter, err);
            break;
        }
        default:
            break;
    }
    if (ends_step(fa.kind)) {
        s.it = iter + 1;
        publish_state(fa, rhs, iter, s);
    }
}

}  // namespace ddafft
