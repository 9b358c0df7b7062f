// Types shared by the host side of the library and its kernels.
#pragma once

#include <stdint.h>

namespace ddafft {

// Lattice description passed by value to every kernel.
struct Dims {
    int ni, nj, nl;   // lattice cells per axis (powers of two, >= 4)
    int nrhs;         // right-hand sides sharing the operator
    long long n;      // ni*nj*nl
};

// Per right-hand-side solver scalars, resident in device memory.  The long
// double coefficients of DDA_FFT (dda_func.cpp:1085-1088) are carried in FP64
// here; every coefficient is rounded to float before it scales a field, as
// the reference does (dda_func.cpp:1219, 1260, 1319, 1362, 1368, 1424, 1464).
struct SolveState {
    double err;        // discrepancy after the last step (ldError)
    double err_min;    // best discrepancy seen at an iteration start
    double rho_prev;   // re/im pairs follow the same pattern below
    double rho_prev_i;
    double rho, rho_i;      // ldcRho of the current iteration
    double A, A_i;          // ldcA
    double W, W_i;          // ldcW
    double den, den_i;      // CG denominator (S, A S)
    double RR_1;            // AA*: |R|^2 of the previous step
    double pad0;
    float2 pro_c1;          // coefficients consumed by the next prologue/update
    float2 pro_c2;
    float2 cA, cW, cB;
    float rA, rB;           // AA* real coefficients
    int step_min;
    int last;               // last step whose discrepancy was recorded
    int stop_at;            // iterations o >= stop_at do not run
    int snapshot;           // copy P into P_min at the next prologue
    int restored;           // P was replaced by P_min at the end
    int it;                 // iteration the solve is in (0: step 0); the scalar step that closes an
                            // iteration advances it, so the kernels of a replayed iteration read their
                            // iteration from here rather than from a launch argument
};

// Host-visible summary of a finished solve.
struct SolveInfo {
    int steps;
    int step_min;
    int restored;
    int matvecs;
};

}  // namespace ddafft
