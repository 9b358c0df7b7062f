/*
 * ddafft.h -- C ABI of the MI355X DDA-FFT solve path (libddafft_amd.so).
 *
 * The reference (physicist2018/dda-fft v7.2) is one C++ program with no
 * plugin or FFI layer; its solve path sits behind the C++ functions cited
 * beside each entry point below.  This header is the drop-in boundary for
 * that path: plain pointers and sizes, integer status codes, no exit().
 *
 * Data layouts (caller-owned host buffers, as the reference stores them):
 *   occupancy  int8  [ni][nj][nl]         material index per cell, 0 = vacuum
 *   field      float [ni][nj][nl][3][2]   complex 3-vector per cell (cvector)
 *   ahat       float [2ni][2nj][2nl][6][2] (e00 e01 e02 e11 e12 e22)
 *   alpha      float [nmat][2]            polarizability per material index,
 *                                         alpha[0] is vacuum (0)
 *   mv         int   [nmat]               1 if material index m is a dipole
 *   directions float [ndirs][3]           unit scattering directions
 *   esc        float [ndirs][3][2]        scattered far field per direction
 *
 * Threading: a context owns one HIP stream and its device buffers; calls on
 * one context must not overlap.  Different contexts may be used from
 * different host threads.
 */
#ifndef DDAFFT_H
#define DDAFFT_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DDAFFT_OK 0
#define DDAFFT_EINVAL (-1)   /* bad argument (sizes, pointers, not a power of two) */
#define DDAFFT_ENOMEM (-2)   /* device or host allocation failed */
#define DDAFFT_EHIP (-3)     /* HIP runtime error */
#define DDAFFT_ENODEV (-4)   /* no usable GPU */
#define DDAFFT_ESTATE (-5)   /* call out of order (e.g. solve before set_target) */

#define DDAFFT_SOLVER_CG 0        /* solver codes of DDA_FFT (options.cpp:80-86) */
#define DDAFFT_SOLVER_AA 1
#define DDAFFT_SOLVER_BICGSTAB 2
#define DDAFFT_SOLVER_SIMPLE 3

typedef struct ddafft_ctx ddafft_ctx;

/* Message of the last failing call on this host thread. */
const char* ddafft_last_error(void);

/* Library version, major*10000 + minor*100 + patch. */
int ddafft_version(void);

/* Page-locked host memory for inputs and results (no counterpart in the
 * reference).  Results copied into such buffers (dipoles, grad_chi, far
 * fields) move at DMA speed instead of through the runtime's staging copies
 * of pageable memory.  Free with ddafft_host_free. */
int ddafft_host_alloc(size_t bytes, void** out);
int ddafft_host_free(void* p);

/* Creates a context for an ni x nj x nl lattice (powers of two, 4..256 per
 * axis) with lattice spacing kd, able to solve up to nrhs_max right-hand
 * sides at once, and builds Â on the device.
 * Replaces main.cpp:390-476 (cmA allocation, CreateAMatrix, FourierPhase
 * tables, FFT_3D(cmatrix)) and dda_workspace_create (dda_func.cpp:1031-1054). */
int ddafft_create(int ni, int nj, int nl, float kd, int nrhs_max, ddafft_ctx** out);

/* Frees every device buffer of the context (dda_workspace_free,
 * dda_func.cpp:1056-1069, and the frees of main.cpp:1030-1077). */
int ddafft_destroy(ddafft_ctx* ctx);

/* Copies Â back in the reference's layout (cmA of main.cpp:390-409). */
int ddafft_get_ahat(ddafft_ctx* ctx, float* ahat);

/* Sets the occupancy map and the dipole flags of the materials
 * (cAlpha / iListMV arguments of DDA_FFT, function.hpp:153-160). */
int ddafft_set_target(ddafft_ctx* ctx, const signed char* occ, int nmat, const int* mv);

/* Makes `dirs` the resident scattering directions used by ddafft_solve when
 * it is given dirs == NULL (ksc argument of DDA_FFT). */
int ddafft_set_directions(ddafft_ctx* ctx, int ndirs, const float* dirs);

/* y = first octant of IFFT(Â ⊙ FFT(pad x)), or with conj(Â(-k)) when conj
 * is non-zero: CopyingBack + Convolution / Convolution_conj
 * (dda_func.cpp:702-718, 150-187, 227-437). */
int ddafft_convolution(ddafft_ctx* ctx, const float* x, float* y, int conj);

/* y = x - alpha_m * (A x) on dipole cells and 0 elsewhere (conj: with
 * conj(alpha_m) and A^H) -- the matvec epilogue of DDA_FFT
 * (dda_func.cpp:1325-1357, 1386-1416). */
int ddafft_matvec(ddafft_ctx* ctx, const float* alpha, const float* x, float* y, int conj);

/* One DDA_FFT solve per right-hand side (dda_func.cpp:1075-1570,
 * function.hpp:153-160): nrhs systems share the lattice, the occupancy and
 * kinc; each has its own alpha [nrhs][nmat][2] and incident field
 * einc [nrhs][3][2].  Outputs per right-hand side: esc [nrhs][ndirs][3][2],
 * cabs [nrhs], cext [nrhs], err_hist [nrhs][max_step+1] (the discrepancy
 * printed after each step; optional) and info [nrhs][4] = {last step,
 * best step, 1 if the best iterate was restored, matvecs} (optional).
 * dirs == NULL uses the directions of ddafft_set_directions. */
int ddafft_solve(ddafft_ctx* ctx, int nrhs, int solver, float tau, int max_step, float err_max,
                 const float* alpha, const float* kinc, const float* einc, int ndirs, const float* dirs,
                 float* esc, float* cabs, float* cext, double* err_hist, int* info);

/* ddafft_solve that allocates nothing.  No counterpart in the reference: it
 * is for callers that cannot allocate memory during a solve (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments and results as ddafft_solve, bit for bit (except the far
 * field of a direction set of scattering planes when the grouped far field's
 * gathered dipole moments, 48 B per dipole for two right-hand sides, do not
 * fit the scratch: then the per-direction kernel, within about 1e-7).  Its
 * scratch (block partials, discrepancy history, far-field partials and phase
 * tables, per-call directions, results on their way back, reference-order
 * sum terms, gathered moments) lives in a fixed 32 MiB device array, one per GPU, that every such call on
 * that GPU shares: calls that overlap on one GPU race.  A solve whose
 * scratch would exceed it fails with DDAFFT_EINVAL and a message giving the
 * size needed; results are limited to 1 MiB per call.  Inputs are copied from
 * the caller's memory (no page-locked staging is allocated). */
int ddafft_solve_noalloc(ddafft_ctx* ctx, int nrhs, int solver, float tau, int max_step, float err_max,
                         const float* alpha, const float* kinc, const float* einc, int ndirs, const float* dirs,
                         float* esc, float* cabs, float* cext, double* err_hist, int* info);
/* Dipole moments P of right-hand side `rhs` after the last solve. */
int ddafft_get_dipoles(ddafft_ctx* ctx, int rhs, float* p);

/* Cabs and Cext of a given dipole field (CabsCext, dda_func.cpp:1679-1727). */
int ddafft_cabs_cext(ddafft_ctx* ctx, const float* alpha, const float* kinc, const float* einc,
                     const float* p, float* cabs, float* cext);

/* Far field of a given dipole field in ndirs directions
 * (FarField, dda_func.cpp:1576-1625). */
int ddafft_far_field(ddafft_ctx* ctx, const float* p, int ndirs, const float* dirs, float* esc);

/* How the observables of every later solve on this context are summed
 * (Cext, Cabs: CabsCext, dda_func.cpp:1679-1727; far field: FarField,
 * dda_func.cpp:1576-1625).  mode 0 (default): the reference's per-dipole
 * float terms summed in FP64.  mode 1: the reference's own sums, sequential
 * float accumulation over the dipoles in its cell order from the same float
 * terms (far-field phases: the correctly rounded float of exp(i arg) for the
 * reference's float argument); a parity mode, slower than mode 0. */
int ddafft_set_sum_mode(ddafft_ctx* ctx, int mode);

/* How Â is stored (replaces the Â of main.cpp:466-476: CreateAMatrix,
 * dda_func.cpp:85-148, and FFT_3D(cmatrix), fft_func.cpp:317-416; rebuilt
 * on the call).  mode 0 (default): the reference's Â over the whole padded
 * grid, 48 B per padded cell; every convolution is bit-identical to the
 * reference's.  mode 1: the padded A(r) with its Nyquist planes (r = -n d
 * along any axis, cells that never pair two lattice sites) set to zero, which
 * makes Â even or odd along every axis; only the octant k <= n of its
 * transform is stored (1/8 of the memory, and of the X pass's Â reads) and
 * mirrored in the X pass.  The operator is mathematically the reference's;
 * results differ from it by float rounding.  ddafft_get_ahat returns the
 * mirrored grid. */
int ddafft_set_ahat_mode(ddafft_ctx* ctx, int mode);
/* Launch schedule of the passes (no counterpart in the reference: the
 * knobs of the autotuner, dda-fft_amd/tune.py, which measures them on the
 * GPU per lattice and right-hand-side count and writes the table the library
 * uses by default, dda-fft_amd/csrc/tuned_table.hpp).  A value set here wins
 * over the table for every later call on this context; 0 (-1 for
 * DDAFFT_TUNE_GRAPH and DDAFFT_TUNE_FUSE_ZY) hands the choice back to the
 * table.
 *   DDAFFT_TUNE_X_RB: the fused X pass works on groups of 1, 2 or 4
 *     right-hand sides that share each Â value read;
 *   DDAFFT_TUNE_X_BLOCKS: each X block takes a tile of (J, L) columns and a
 *     run of groups, reading the tile's Â once for the whole run; groups are
 *     spread over more blocks (shorter runs) until a launch has this many
 *     blocks;
 *   DDAFFT_TUNE_GRAPH: 1 replays each solver iteration as a captured HIP
 *     graph (its kernels read their iteration from the device-resident
 *     solver states), 0 launches every kernel of every iteration. */
#define DDAFFT_TUNE_X_BLOCKS 0
#define DDAFFT_TUNE_X_RB 1
#define DDAFFT_TUNE_GRAPH 2
/* DDAFFT_TUNE_FUSE_ZY: 1 runs the z- and y-forward transforms of one
 * i-plane in one kernel where n_j, n_l <= 64, 0 always as two passes, -1
 * (default) as the table says. */
#define DDAFFT_TUNE_FUSE_ZY 3
/* DDAFFT_TUNE_FF_MFMA: 1 sums the far field (FP64-summed mode) on the matrix
 * cores, 16 directions per block, where the phase rows fit (ni + nj + nl <=
 * 384); 0 with the direct kernel; -1 (default) as the library chooses
 * (the direct kernel).  The two agree to about 1e-7 relative, not bit for
 * bit. */
#define DDAFFT_TUNE_FF_MFMA 4
/* DDAFFT_TUNE_FF_GROUP: 1 (default, -1) sums the far field (FP64-summed mode)
 * of a direction set whose z components repeat with a period ng, i.e. ng
 * groups of 2..16 directions sharing z (the scattering planes of a fixed
 * orientation), by groups: the sum over l once per group, then per direction
 * over the (i, j) columns; lattices with n_i + n_j <= 512 and n_l <= 256.
 * 0 always uses the per-direction kernel.  The two agree to about 1e-7
 * relative, not bit for bit. */
#define DDAFFT_TUNE_FF_GROUP 5
int ddafft_set_tuning(ddafft_ctx* ctx, int key, int value);
/* One instance's solve for ddafft_solve_ex / ddafft_solve_batch: nrhs
 * right-hand sides, each with its own incidence (Monte-Carlo orientations x
 * polarizations of one target, main.cpp:875-940), optionally each with its
 * own (rotated) scattering directions.  Outputs as ddafft_solve's. */
typedef struct {
    int nrhs;
    const float* alpha;   /* [nrhs][nmat][2] */
    const float* kinc;    /* [nrhs][3] */
    const float* einc;    /* [nrhs][3][2] */
    int ndirs;
    const float* dirs;    /* [ndirs][3], or [nrhs][ndirs][3] when dirs_per_rhs; NULL: resident set */
    int dirs_per_rhs;
    float* esc;           /* [nrhs][ndirs][3][2] */
    float* cabs;          /* [nrhs] (may be NULL) */
    float* cext;          /* [nrhs] (may be NULL) */
    int* info;            /* [nrhs][4] (may be NULL) */
    int exact_phases;     /* 1: every right-hand side's incident phases exp(i k.(r - c)) are evaluated on
                             the host with the C library's cosf/sinf on the reference's float argument
                             (dda_func.cpp:663-675, 1694-1700), so the first approximation and the cross
                             sections match the reference bit for bit for any incidence, at a host cost of
                             two sincos per dipole and right-hand side; 0: only axis-aligned incidences
                             are tabulated so, others use FP64 sincos rounded to float */
} ddafft_job;

/* ddafft_solve with per-right-hand-side incidence and directions. */
int ddafft_solve_ex(ddafft_ctx* ctx, const ddafft_job* job, int solver, float tau, int max_step, float err_max,
                    double* err_hist);

/* Solves n instances (distinct contexts, e.g. random realizations of uneven
 * lattices, each with its own A-hat) together: up to `width` of them (0: the
 * autotuner's choice, dda-fft_amd/csrc/tuned_table.hpp)
 * run at once, each on its own stream, one iteration per instance per round,
 * each right-hand side stopping on its own (the reference's sample loop,
 * main.cpp:774-996, batched).  Results as ddafft_solve_ex for each job. */
int ddafft_solve_batch(int n, ddafft_ctx* const* ctxs, const ddafft_job* jobs, int solver, float tau, int max_step,
                       float err_max, int width);

/* Forward + adjoint gradients (no counterpart in the reference, whose forward
 * model DDA_FFT, dda_func.cpp:1075-1570, CabsCext :1679-1727, FarField
 * :1576-1625 and BuildMatrixJones/JonesToMuller, mmuller.cpp:175-314, they
 * differentiate).  Solves the npol (1 or 2) polarizations as ddafft_solve
 * does, keeps their dipoles, then solves the adjoint systems of nobj
 * objectives as npol*nobj more right-hand sides of the same operator
 * (K = diag(alpha)^-1 - A is complex symmetric) and assembles
 *   f[o] = w_ext * mean_p Cext_p + w_abs * mean_p Cabs_p + sum_a w_a * S11(a)
 * with weights[o] = {w_ext, w_abs, w_s11[ndirs]} and S11(a) = M00 of
 * direction a (before plane averaging).  Outputs:
 *   grad_m   [nobj][nmat][2]        df/dRe m, df/dIm m per material
 *   grad_chi [nobj][ni][nj][nl][2]  df/dRe chi, df/dIm chi per dipole,
 *                                    chi = m^2 - 1 (0 on empty cells)
 *   f [nobj], s11 [ndirs], esc/cabs/cext of the forward solve,
 *   info [npol + npol*nobj][4] (forward, then adjoint right-hand sides).
 * alpha and dalpha_dm are [npol][nmat][2] (ddafft_polarizability and
 * ddafft_polarizability_deriv per polarization), m is [nmat][2];
 * normals [ndirs][3] is the scattering-plane normal of each direction
 * (needed for S11).  npol * nobj <= nrhs_max.  Any output may be NULL. */
int ddafft_gradient(ddafft_ctx* ctx, int npol, int nobj, int solver, float tau, int max_step, float err_max,
                    const float* alpha, const float* dalpha_dm, const float* m, const float* kinc,
                    const float* einc, int ndirs, const float* dirs, const float* normals, const float* weights,
                    float* f, float* esc, float* cabs, float* cext, float* s11, float* grad_m, float* grad_chi,
                    int* info);

/* Opt-in timing of every matvec pass with HIP events on the context's
 * stream.  Returns, in pass_ms[5] and *matvecs, the summed durations of the
 * five passes (z-fwd, y-fwd, fused x, y-inv, z-inv) and the number of matvecs
 * timed since profiling was last enabled; then enables (enable != 0, with
 * the sums reset) or disables profiling.  Benchmark aid; no counterpart in
 * the reference. */
int ddafft_profile(ddafft_ctx* ctx, int enable, double* pass_ms, long long* matvecs);

/* Times `reps` matvecs of nrhs right-hand sides on device-resident data with
 * HIP events on the context's stream.  kernel_ms[5] receives the average
 * duration of each pass (z-fwd, y-fwd, fused x, y-inv, z-inv) and
 * total_ms the average wall time of one matvec.  Test and benchmark aid. */
int ddafft_time_matvec(ddafft_ctx* ctx, int nrhs, int reps, float* kernel_ms, float* total_ms);

/* ------------------------------------------------------------------------
 * Host-side pieces (no GPU needed), bit-exact with the reference's float
 * arithmetic.
 * ------------------------------------------------------------------------ */

/* Occupancy of the deterministic targets; return the number of occupied
 * cells or -1 if the body does not fit: iGenerateEllipsoid (particle.cpp:408-444,
 * targets 1 and 2), iGenerateParallelepiped (:501-527, target 3),
 * iGenerateBisphere (:448-497, target 12). */
int ddafft_target_ellipsoid(int ni, int nj, int nl, double b_a, double c_a, signed char* occ);
int ddafft_target_parallelepiped(int ni, int nj, int nl, double b_a, double c_a, signed char* occ);
int ddafft_target_bisphere(int ni, int nj, int nl, double r1_r2, signed char* occ);

/* The reference's random stream (xorshift64*, randomize.cpp:44-58): a state
 * of 0 is replaced by 0x9E3779B97F4A7C15 as dda_srand does.  A worker's seed
 * is base + 0x9E3779B97F4A7C15 * (worker + 1) (main.cpp:803-804). */
unsigned long long ddafft_rng_worker_seed(unsigned long long base_seed, int worker);
int ddafft_rng_next(unsigned long long* state);

/* Random and parametric targets with the reference's semantics and float
 * rounding; *state is the random stream, advanced as the reference advances
 * it.  Each returns what the reference's generator returns (its dipole count
 * convention, -1 if the body does not fit), or -2 on bad arguments.
 *   random ellipsoid      iGenerateRandomEllipsoid (particle.cpp:531-613), targets 4 and 5
 *   random parallelepiped iGenerateRandomParallelepiped (:616-694), target 9
 *   nibbled ellipsoid     iGenerateNibbleEllipsoid (:696-822), target 10
 *   hierarchical          iGenerateHierarchicalEllipsoid (:824-962), target 13
 *   partial sphere        iGenerateSphereSpec (:965-994), target 14 (radius = ni / 2)
 *   worms, kind 6/7/8     iGenerateAnthrax / Hedgehog / DeadAnthrax (:21-231); with
 *                         max_attempts > 0 the placement gives up (-3) after that
 *                         many draws, where the reference can loop forever. */
int ddafft_target_random_ellipsoid(int ni, int nj, int nl, double b_a, double c_a, int nsorts, const int* nseeds,
                                   unsigned long long* state, signed char* occ);
int ddafft_target_random_parallelepiped(int ni, int nj, int nl, double b_a, double c_a, int nsorts,
                                        const int* nseeds, unsigned long long* state, signed char* occ);
int ddafft_target_nibble_ellipsoid(int ni, int nj, int nl, double b_a, double c_a, int nsorts, const int* nseeds,
                                   int bites, double depth, const int* mv, unsigned long long* state,
                                   signed char* occ);
int ddafft_target_hierarchical_ellipsoid(int levels, int ni, int nj, int nl, double b_a, double c_a, int nsorts,
                                         const int* nseeds, unsigned long long* state, signed char* occ);
int ddafft_target_sphere_spec(int ni, int nj, int nl, double radius, double rho, signed char* occ);
int ddafft_worm_count(int ni, int nj, int nl, double rho, double radius_worm, double length_ratio);
int ddafft_target_worms(int kind, int ni, int nj, int nl, double radius_worm, double length_worm, int nworms,
                        long long max_attempts, unsigned long long* state, signed char* occ);

/* Generators of benchmark workloads the reference has no target for; their
 * maps reach the reference through the target 11 shape file.
 *   mixture       occ[k] = base[k], and every cell with base[k] == 1 becomes
 *                 `material` when its draw u = next / (2^31 - 1) is below
 *                 `fraction` (one draw per such cell, in storage order).
 *                 Returns the number of cells changed.
 *   agglomerate   ddafft_agglomerate_count(...) = round(packing (R / r_mono)^3)
 *                 non-overlapping spheres of radius r_mono (cells) with
 *                 centres drawn uniformly in the ball of radius R - r_mono,
 *                 R = min(ni, nj, nl) / 2, about the lattice centre
 *                 ((ni - 1) / 2, ...), by rejection: each centre takes three
 *                 draws c = (R - r_mono)(2u - 1) per axis and is kept when it
 *                 lies in that ball and at least 2 r_mono from every kept one;
 *                 after 100000 consecutive rejected draws (a jammed packing)
 *                 all kept centres are dropped and placement starts over.
 *                 A cell is material 1 when within r_mono of a centre.
 *                 Returns the occupied count, -3 after max_attempts draws of a
 *                 centre (max_attempts > 0).  centres, when not NULL, receives
 *                 the 3 * ddafft_agglomerate_count(...) centre coordinates. */
int ddafft_target_mixture(int n, const signed char* base, int material, double fraction,
                          unsigned long long* state, signed char* occ);
int ddafft_agglomerate_count(int ni, int nj, int nl, double r_mono, double packing);
int ddafft_target_agglomerate(int ni, int nj, int nl, double r_mono, double packing, long long max_attempts,
                              unsigned long long* state, signed char* occ, double* centres);

/* Target 11 shape file (main.cpp:645-677): the count N, then N lines
 * "X Y Z T".  read returns N, -3 for a cell outside the lattice (the
 * reference's exit code 3), -4 for a short file; write lists the nonzero cells. */
int ddafft_target_file_read(const char* path, int ni, int nj, int nl, signed char* occ);
int ddafft_target_file_write(const char* path, int ni, int nj, int nl, const signed char* occ);

/* Monte-Carlo orientation (main.cpp:875-893): draws the Euler angles from
 * *state and returns the 3x3 matrix rot[9] (row-major) that main.cpp applies
 * to kinc, Einc, the directions and the plane normals; ddafft_rotate applies
 * it to n vectors as main.cpp:895-921 does. */
int ddafft_orientation_mc(unsigned long long* state, float* rot);
int ddafft_rotate(const float* rot, int n, const float* in, float* out);

/* Polarizability model 1..5 (CM, CMRR, CMNR, ILDR, LDR) of refractive index
 * m[2] at lattice spacing kd (Polarizability_*, dda_func.cpp:1631-1677, as
 * selected by main.cpp:606-631 and 926-931); kinc[3] and einc[3][2] are used
 * by model 5 only. */
int ddafft_polarizability(int model, float kd, const float* m, const float* kinc, const float* einc, float* alpha);

/* Holomorphic derivative d alpha / d m [2] of the same models, evaluated in
 * double (the formulas of dda_func.cpp:1631-1677 differentiated); for the
 * gradient entry point. */
int ddafft_polarizability_deriv(int model, float kd, const float* m, const float* kinc, const float* einc,
                                float* dalpha_dm);

/* Scattering planes, directions and incident field for a fixed orientation
 * (main.cpp:354-387 and 516-542): dirs [nplanes][nangles][3],
 * normals [nplanes][3], kinc [3], einc [2][3][2].  Angles in degrees. */
int ddafft_geometry_fixed(int nplanes, int nangles, double theta_min, double theta_step, double psi,
                          double iota, double kappa, float* dirs, float* normals, float* kinc, float* einc);

/* Mueller matrix [4][4] of one direction from the far fields of the two
 * incident polarizations (BuildMatrixJones + JonesToMuller, mmuller.cpp:175-314). */
int ddafft_muller(const float* kinc, const float* ksc, const float* normal, const float* einc0,
                  const float* einc1, const float* esc0, const float* esc1, float* mm);

/* Plane-averaged Mueller matrices of one sample (main.cpp:943-954):
 * mm [nangles][4][4]; esc0/esc1 [nplanes][nangles][3][2]. */
int ddafft_muller_average(int nplanes, int nangles, const float* kinc, const float* dirs,
                          const float* normals, const float* einc, const float* esc0, const float* esc1,
                          float* mm);

#ifdef __cplusplus
}
#endif

#endif /* DDAFFT_H */
