// Test-only harness around the reference's own functions.
//
// This file is oracle infrastructure: it is compiled together with the
// reference sources (dda_func.cpp, fft_func.cpp, complex.cpp, cvector.cpp,
// mmuller.cpp, particle.cpp, randomize.cpp, read where they lie) into
// oracle/_ref/libddaref.so by oracle/Makefile, and is used only to write the
// golden fixtures under tests/golden/ and to cross-check the C restatement in
// oracle/ddaoracle.c.  Nothing in the product links it.
//
// Every entry point takes flat, caller-owned arrays:
//   field  : [ni][nj][nl][3 components][re,im]   float
//   padded : [2ni][2nj][2nl][3][re,im]            float
//   ahat   : [2ni][2nj][2nl][6 (e00 e01 e02 e11 e12 e22)][re,im] float
// and builds the reference's pointer-of-pointer views over them.

#include "function.hpp"
#include "fft_kernel.hpp"

#include <fcntl.h>
#include <unistd.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace {

template <class T>
struct view3 {
    std::vector<T*> rows;
    std::vector<T**> planes;
    T*** p = nullptr;
    view3(T* base, int ni, int nj, int nl) : rows((size_t)ni * nj), planes(ni) {
        for (int i = 0; i < ni; ++i) {
            planes[i] = rows.data() + (size_t)i * nj;
            for (int j = 0; j < nj; ++j)
                rows[(size_t)i * nj + j] = base + ((size_t)i * nj + j) * nl;
        }
        p = planes.data();
    }
};

struct tables {
    std::vector<cplx> w2i, w2j, w2l, wi, wj, wl;
    tables(int ni, int nj, int nl)
        : w2i(ni / 2 + 1), w2j(nj / 2 + 1), w2l(nl / 2 + 1), wi(ni), wj(nj), wl(nl) {
        // main.cpp:468-475
        FourierPhase(ni / 2, w2i.data());
        FourierPhase(nj / 2, w2j.data());
        FourierPhase(nl / 2, w2l.data());
        FourierPhase(ni, wi.data());
        FourierPhase(nj, wj.data());
        FourierPhase(nl, wl.data());
    }
};

struct stdout_to_file {
    int saved = -1;
    explicit stdout_to_file(const char* path) {
        if (!path || !*path) return;
        fflush(stdout);
        int fd = open(path, O_WRONLY | O_CREAT | O_TRUNC, 0644);
        if (fd < 0) return;
        saved = dup(1);
        dup2(fd, 1);
        close(fd);
    }
    ~stdout_to_file() {
        if (saved < 0) return;
        fflush(stdout);
        dup2(saved, 1);
        close(saved);
    }
};

}  // namespace

extern "C" {

// FourierPhase(n) table: n complex values (fft_func.cpp:77-88).
void ref_fourier_phase(int n, float* out) {
    std::vector<cplx> t(n > 2 ? n : 2);
    FourierPhase(n, t.data());
    memcpy(out, t.data(), sizeof(cplx) * n);
}

// One natural-order DFT line of the mixed-radix kernel (fft_kernel.hpp:287-322)
// on a line of cvectors; inverse = conj-DFT-conj without 1/N.
int ref_fft_line(int n, const float* in, float* out, int inverse) {
    const ddafft::plan* p = ddafft::get_plan(n);
    if (!p) return -1;
    if (inverse)
        ddafft::ifft_line(p, (cvector*)out, (const cvector*)in);
    else
        ddafft::fft_line(p, (cvector*)out, (const cvector*)in);
    return 0;
}

// Legacy radix-2 kernel FFT/iFFT on a line of cvectors (fft_func.cpp:188-311),
// using FourierPhase(n/2) twiddles as main.cpp:468-470 sets them up.
int ref_fft_line_legacy(int n, const float* in, float* out, int inverse) {
    std::vector<cplx> w(n / 2 + 1);
    std::vector<cvector> a((const cvector*)in, (const cvector*)in + n), b(n);
    int nb;
    FourierPhase(n / 2, w.data());
    Get_number_bits(n - 1, &nb);
    if (inverse)
        iFFT(nb, n, w.data(), a.data(), b.data());
    else
        FFT(nb, n, w.data(), a.data(), b.data());
    memcpy(out, b.data(), sizeof(cvector) * n);
    return 0;
}

// The Green tensor on the padded grid before any transform (dda_func.cpp:85-148).
void ref_amatrix(int ni, int nj, int nl, float kd, float* out) {
    view3<cmatrix> a((cmatrix*)out, 2 * ni, 2 * nj, 2 * nl);
    CreateAMatrix(ni, nj, nl, kd, a.p);
}

// Â = FFT_3D(A)/(8N) exactly as main.cpp:466-476 builds it.
void ref_build_ahat(int ni, int nj, int nl, float kd, float* out) {
    view3<cmatrix> a((cmatrix*)out, 2 * ni, 2 * nj, 2 * nl);
    tables t(ni, nj, nl);
    CreateAMatrix(ni, nj, nl, kd, a.p);
    FFT_3D(2 * ni, 2 * nj, 2 * nl, t.wi.data(), t.wj.data(), t.wl.data(), a.p);
}

// Pruned forward transform of a zero-padded field (fft_func.cpp:418-550).
// x: field [ni][nj][nl]; spec: padded [2ni][2nj][2nl] natural-order spectrum.
void ref_fft3d(int ni, int nj, int nl, const float* x, float* spec) {
    size_t P = (size_t)8 * ni * nj * nl;
    memset(spec, 0, P * sizeof(cvector));
    view3<cvector> w((cvector*)spec, 2 * ni, 2 * nj, 2 * nl);
    view3<cvector> f((cvector*)const_cast<float*>(x), ni, nj, nl);
    CopyingBack(ni, nj, nl, f.p, w.p);
    tables t(ni, nj, nl);
    FFT_3D(2 * ni, 2 * nj, 2 * nl, t.w2i.data(), t.w2j.data(), t.w2l.data(), t.wi.data(),
           t.wj.data(), t.wl.data(), w.p);
}

// Pruned inverse transform (fft_func.cpp:556-653); returns the first octant.
void ref_ifft3d(int ni, int nj, int nl, const float* spec, float* y) {
    size_t P = (size_t)8 * ni * nj * nl;
    std::vector<cvector> buf((const cvector*)spec, (const cvector*)spec + P);
    view3<cvector> w(buf.data(), 2 * ni, 2 * nj, 2 * nl);
    tables t(ni, nj, nl);
    iFFT_3D(2 * ni, 2 * nj, 2 * nl, t.wi.data(), t.wj.data(), t.wl.data(), w.p);
    cvector* o = (cvector*)y;
    for (int i = 0; i < ni; ++i)
        for (int j = 0; j < nj; ++j)
            for (int l = 0; l < nl; ++l) o[((size_t)i * nj + j) * nl + l] = w.p[i][j][l];
}

// y = first octant of IFFT(Â ⊙ FFT(pad x)) (dda_func.cpp:150-187), or with
// conj(Â(-k)) when conj != 0 (dda_func.cpp:227-437).
void ref_convolution(int ni, int nj, int nl, const float* ahat, const float* x, float* y,
                     int conj) {
    size_t P = (size_t)8 * ni * nj * nl;
    std::vector<cvector> buf(P);
    view3<cvector> w(buf.data(), 2 * ni, 2 * nj, 2 * nl);
    view3<cvector> f((cvector*)const_cast<float*>(x), ni, nj, nl);
    view3<cmatrix> a((cmatrix*)const_cast<float*>(ahat), 2 * ni, 2 * nj, 2 * nl);
    tables t(ni, nj, nl);
    CopyingBack(ni, nj, nl, f.p, w.p);
    if (conj)
        Convolution_conj(2 * ni, 2 * nj, 2 * nl, t.w2i.data(), t.w2j.data(), t.w2l.data(),
                         t.wi.data(), t.wj.data(), t.wl.data(), a.p, w.p);
    else
        Convolution(2 * ni, 2 * nj, 2 * nl, t.w2i.data(), t.w2j.data(), t.w2l.data(),
                    t.wi.data(), t.wj.data(), t.wl.data(), a.p, w.p);
    cvector* o = (cvector*)y;
    for (int i = 0; i < ni; ++i)
        for (int j = 0; j < nj; ++j)
            for (int l = 0; l < nl; ++l) o[((size_t)i * nj + j) * nl + l] = w.p[i][j][l];
}

// P0 = α e^{ik(r-c)} E on occupied cells (dda_func.cpp:654-689).
void ref_first_approx(int ni, int nj, int nl, const int* mv, const float* alpha,
                      const signed char* occ, float kd, const float* kinc, const float* einc,
                      float* p0) {
    view3<char> o((char*)const_cast<signed char*>(occ), ni, nj, nl);
    view3<cvector> p((cvector*)p0, ni, nj, nl);
    vector k = {kinc[0], kinc[1], kinc[2]};
    cvector e;
    memcpy(&e, einc, sizeof(e));
    FirstApproximation(ni, nj, nl, const_cast<int*>(mv), (cplx*)const_cast<float*>(alpha), o.p,
                       kd, k, e, p.p);
}

// The full DDA_FFT solve (dda_func.cpp:1075-1570) with the Cooley-Tukey tables.
// dirs: [nplanes][nangles][3]; esc out: [nplanes][nangles][3][2].  The
// per-step discrepancy lines are written to log_path when it is given.
int ref_dda_fft(int solver, float tau, int ni, int nj, int nl, const float* alpha,
                const int* mv, const signed char* occ, int max_step, float err_max, float kd,
                const float* kinc, const float* einc, const float* ahat, int nplanes,
                int nangles, const float* dirs, float* esc, float* cabs, float* cext,
                const char* log_path) {
    view3<char> o((char*)const_cast<signed char*>(occ), ni, nj, nl);
    view3<cmatrix> a((cmatrix*)const_cast<float*>(ahat), 2 * ni, 2 * nj, 2 * nl);
    tables t(ni, nj, nl);
    std::vector<vector*> ksc(nplanes);
    std::vector<cvector*> es(nplanes);
    for (int m = 0; m < nplanes; ++m) {
        ksc[m] = (vector*)const_cast<float*>(dirs) + (size_t)m * nangles;
        es[m] = (cvector*)esc + (size_t)m * nangles;
    }
    vector k = {kinc[0], kinc[1], kinc[2]};
    cvector e;
    memcpy(&e, einc, sizeof(e));
    dda_workspace ws;
    dda_workspace_create(&ws, ni, nj, nl, solver);
    ws.iVerbose = 1;
    int rc;
    {
        stdout_to_file redirect(log_path);
        rc = DDA_FFT(0, solver, tau, ni, nj, nl, (cplx*)const_cast<float*>(alpha),
                     const_cast<int*>(mv), o.p, max_step, err_max, kd, k, e, t.w2i.data(),
                     t.w2j.data(), t.w2l.data(), t.wi.data(), t.wj.data(), t.wl.data(), nullptr,
                     nullptr, nullptr, nullptr, nullptr, nullptr, 0, a.p, nplanes, nangles,
                     ksc.data(), es.data(), cabs, cext, &ws);
    }
    dda_workspace_free(&ws);
    return rc;
}

void ref_cabs_cext(float kd, const float* kinc, const float* einc, int ni, int nj, int nl,
                   const float* alpha, const signed char* occ, const float* p, float* cabs,
                   float* cext) {
    view3<char> o((char*)const_cast<signed char*>(occ), ni, nj, nl);
    view3<cvector> pv((cvector*)const_cast<float*>(p), ni, nj, nl);
    vector k = {kinc[0], kinc[1], kinc[2]};
    cvector e;
    memcpy(&e, einc, sizeof(e));
    CabsCext(kd, k, e, ni, nj, nl, (cplx*)const_cast<float*>(alpha), o.p, pv.p, cabs, cext);
}

void ref_far_field(int ni, int nj, int nl, const int* mv, const signed char* occ, float kd,
                   const float* p, const float* dir, float* esc) {
    view3<char> o((char*)const_cast<signed char*>(occ), ni, nj, nl);
    view3<cvector> pv((cvector*)const_cast<float*>(p), ni, nj, nl);
    vector d = {dir[0], dir[1], dir[2]};
    cvector r = FarField(ni, nj, nl, const_cast<int*>(mv), o.p, kd, pv.p, d);
    memcpy(esc, &r, sizeof(r));
}

// Polarizability models 1-5 (dda_func.cpp:1631-1677, selection main.cpp:606-631, 926-931).
void ref_polarizability(int model, float kd, const float* m, const float* kinc,
                        const float* einc, float* alpha) {
    cplx n = {m[0], m[1]};
    cplx a = Polarizability_CM(kd, n);
    if (model == 2) a = Polarizability_CMRR(a);
    if (model == 3) a = Polarizability_CMNR(kd, a);
    if (model == 4) a = Polarizability_ILDR(kd, n, a);
    if (model == 5) {
        vector k = {kinc[0], kinc[1], kinc[2]};
        cvector e;
        memcpy(&e, einc, sizeof(e));
        a = Polarizability_LDR(kd, n, a, k, e);
    }
    alpha[0] = a.r;
    alpha[1] = a.i;
}

// One Jones -> Mueller conversion (mmuller.cpp:175-314).
void ref_muller(const float* kinc, const float* ksc, const float* normal, const float* einc0,
                const float* einc1, const float* esc0, const float* esc1, float* mm) {
    vector k = {kinc[0], kinc[1], kinc[2]}, s = {ksc[0], ksc[1], ksc[2]},
           n = {normal[0], normal[1], normal[2]};
    cvector e0, e1, s0, s1;
    memcpy(&e0, einc0, sizeof(e0));
    memcpy(&e1, einc1, sizeof(e1));
    memcpy(&s0, esc0, sizeof(s0));
    memcpy(&s1, esc1, sizeof(s1));
    mmuller r = JonesToMuller(BuildMatrixJones(k, s, n, e0, e1, s0, s1));
    memcpy(mm, &r, sizeof(r));
}

// Deterministic ellipsoid / parallelepiped / bisphere targets (particle.cpp:408-527).
int ref_target(int kind, int ni, int nj, int nl, float p1, float p2, signed char* occ) {
    view3<char> o((char*)occ, ni, nj, nl);
    if (kind == 2) return iGenerateEllipsoid(ni, nj, nl, p1, p2, o.p);
    if (kind == 3) return iGenerateParallelepiped(ni, nj, nl, p1, p2, o.p);
    if (kind == 12) return iGenerateBisphere(ni, nj, nl, p1, o.p);
    return -1;
}

// Random and parametric targets (particle.cpp:21-995) after dda_srand(*state);
// *state is left where the reference's stream stands afterwards, as the
// next dda_rand() would continue from it.  kind: 5 random ellipsoid,
// 9 random parallelepiped, 10 nibble, 13 hierarchical, 14 sphere spec,
// 6/7/8 worms.
int ref_target_rand(int kind, int ni, int nj, int nl, float b_a, float c_a, int nsorts, int* nseeds, int bites,
                    float depth, int* mv, int levels, float rho, float radius_worm, float length_worm, int nworms,
                    unsigned long long seed, signed char* occ) {
    view3<char> o((char*)occ, ni, nj, nl);
    dda_srand(seed);
    switch (kind) {
        case 5: return iGenerateRandomEllipsoid(ni, nj, nl, b_a, c_a, nsorts, nseeds, o.p);
        case 9: return iGenerateRandomParallelepiped(ni, nj, nl, b_a, c_a, nsorts, nseeds, o.p);
        case 10: return iGenerateNibbleEllipsoid(ni, nj, nl, b_a, c_a, nsorts, nseeds, bites, depth, mv, o.p);
        case 13: return iGenerateHierarchicalEllipsoid(levels, ni, nj, nl, b_a, c_a, nsorts, nseeds, o.p);
        case 14: return iGenerateSphereSpec(ni, nj, nl, ni / 2., rho, o.p);
        case 6:
        case 7:
        case 8: {
            std::vector<cylinder> w((size_t)nworms);
            if (kind == 6) return iGenerateAnthrax(ni, nj, nl, radius_worm, length_worm, nworms, w.data(), o.p);
            if (kind == 7) return iGenerateHedgehog(ni, nj, nl, radius_worm, length_worm, nworms, w.data(), o.p);
            return iGenerateDeadAnthrax(ni, nj, nl, radius_worm, length_worm, nworms, w.data(), o.p);
        }
    }
    return -100;
}

// The reference's next draw from its current stream (after ref_target_rand).
int ref_rand_next(void) { return dda_rand(); }

// The next value of the reference's stream after dda_srand(seed) and `skip` draws.
int ref_rand_after(unsigned long long seed, int skip) {
    dda_srand(seed);
    for (int k = 0; k < skip; ++k) dda_rand();
    return dda_rand();
}

}  // extern "C"
