// Host-side pieces of the solve path that run once per sample or per solve:
// target occupancy, polarizabilities, scattering geometry and the Jones ->
// Mueller conversion.  They keep the reference's float semantics (operation
// order and float/double promotions), so occupancy maps are bit-exact and
// alpha, directions and Mueller elements round as the reference's do.
// Compiled with -ffp-contract=off.
#include <cmath>
#include <complex>
#include <cstring>

#include "host_internal.hpp"

#include "../../include/ddafft.h"

namespace {

const double kPi = 3.1415926535897932384626433832795;

struct cf {
    float r, i;
};
inline cf cmul(cf a, cf b) { return {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; }
inline cf cadd(cf a, cf b) { return {a.r + b.r, a.i + b.i}; }
inline cf csub(cf a, cf b) { return {a.r - b.r, a.i - b.i}; }
inline cf rmul(float r, cf a) { return {r * a.r, r * a.i}; }
inline cf cconj(cf a) { return {a.r, -a.i}; }
inline cf cdiv(cf a, cf b) {   // complex.cpp:81-90
    const float n = b.r * b.r + b.i * b.i;
    return {(a.r * b.r + a.i * b.i) / n, (a.i * b.r - a.r * b.i) / n};
}
inline cf cdiv_macro(cf x, cf y) {   // cplx.h Cdiv
    const float n = y.r * y.r + y.i * y.i;
    return {(x.r * y.r + x.i * y.i) / n, (-x.r * y.i + x.i * y.r) / n};
}

struct v3 {
    float x, y, z;
};
inline v3 cross(v3 a, v3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }

// Rotations about x, y, z (dda_func.cpp:7-34)
inline v3 rot_x(float c, float s, v3 v) { return {v.x, c * v.y + s * v.z, -s * v.y + c * v.z}; }
inline v3 rot_y(float c, float s, v3 v) { return {c * v.x - s * v.z, v.y, s * v.x + c * v.z}; }
inline v3 rot_z(float c, float s, v3 v) { return {c * v.x + s * v.y, -s * v.x + c * v.y, v.z}; }

// alpha_CM (dda_func.cpp:1631-1637)
cf pol_cm(float d0, cf m) {
    const cf mm = cmul(m, m);
    const float f = (float)((double)(d0 * d0 * d0) * (0.75 / kPi));
    const cf num = {mm.r - 1.f, mm.i}, den = {2.f + mm.r, mm.i};
    return cdiv(rmul(f, num), den);
}

// a / (1 - (2/3) i a)  (dda_func.cpp:1639-1646)
cf radiative(cf a) {
    const cf i23 = rmul((float)(2. / 3.), cf{0.f, 1.f});
    const cf t = cmul(i23, a);
    return cdiv(a, cf{1.f - t.r, -t.i});
}

// lattice dispersion relation core (dda_func.cpp:1659-1677)
cf ldr_core(float d0, cf m, cf a, float S) {
    const float b1 = -1.8915316f, b2 = 0.1648469f, b3 = -1.7700004f;
    const cf eps = cmul(m, m);
    cf t = rmul(b2, eps);
    t.r = b1 + t.r;
    t = cadd(t, rmul(S * b3, eps));
    t = csub(t, rmul((float)((2. / 3.) * d0), cf{0.f, 1.f}));
    const cf ad = {a.r / d0, a.i / d0};
    cf u = cmul(ad, t);
    u.r = 1.f + u.r;
    return cdiv(a, u);
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------------------
// Targets (particle.cpp:408-527): returns the number of occupied cells, or
// -1 when the body does not fit (the reference prints a message and returns
// -1 as well).
// ---------------------------------------------------------------------------

int ddafft_target_ellipsoid(int ni, int nj, int nl, double b_a_in, double c_a_in, signed char* occ) {
    if (!occ || ni < 1 || nj < 1 || nl < 1) return -1;
    const float b_a = (float)b_a_in, c_a = (float)c_a_in;
    const float cx = (float)((float)ni / 2. - .5), cy = (float)((float)nj / 2. - .5), cz = (float)((float)nl / 2. - .5);
    float a = (float)ni, b = a * b_a, c = a * c_a;
    if (b > (float)nj || c > (float)nl) return -1;
    a = (float)(a / 2.);
    b = (float)(b / 2.);
    c = (float)(c / 2.);
    a = a * a;
    b = b * b;
    c = c * c;
    int cnt = 0;
    for (int i = 0; i < ni; ++i)
        for (int j = 0; j < nj; ++j)
            for (int l = 0; l < nl; ++l) {
                const float x = (float)i - cx, y = (float)j - cy, z = (float)l - cz;
                const bool in = (x * x) / a + (y * y) / b + (z * z) / c <= 1.f;
                occ[((size_t)i * nj + j) * nl + l] = in ? 1 : 0;
                cnt += in;
            }
    return cnt;
}

int ddafft_target_parallelepiped(int ni, int nj, int nl, double b_a_in, double c_a_in, signed char* occ) {
    if (!occ || ni < 1 || nj < 1 || nl < 1) return -1;
    const float a = (float)ni, b = a * (float)b_a_in, c = a * (float)c_a_in;
    if (b > (float)nj || c > (float)nl) return -1;
    int cnt = 0;
    for (int i = 0; i < ni; ++i)
        for (int j = 0; j < nj; ++j)
            for (int l = 0; l < nl; ++l) {
                const bool in = (double)(a - (float)i) >= 0.001 && (double)(b - (float)j) >= 0.001 &&
                                (double)(c - (float)l) >= 0.001;
                occ[((size_t)i * nj + j) * nl + l] = in ? 1 : 0;
                cnt += in;
            }
    return cnt;
}

int ddafft_target_bisphere(int ni, int nj, int nl, double r1_r2_in, signed char* occ) {
    if (!occ || ni < 1 || nj < 1 || nl < 1) return -1;
    const float r1_r2 = (float)r1_r2_in;
    float R1, R2, c1x, c2x;
    if (r1_r2 > 1.f) {
        R1 = (float)((float)nj / 2.);
        R2 = R1 / r1_r2;
        c1x = (float)((float)ni / 4. - .5);
        c2x = (R1 + R2) + c1x;
    } else {
        R2 = (float)((float)nj / 2.);
        R1 = R2 * r1_r2;
        c2x = (float)((float)ni * 3. / 4. - .5);
        c1x = c2x - (R1 + R2);
    }
    const float cy = (float)((float)nj / 2. - .5), cz = (float)((float)nl / 2. - .5);
    const float R1_2 = R1 * R1, R2_2 = R2 * R2;
    int cnt = 0;
    for (int i = 0; i < ni; ++i)
        for (int j = 0; j < nj; ++j)
            for (int l = 0; l < nl; ++l) {
                const float x1 = (float)i - c1x, y = (float)j - cy, z = (float)l - cz, x2 = (float)i - c2x;
                const float d1 = (x1 * x1) + (y * y) + (z * z), d2 = (x2 * x2) + (y * y) + (z * z);
                signed char v = 0;
                if (!(d1 > R1_2 && d2 > R2_2)) {
                    ++cnt;
                    if (d1 <= R1_2) v = 1;
                    if (d2 <= R2_2) v = 2;
                }
                occ[((size_t)i * nj + j) * nl + l] = v;
            }
    return cnt;
}

// ---------------------------------------------------------------------------
// Polarizability models 1..5 (dda_func.cpp:1631-1677, selection
// main.cpp:606-631 and 926-931).  kinc/einc are used by model 5 only.
// ---------------------------------------------------------------------------

int ddafft_polarizability(int model, float kd, const float* m, const float* kinc, const float* einc, float* alpha) {
    if (!m || !alpha || model < 1 || model > 5) return DDAFFT_EINVAL;
    const cf n = {m[0], m[1]};
    cf a = pol_cm(kd, n);
    if (model == 2) a = radiative(a);
    if (model == 3) {
        const float b = (float)(-std::sqrt(4. * kPi / 3.) / kd);
        cf t = rmul(b, a);
        t.r = 1.f + t.r;
        a = radiative(cdiv(a, t));
    }
    if (model == 4) a = ldr_core(kd, n, a, .2f);
    if (model == 5) {
        if (!kinc || !einc) return DDAFFT_EINVAL;
        const float rx = kinc[0] * einc[0], ry = kinc[1] * einc[2], rz = kinc[2] * einc[4];
        a = ldr_core(kd, n, a, rx * rx + ry * ry + rz * rz);
    }
    alpha[0] = a.r;
    alpha[1] = a.i;
    return DDAFFT_OK;
}

// Holomorphic derivative d alpha / d m of the five models, in double
// (the formulas of dda_func.cpp:1631-1677 differentiated by hand).
int ddafft_polarizability_deriv(int model, float kd, const float* m, const float* kinc, const float* einc,
                                float* dalpha_dm) {
    if (!m || !dalpha_dm || model < 1 || model > 5) return DDAFFT_EINVAL;
    if (model == 5 && (!kinc || !einc)) return DDAFFT_EINVAL;
    using C = std::complex<double>;
    const double d = kd;
    const C n(m[0], m[1]), e = n * n;
    const double f = d * d * d * (0.75 / kPi);
    C a = f * (e - 1.) / (e + 2.), da = f * 6. * n / ((e + 2.) * (e + 2.));
    if (model == 3) {
        const double b = -std::sqrt(4. * kPi / 3.) / d;
        const C u = 1. + b * a;
        da = da / (u * u);
        a = a / u;
    }
    if (model == 2 || model == 3) {
        const C u = 1. - C(0., 2. / 3.) * a;
        da = da / (u * u);
        a = a / u;
    }
    if (model == 4 || model == 5) {
        double S = .2;
        if (model == 5) {
            const double rx = (double)kinc[0] * einc[0], ry = (double)kinc[1] * einc[2],
                         rz = (double)kinc[2] * einc[4];
            S = rx * rx + ry * ry + rz * rz;
        }
        const double b1 = -1.8915316, b2 = 0.1648469, b3 = -1.7700004;
        const C t = b1 + b2 * e + S * b3 * e - C(0., (2. / 3.) * d);
        const C dt = (b2 + S * b3) * 2. * n;
        const C u = 1. + (a / d) * t;
        const C du = (da / d) * t + (a / d) * dt;
        da = (da * u - a * du) / (u * u);
    }
    dalpha_dm[0] = (float)da.real();
    dalpha_dm[1] = (float)da.imag();
    return DDAFFT_OK;
}

// ---------------------------------------------------------------------------
// Scattering geometry for a fixed orientation (main.cpp:354-387, 516-542):
// planes rotated about z by m*180/nplanes, phase angles theta_min + m*step,
// then the -psi(z), -iota(y), -kappa(x) rotations (angles in degrees).
//   dirs   [nplanes][nangles][3]   normals [nplanes][3]
//   kinc   [3]                     einc [2][3][2] (the two polarizations)
// ---------------------------------------------------------------------------

int ddafft_geometry_fixed(int nplanes, int nangles, double theta_min, double theta_step, double psi_deg,
                          double iota_deg, double kappa_deg, float* dirs, float* normals, float* kinc, float* einc) {
    if (nplanes < 1 || nangles < 1 || !dirs || !normals || !kinc || !einc) return DDAFFT_EINVAL;
    const float tmin = (float)theta_min, tstep = (float)theta_step;
    // first plane (SphCart_d with r = 1, phi = 0)
    for (int m = 0; m < nangles; ++m) {
        const float th = tmin + tstep * (float)m;
        v3 d = {(float)(1. * std::cos(0. * kPi / 180.) * std::sin(th * kPi / 180.)),
                (float)(1. * std::sin(0. * kPi / 180.) * std::sin(th * kPi / 180.)), (float)(1. * std::cos(th * kPi / 180.))};
        std::memcpy(&dirs[3 * m], &d, sizeof(d));
    }
    v3 n0 = {0.f, 1.f, 0.f};
    std::memcpy(&normals[0], &n0, sizeof(n0));
    const float r0 = (float)(180. / (double)(float)nplanes);
    float r1 = 0.f;
    for (int m = 1; m < nplanes; ++m) {
        r1 += r0;
        const float c = (float)std::cos(r1 * (kPi / 180.)), s = (float)std::sin(r1 * (kPi / 180.));
        const v3 nm = rot_z(c, s, n0);
        std::memcpy(&normals[3 * m], &nm, sizeof(nm));
        for (int i = 0; i < nangles; ++i) {
            v3 d;
            std::memcpy(&d, &dirs[3 * i], sizeof(d));
            d = rot_z(c, s, d);
            std::memcpy(&dirs[3 * ((size_t)m * nangles + i)], &d, sizeof(d));
        }
    }
    const float psi = (float)((float)psi_deg * kPi / 180.), iota = (float)((float)iota_deg * kPi / 180.),
                kap = (float)((float)kappa_deg * kPi / 180.);
    const float cp = std::cos(-psi), sp = std::sin(-psi), ci = std::cos(-iota), si = std::sin(-iota),
                ck = std::cos(-kap), sk = std::sin(-kap);
    auto orient = [&](v3 v) { return rot_x(ck, sk, rot_y(ci, si, rot_z(cp, sp, v))); };
    for (int m = 0; m < nplanes; ++m) {
        v3 nm;
        std::memcpy(&nm, &normals[3 * m], sizeof(nm));
        nm = orient(nm);
        std::memcpy(&normals[3 * m], &nm, sizeof(nm));
        for (int i = 0; i < nangles; ++i) {
            v3 d;
            std::memcpy(&d, &dirs[3 * ((size_t)m * nangles + i)], sizeof(d));
            d = orient(d);
            std::memcpy(&dirs[3 * ((size_t)m * nangles + i)], &d, sizeof(d));
        }
    }
    const v3 k = orient(v3{0.f, 0.f, 1.f});
    std::memcpy(kinc, &k, sizeof(k));
    // Einc rotated as cvectors (real parts only; imaginary parts stay 0)
    for (int p = 0; p < 2; ++p) {
        cf e[3] = {{p == 0 ? 1.f : 0.f, 0.f}, {p == 1 ? 1.f : 0.f, 0.f}, {0.f, 0.f}};
        auto rc = [](float c, float s, int axis, cf* v) {
            cf x = v[0], y = v[1], z = v[2];
            if (axis == 2) {
                v[0] = cadd(rmul(c, x), rmul(s, y));
                v[1] = cadd(rmul(-s, x), rmul(c, y));
            } else if (axis == 1) {
                v[0] = csub(rmul(c, x), rmul(s, z));
                v[2] = cadd(rmul(s, x), rmul(c, z));
            } else {
                v[1] = cadd(rmul(c, y), rmul(s, z));
                v[2] = cadd(rmul(-s, y), rmul(c, z));
            }
        };
        rc(cp, sp, 2, e);
        rc(ci, si, 1, e);
        rc(ck, sk, 0, e);
        for (int q = 0; q < 3; ++q) {
            einc[6 * p + 2 * q] = e[q].r;
            einc[6 * p + 2 * q + 1] = e[q].i;
        }
    }
    return DDAFFT_OK;
}

// ---------------------------------------------------------------------------
// Jones matrix of one direction and its Mueller matrix (mmuller.cpp:175-314).
// mm: [4][4] row-major.
// ---------------------------------------------------------------------------

int ddafft_muller(const float* kinc, const float* ksc, const float* normal, const float* einc0, const float* einc1,
                  const float* esc0, const float* esc1, float* mm) {
    if (!kinc || !ksc || !normal || !einc0 || !einc1 || !esc0 || !esc1 || !mm) return DDAFFT_EINVAL;
    const v3 k = {kinc[0], kinc[1], kinc[2]}, s = {ksc[0], ksc[1], ksc[2]}, nv = {normal[0], normal[1], normal[2]};
    v3 v0 = cross(k, nv);
    const float a0 = std::sqrt(v0.x * v0.x + v0.y * v0.y + v0.z * v0.z);
    const float f0 = (float)(1. / a0);
    const v3 pin = {f0 * v0.x, f0 * v0.y, f0 * v0.z};
    v0 = cross(s, nv);
    const float a1 = std::sqrt(v0.x * v0.x + v0.y * v0.y + v0.z * v0.z);
    const float f1 = (float)(-1. / a1);
    const v3 psc = {f1 * v0.x, f1 * v0.y, f1 * v0.z};
    auto proj = [](v3 v, const float* e) {
        const cf c0 = rmul(v.x, cf{e[0], e[1]}), c1 = rmul(v.y, cf{e[2], e[3]}), c2 = rmul(v.z, cf{e[4], e[5]});
        return cadd(c2, cadd(c0, c1));
    };
    const cf cc0 = proj(nv, einc0), cc1 = proj(nv, einc1), cc2 = proj(pin, einc0), cc3 = proj(pin, einc1);
    const cf c3 = csub(cmul(cc0, cc3), cmul(cc1, cc2));
    const cf den = {-c3.i, c3.r};
    const cf cc4 = proj(nv, esc0), cc5 = proj(nv, esc1), cc6 = proj(psc, esc0), cc7 = proj(psc, esc1);
    const cf J00 = cdiv_macro(csub(cmul(cc7, cc0), cmul(cc6, cc1)), den);
    const cf J01 = cdiv_macro(csub(cmul(cc6, cc3), cmul(cc7, cc2)), den);
    const cf J10 = cdiv_macro(csub(cmul(cc5, cc0), cmul(cc4, cc1)), den);
    const cf J11 = cdiv_macro(csub(cmul(cc4, cc3), cmul(cc5, cc2)), den);
    const cf a = cmul(J00, cconj(J00)), b = cmul(J01, cconj(J01)), c = cmul(J10, cconj(J10)), d = cmul(J11, cconj(J11));
    const cf x01 = cmul(J00, cconj(J01)), x1110 = cmul(J11, cconj(J10));
    const cf x0010 = cmul(J00, cconj(J10)), x1101 = cmul(J11, cconj(J01));
    const cf x1100 = cmul(J11, cconj(J00)), x0110 = cmul(J01, cconj(J10));
    const cf x0011 = cmul(J00, cconj(J11)), x1001 = cmul(J10, cconj(J01));
    const cf x1000 = cmul(J10, cconj(J00));
    float(*M)[4] = reinterpret_cast<float(*)[4]>(mm);
    M[0][0] = (float)(.5 * cadd(cadd(cadd(a, b), c), d).r);
    M[0][1] = (float)(.5 * csub(cadd(csub(a, b), c), d).r);
    M[0][2] = cadd(x01, x1110).r;
    M[0][3] = csub(x01, x1110).i;
    M[1][0] = (float)(.5 * csub(csub(cadd(a, b), c), d).r);
    M[1][1] = (float)(.5 * cadd(csub(csub(a, b), c), d).r);
    M[1][2] = csub(x01, x1110).r;
    M[1][3] = cadd(x01, x1110).i;
    M[2][0] = cadd(x0010, x1101).r;
    M[2][1] = csub(x0010, x1101).r;
    M[2][2] = cadd(x1100, x0110).r;
    M[2][3] = cadd(x0011, x1001).i;
    M[3][0] = cadd(x1000, x1101).i;
    M[3][1] = csub(x1000, x1101).i;
    M[3][2] = csub(x1100, x0110).i;
    M[3][3] = csub(x1100, x0110).r;
    return DDAFFT_OK;
}

// Plane-averaged Mueller matrices of one sample (main.cpp:943-954):
// mm_out[nangles][4][4] = (1/nplanes) sum_planes Muller(plane, angle).
// esc0/esc1: [nplanes][nangles][3][2]; dirs [nplanes][nangles][3].
int ddafft_muller_average(int nplanes, int nangles, const float* kinc, const float* dirs, const float* normals,
                          const float* einc, const float* esc0, const float* esc1, float* mm_out) {
    if (nplanes < 1 || nangles < 1 || !kinc || !dirs || !normals || !einc || !esc0 || !esc1 || !mm_out)
        return DDAFFT_EINVAL;
    for (int m = 0; m < nangles; ++m) {
        float acc[16] = {0};
        for (int p = 0; p < nplanes; ++p) {
            const size_t d = (size_t)p * nangles + m;
            float one[16];
            ddafft_muller(kinc, dirs + 3 * d, normals + 3 * p, einc, einc + 6, esc0 + 6 * d, esc1 + 6 * d, one);
            for (int q = 0; q < 16; ++q) acc[q] = acc[q] + one[q];
        }
        const float np = (float)nplanes;
        for (int q = 0; q < 16; ++q) mm_out[16 * m + q] = acc[q] / np;
    }
    return DDAFFT_OK;
}

}  // extern "C"

namespace ddafft_host {

// M00 of one direction (mmuller.cpp:175-314, evaluated in double) and its
// Wirtinger derivatives g0 = dM00/dEsc_0, g1 = dM00/dEsc_1 (3 complex each).
// M00 = (|J00|^2 + |J01|^2 + |J10|^2 + |J11|^2) / 2 with every J linear in the
// projections of Esc_0 and Esc_1 on the plane normal and on psc.
double m00_grad(const float* kinc, const float* ksc, const float* normal, const float* einc0, const float* einc1,
                const float* esc0, const float* esc1, std::complex<double>* g0, std::complex<double>* g1) {
    using C = std::complex<double>;
    const double k[3] = {kinc[0], kinc[1], kinc[2]}, s[3] = {ksc[0], ksc[1], ksc[2]},
                 nv[3] = {normal[0], normal[1], normal[2]};
    auto cr = [](const double* a, const double* b, double* o) {
        o[0] = a[1] * b[2] - a[2] * b[1];
        o[1] = a[2] * b[0] - a[0] * b[2];
        o[2] = a[0] * b[1] - a[1] * b[0];
    };
    double pin[3], psc[3];
    cr(k, nv, pin);
    cr(s, nv, psc);
    const double n0 = std::sqrt(pin[0] * pin[0] + pin[1] * pin[1] + pin[2] * pin[2]);
    const double n1 = std::sqrt(psc[0] * psc[0] + psc[1] * psc[1] + psc[2] * psc[2]);
    for (int q = 0; q < 3; ++q) {
        pin[q] /= n0;
        psc[q] = -psc[q] / n1;
    }
    auto dot = [](const double* v, const float* e) {
        return v[0] * C(e[0], e[1]) + v[1] * C(e[2], e[3]) + v[2] * C(e[4], e[5]);
    };
    const C cc0 = dot(nv, einc0), cc1 = dot(nv, einc1), cc2 = dot(pin, einc0), cc3 = dot(pin, einc1);
    const C den = C(0., 1.) * (cc0 * cc3 - cc1 * cc2);
    const C cc4 = dot(nv, esc0), cc5 = dot(nv, esc1), cc6 = dot(psc, esc0), cc7 = dot(psc, esc1);
    const C J00 = (cc7 * cc0 - cc6 * cc1) / den, J01 = (cc6 * cc3 - cc7 * cc2) / den;
    const C J10 = (cc5 * cc0 - cc4 * cc1) / den, J11 = (cc4 * cc3 - cc5 * cc2) / den;
    const double m00 = .5 * (std::norm(J00) + std::norm(J01) + std::norm(J10) + std::norm(J11));
    const C a0 = .5 * (std::conj(J00) * -cc1 + std::conj(J01) * cc3) / den;   // psc part of dM00/dEsc_0
    const C b0 = .5 * (std::conj(J10) * -cc1 + std::conj(J11) * cc3) / den;   // normal part
    const C a1 = .5 * (std::conj(J00) * cc0 - std::conj(J01) * cc2) / den;
    const C b1 = .5 * (std::conj(J10) * cc0 - std::conj(J11) * cc2) / den;
    for (int q = 0; q < 3; ++q) {
        g0[q] = a0 * psc[q] + b0 * nv[q];
        g1[q] = a1 * psc[q] + b1 * nv[q];
    }
    return m00;
}

}  // namespace ddafft_host
