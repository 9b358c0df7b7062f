// Random and parametric targets, the reference's random stream, and the
// Monte-Carlo orientation, on the host.
//
// Every expression keeps the reference's types: `real` is float, literals such
// as 2. and PI are double, so a product like 2.*PI*(real)r/(real)RAND_MAX is
// formed in double and rounded once to float, while (real)r/(real)RAND_MAX is a
// float division; float arguments go to the float libm functions (acosf, sinf,
// cosf, powf, sqrtf) as C++ overload resolution sends them in the reference.
// With the same seed the occupancy maps are therefore identical to the
// reference's (particle.cpp:21-995, randomize.cpp:44-58, main.cpp:875-922).
// Compiled with -ffp-contract=off.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/ddafft.h"

namespace {

const double kPi = 3.1415926535897932384626433832795;
const double kZero = 0.0000000001;   // function.hpp:35
const float kRandMaxF = (float)2147483647;

// xorshift64* (randomize.cpp:44-58): 31-bit outputs like rand()
struct Stream {
    unsigned long long* s;
    int next() {
        unsigned long long t = *s;
        t ^= t >> 12;
        t ^= t << 25;
        t ^= t >> 27;
        *s = t;
        return (int)((t * 2685821657736338717ULL) >> 33);
    }
    float unit() { return (float)next() / kRandMaxF; }   // (real)dda_rand()/(real)DDA_RAND_MAX
};

struct V {
    float x, y, z;
};
inline V operator+(V a, V b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V operator-(V a, V b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline float dot(V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V scale(float r, V a) { return {r * a.x, r * a.y, r * a.z}; }
inline V cross(V a, V b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline float norm(V a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }

// rounding() (complex.cpp:213-223): fractional part above one half rounds up
inline int round_half_up(float r0) {
    double ip;
    const double fr = std::modf((double)r0, &ip);
    return fr > 0.5 ? (int)ip + 1 : (int)ip;
}

inline V centre(int ni, int nj, int nl) {
    return {(float)((float)ni / 2. - .5), (float)((float)nj / 2. - .5), (float)((float)nl / 2. - .5)};
}

inline size_t at(int i, int j, int l, int nj, int nl) { return ((size_t)i * nj + j) * nl + l; }

// a lattice point drawn as in the seed loops (particle.cpp:557-559)
inline V draw_point(Stream& rng, int ni, int nj, int nl) {
    V v;
    v.x = (float)round_half_up((float)(rng.unit() * ((float)ni - 1.)));
    v.y = (float)round_half_up((float)(rng.unit() * ((float)nj - 1.)));
    v.z = (float)round_half_up((float)(rng.unit() * ((float)nl - 1.)));
    return v;
}

// index of the nearest seed (first one on ties), particle.cpp:578-585
inline int nearest(V p, const std::vector<V>& seeds, size_t count, int ni, int nj, int nl) {
    int best = 0;
    float dmin = std::sqrt((float)(ni * ni + nj * nj + nl * nl));
    for (size_t m = 0; m < count; ++m) {
        const float dd = norm(p - seeds[m]);
        if (dd < dmin) {
            best = (int)m;
            dmin = dd;
        }
    }
    return best;
}

// material of the seed with index `flag` (seed ranges per material in order)
inline int material_of(int flag, int nsorts, const int* nseeds) {
    int first = nseeds[0];
    if (flag < first) return 1;
    for (int m = 1; m < nsorts; ++m) {
        if (flag >= first && flag < first + nseeds[m]) return m + 1;
        first += nseeds[m];
    }
    return -1;   // no material: the reference leaves the cell as it was
}

inline bool duplicate(const std::vector<V>& seeds, size_t from, size_t to, V v) {
    for (size_t i = from; i < to; ++i)
        if (norm(seeds[i] - v) < kZero) return true;
    return false;
}

struct Ell {
    float a, b, c;   // squared semi-axes
    bool inside(V v) const { return (v.x * v.x) / a + (v.y * v.y) / b + (v.z * v.z) / c <= 1.; }
};

Ell semi_axes2(float a, float b, float c) {
    a = (float)(a / 2.);
    b = (float)(b / 2.);
    c = (float)(c / 2.);
    return {a * a, b * b, c * c};
}

// ---------------------------------------------------------------------------
// worms (particle.cpp:233-405)
// ---------------------------------------------------------------------------

struct Cyl {
    float r, l;
    V b, e, n;
};

int point_to_worm(const Cyl& w, V p, float radius) {
    const float t = dot(p - w.b, w.n), max_d = w.r + radius;
    float d = norm((w.b - p) + scale(t, w.n));
    if (d > max_d) return 0;
    if (t >= 0. && t <= w.l) return 1;
    d = t < 0 ? norm(w.b - p) : norm(w.e - p);
    return d > max_d ? 0 : 1;
}

int worms_cross(const Cyl& c0, const Cyl& c1) {
    const V v0 = cross(c0.n, c1.n);
    float r0 = norm(v0);
    if (r0 <= 1.e-20) return 1;   // parallel
    const V v1 = c1.b - c0.b;
    const float d = std::fabs(dot(v1, v0) / r0);
    float r2 = c0.r + c1.r;
    if (d > r2) return 0;
    const float b1 = dot(v1, c0.n), b2 = dot(v1, c1.n);
    const float a11 = 1., a12 = dot(c1.n, c0.n), a21 = -a12, a22 = -a11;
    r2 = a11 * a22 - a12 * a21;
    r0 = (b1 * a22 - b2 * a12) / r2;
    const float r1 = (b2 * a11 - b1 * a21) / r2;
    if (r0 >= 0. && r0 <= c0.l && r1 >= 0. && r1 <= c1.l) return 1;
    V nearp{0.f, 0.f, 0.f};
    int i0 = 0, i1 = 0;
    if (r0 >= 0. && r0 <= c0.l) {
        nearp = c0.b + scale(r0, c0.n);
        i0 = 1;
    }
    if (r1 >= 0. && r1 <= c1.l) {
        nearp = c1.b + scale(r1, c1.n);
        i1 = 1;
    }
    if (point_to_worm(c0, c1.b, c1.r) == 1) return 1;
    if (point_to_worm(c0, c1.e, c1.r) == 1) return 1;
    if (point_to_worm(c1, c0.b, c0.r) == 1) return 1;
    if (point_to_worm(c1, c0.e, c0.r) == 1) return 1;
    if (i0 == 1 && point_to_worm(c1, nearp, c0.r) == 1) return 1;
    if (i1 == 1 && point_to_worm(c0, nearp, c1.r) == 1) return 1;
    return 0;
}

char seed_to_worm(const Cyl& w, V p) {
    const float t = dot(p - w.b, w.n);
    const float r2 = w.r * w.r;
    if (t >= 0. && t <= w.l) {
        const V v0 = (w.b - p) + scale(t, w.n);
        return dot(v0, v0) <= r2 ? 1 : 0;
    }
    if (t > -w.r && t < 0.) {
        const V v0 = w.b - p;
        return dot(v0, v0) <= r2 ? 1 : 0;
    }
    if (t > w.l && t < w.l + w.r) {
        const V v0 = w.e - p;
        return dot(v0, v0) <= r2 ? 1 : 0;
    }
    return 0;
}

// worm() (particle.cpp:302-344): r < 0 when an end leaves the sphere R2
Cyl make_worm(float r1, float theta1, float phi1, float r2, float theta2, float phi2, float length, float radius,
              float R2) {
    const float ct = std::cos(theta1), st = std::sin(theta1), cp = std::cos(phi1), sp = std::sin(phi1);
    const V begin_sph = {r1, phi1, theta1}, end_sph = {r2, phi2, theta2};
    const V dir_sph = end_sph - begin_sph;
    Cyl c{};
    const V begin = {r1 * st * cp, r1 * st * sp, r1 * ct};
    if (dot(begin, begin) > R2) {
        c.r = -1.f;
        return c;
    }
    const V v0 = {st * cp * dir_sph.x - sp * dir_sph.y + ct * cp * dir_sph.z,
                  st * sp * dir_sph.x + cp * dir_sph.y + ct * sp * dir_sph.z, ct * dir_sph.x - st * dir_sph.z};
    const V dir = scale((float)(1. / norm(v0)), v0);
    const V end = begin + scale(length, dir);
    if (dot(end, end) > R2) {
        c.r = -1.f;
        return c;
    }
    c.r = radius;
    c.l = length;
    c.b = begin;
    c.e = end;
    c.n = dir;
    return c;
}

// vrand() (dda_func.cpp:68-83)
V random_unit(Stream& rng) {
    const float alpha = (float)((double)((float)rng.next() * 2) * kPi / 2147483647);
    const float cb = (float)(2. * (float)rng.next() / kRandMaxF - 1.);
    const float beta = std::acos(cb);
    const float sb = std::sin(beta);
    return {std::cos(alpha) * sb, std::sin(alpha) * sb, cb};
}

}  // namespace

extern "C" {

unsigned long long ddafft_rng_worker_seed(unsigned long long base_seed, int worker) {
    // main.cpp:803-804
    return base_seed + 0x9E3779B97F4A7C15ULL * (unsigned long long)(worker + 1);
}

int ddafft_rng_next(unsigned long long* state) {
    if (!state) return -1;
    if (*state == 0) *state = 0x9E3779B97F4A7C15ULL;   // dda_srand(0)
    Stream s{state};
    return s.next();
}

int ddafft_target_random_ellipsoid(int ni, int nj, int nl, double b_a, double c_a, int nsorts, const int* nseeds,
                                   unsigned long long* state, signed char* occ) {
    if (ni < 1 || nj < 1 || nl < 1 || nsorts < 1 || !nseeds || !state || !occ) return -2;
    if (*state == 0) *state = 0x9E3779B97F4A7C15ULL;
    Stream rng{state};
    const float ba = (float)b_a, ca = (float)c_a;
    const float a = (float)ni, b = a * ba, c = a * ca;
    if (b > (float)nj || c > (float)nl) return -1;
    size_t total = 0;
    for (int m = 0; m < nsorts; ++m) total += (size_t)nseeds[m];
    const V ctr = centre(ni, nj, nl);
    const Ell e = semi_axes2(a, b, c);
    std::vector<V> seeds(total);
    for (size_t m = 0; m < total; ++m) {
        V v0;
        bool dup;
        do {   // retries only while both outside and duplicated (particle.cpp:555-570)
            v0 = draw_point(rng, ni, nj, nl);
            dup = duplicate(seeds, 0, m, v0);
        } while (!e.inside(v0 - ctr) && dup);
        seeds[m] = v0;
    }
    int count = 0;
    for (int i = 0; i < ni; ++i)
        for (int j = 0; j < nj; ++j)
            for (int l = 0; l < nl; ++l) {
                const V p = {(float)i, (float)j, (float)l};
                signed char& o = occ[at(i, j, l, nj, nl)];
                if (e.inside(p - ctr)) {
                    const int mat = material_of(nearest(p, seeds, total, ni, nj, nl), nsorts, nseeds);
                    if (mat == 1) ++count;
                    if (mat > 0) o = (signed char)mat;
                } else {
                    o = 0;
                }
            }
    return count;
}

int ddafft_target_random_parallelepiped(int ni, int nj, int nl, double b_a, double c_a, int nsorts,
                                        const int* nseeds, unsigned long long* state, signed char* occ) {
    if (ni < 1 || nj < 1 || nl < 1 || nsorts < 1 || !nseeds || !state || !occ) return -2;
    if (*state == 0) *state = 0x9E3779B97F4A7C15ULL;
    Stream rng{state};
    const float ba = (float)b_a, ca = (float)c_a;
    const float a = (float)ni, b = a * ba, c = a * ca;
    if (b > (float)nj || c > (float)nl) return -1;
    size_t total = 0;
    for (int m = 0; m < nsorts; ++m) total += (size_t)nseeds[m];
    std::vector<V> seeds(total);
    for (size_t m = 0; m < total; ++m) {
        V v0;
        do {
            v0 = draw_point(rng, ni, nj, nl);
        } while (duplicate(seeds, 0, m, v0));
        seeds[m] = v0;
    }
    int count = 0;
    for (int i = 0; i < ni; ++i)
        for (int j = 0; j < nj; ++j)
            for (int l = 0; l < nl; ++l) {
                const V p = {(float)i, (float)j, (float)l};
                const int mat = material_of(nearest(p, seeds, total, ni, nj, nl), nsorts, nseeds);
                if (mat == 1) ++count;
                if (mat > 0) occ[at(i, j, l, nj, nl)] = (signed char)mat;
            }
    return count;
}

int ddafft_target_nibble_ellipsoid(int ni, int nj, int nl, double b_a, double c_a, int nsorts, const int* nseeds,
                                   int bites, double depth, const int* mv, unsigned long long* state,
                                   signed char* occ) {
    if (ni < 1 || nj < 1 || nl < 1 || nsorts < 1 || bites < 0 || !nseeds || !mv || !state || !occ) return -2;
    if (*state == 0) *state = 0x9E3779B97F4A7C15ULL;
    Stream rng{state};
    const float ba = (float)b_a, ca = (float)c_a, dp = (float)depth;
    const float a = (float)ni, b = a * ba, c = a * ca;
    const float a1 = (float)((1. - dp) * a), b1 = (float)((1. - dp) * b), c1 = (float)((1. - dp) * c);
    if (b > (float)nj || c > (float)nl) return -1;
    size_t total = 0;
    for (int m = 0; m < nsorts; ++m) total += (size_t)nseeds[m];
    const V ctr = centre(ni, nj, nl);
    const Ell e = semi_axes2(a, b, c), e1 = semi_axes2(a1, b1, c1);
    std::vector<V> seeds(total + bites);
    for (size_t m = 0; m < total; ++m) {
        V v0;
        bool dup;
        do {
            v0 = draw_point(rng, ni, nj, nl);
            dup = duplicate(seeds, 0, m, v0);
        } while (!e1.inside(v0 - ctr) && dup);
        seeds[m] = v0;
    }
    for (size_t m = total; m < total + bites; ++m) {
        V v0;
        bool again;
        do {   // bites sit in the shell between the two ellipsoids (particle.cpp:742-764)
            v0 = draw_point(rng, ni, nj, nl);
            const V v1 = v0 - ctr;
            const float x2 = v1.x * v1.x, y2 = v1.y * v1.y, z2 = v1.z * v1.z;
            if ((x2 / e.a + y2 / e.b + z2 / e.c) <= 1. && (x2 / e1.a + y2 / e1.b + z2 / e1.c) > 1.)
                again = duplicate(seeds, total, m, v0);
            else
                again = true;
        } while (again);
        seeds[m] = v0;
    }
    int count = 0;
    for (int i = 0; i < ni; ++i)
        for (int j = 0; j < nj; ++j)
            for (int l = 0; l < nl; ++l) {
                const V p = {(float)i, (float)j, (float)l};
                signed char& o = occ[at(i, j, l, nj, nl)];
                if (e.inside(p - ctr)) {
                    const int flag = nearest(p, seeds, total + bites, ni, nj, nl);
                    if (flag >= (int)total) {
                        o = 0;
                    } else {
                        const int mat = material_of(flag, nsorts, nseeds);
                        if (mat > 0) o = (signed char)mat;
                    }
                } else {
                    o = 0;
                }
                if (mv[(int)o] != 0) ++count;
            }
    return count;
}

int ddafft_target_hierarchical_ellipsoid(int levels, int ni, int nj, int nl, double b_a, double c_a, int nsorts,
                                         const int* nseeds, unsigned long long* state, signed char* occ) {
    const int rc = ddafft_target_random_ellipsoid(ni, nj, nl, b_a, c_a, nsorts, nseeds, state, occ);
    if (rc < 0 || levels <= 1) return rc;
    // one more level inside material 1 (particle.cpp:908-955; the reference
    // runs the refinement once, whatever the number of levels above one)
    Stream rng{state};
    size_t total = 0;
    for (int m = 0; m < nsorts; ++m) total += (size_t)nseeds[m];
    std::vector<V> seeds(total);
    // the same seed array carries over from the first level, so duplicates are
    // tested against the seeds drawn so far at this level only
    for (size_t m = 0; m < total; ++m) {
        V v0;
        do {
            int i0;
            do {
                v0 = draw_point(rng, ni, nj, nl);
                i0 = occ[at((int)v0.x, (int)v0.y, (int)v0.z, nj, nl)];
            } while (i0 != 1);
        } while (duplicate(seeds, 0, m, v0));
        seeds[m] = v0;
    }
    int count = 0;
    for (int i = 0; i < ni; ++i)
        for (int j = 0; j < nj; ++j)
            for (int l = 0; l < nl; ++l) {
                signed char& o = occ[at(i, j, l, nj, nl)];
                if ((int)o != 1) continue;
                const V p = {(float)i, (float)j, (float)l};
                const int mat = material_of(nearest(p, seeds, total, ni, nj, nl), nsorts, nseeds);
                if (mat == 1) ++count;
                if (mat > 0) o = (signed char)mat;
            }
    return count;
}

int ddafft_target_sphere_spec(int ni, int nj, int nl, double radius, double rho, signed char* occ) {
    if (ni < 1 || nj < 1 || nl < 1 || !occ) return -2;
    const V ctr = centre(ni, nj, nl);
    const float R = (float)radius;
    const float r = std::pow((float)rho, (float)(1. / 3.)) * R, r2 = r * r;
    int count = 0;
    for (int i = 0; i < ni; ++i)
        for (int j = 0; j < nj; ++j)
            for (int l = 0; l < nl; ++l) {
                const V v = V{(float)i, (float)j, (float)l} - ctr;
                const bool in = (v.x * v.x + v.y * v.y + v.z * v.z) <= r2;
                occ[at(i, j, l, nj, nl)] = in ? 1 : 0;
                count += in;
            }
    return count;
}

int ddafft_worm_count(int ni, int nj, int nl, double rho, double radius_worm, double length_ratio) {
    // main.cpp:580-581
    const float r = (float)rho, rw = (float)radius_worm, lr = (float)length_ratio;
    return (int)(r * (float)(ni * nj * nl) / (8. * (rw * rw * rw) * (0.75 * lr + 1))) + 1;
}

int ddafft_target_worms(int kind, int ni, int nj, int nl, double radius_worm, double length_worm, int nworms,
                        long long max_attempts, unsigned long long* state, signed char* occ) {
    if ((kind != 6 && kind != 7 && kind != 8) || ni < 1 || nj < 1 || nl < 1 || nworms < 1 || !state || !occ)
        return -2;
    if (*state == 0) *state = 0x9E3779B97F4A7C15ULL;
    Stream rng{state};
    const float rw = (float)radius_worm, lw = (float)length_worm;
    const float R = (float)((float)ni / 2.);
    float R2 = (R - rw) * (R - rw);
    const V ctr = centre(ni, nj, nl);
    std::vector<Cyl> worms((size_t)nworms);
    long long attempts = 0;
    int placed = 0;
    while (placed < nworms) {
        Cyl c{};
        c.r = -1.f;
        while (c.r < 0.) {
            if (max_attempts > 0 && ++attempts > max_attempts) return -3;
            if (kind == 6) {   // anthrax, particle.cpp:36-50
                const float r1 = R * (float)rng.next() / kRandMaxF;
                const float r2 = r1;
                const float phi1 = (float)(2. * kPi * (float)rng.next() / kRandMaxF);
                const float phi2 = (float)(2. * kPi * (float)rng.next() / kRandMaxF);
                float r0 = (float)(2. * (float)rng.next() / kRandMaxF - 1.);
                const float theta1 = std::acos(r0);
                r0 = (float)(2. * (float)rng.next() / kRandMaxF - 1.);
                const float theta2 = std::acos(r0);
                c = make_worm(r1, theta1, phi1, r2, theta2, phi2, lw, rw, R2);
            } else if (kind == 7) {   // hedgehog, particle.cpp:105-118
                const float r1 = R * (float)rng.next() / kRandMaxF;
                const float r2 = (2. * (float)rng.next() / kRandMaxF - 1.) <= 0. ? (float)(r1 - 0.1) : (float)(r1 + 0.1);
                const float phi1 = (float)(2. * kPi * (float)rng.next() / kRandMaxF);
                const float theta1 = (float)std::acos(2. * (float)rng.next() / kRandMaxF - 1.);
                c = make_worm(r1, theta1, phi1, r2, theta1, phi1, lw, rw, R2);
            } else {   // dead anthrax, particle.cpp:174-193
                const float r1 = R * (float)rng.next() / kRandMaxF;
                const float phi1 = (float)(2. * kPi * (float)rng.next() / kRandMaxF);
                const float r0 = (float)(2. * (float)rng.next() / kRandMaxF - 1.);
                const float theta1 = std::acos(r0);
                const float r2 = std::sin(theta1);
                c.l = lw;
                c.n = random_unit(rng);
                c.b = {r1 * r2 * std::cos(phi1), r1 * r2 * std::sin(phi1), r1 * r0};
                c.e = c.b + scale(c.l, c.n);
                c.r = dot(c.e, c.e) <= R2 ? rw : -1.f;
            }
        }
        bool crossing = false;
        for (int j = 0; j < placed && !crossing; ++j) crossing = worms_cross(c, worms[j]) == 1;
        if (!crossing) worms[placed++] = c;
    }
    R2 = R * R;
    int count = 0;
    for (int i = 0; i < ni; ++i)
        for (int j = 0; j < nj; ++j)
            for (int l = 0; l < nl; ++l) {
                const V v1 = V{(float)i, (float)j, (float)l} - ctr;
                signed char& o = occ[at(i, j, l, nj, nl)];
                o = 0;
                if ((v1.x * v1.x) + (v1.y * v1.y) + (v1.z * v1.z) <= R2)
                    for (int m = 0; m < nworms; ++m) {
                        o = seed_to_worm(worms[m], v1);
                        if (o != 0) {
                            ++count;
                            break;
                        }
                    }
            }
    return count;
}

int ddafft_orientation_mc(unsigned long long* state, float* rot) {
    // main.cpp:875-893: Euler angles with cos(psi) ~ U(-1, 1), iota, kappa ~ U(0, 2 pi)
    if (!state || !rot) return DDAFFT_EINVAL;
    if (*state == 0) *state = 0x9E3779B97F4A7C15ULL;
    Stream rng{state};
    const float cps = (float)((double)(float)rng.next() * 2. / (double)kRandMaxF - 1.);
    const float psi = std::acos(cps);
    const float sps = std::sin(psi);
    const float iota = (float)((double)(float)rng.next() * 2. * kPi / (double)kRandMaxF);
    const float ci = std::cos(iota), si = std::sin(iota);
    const float kap = (float)((double)(float)rng.next() * 2. * kPi / (double)kRandMaxF);
    const float ck = std::cos(kap), sk = std::sin(kap);
    rot[0] = ci * ck - cps * si * sk;
    rot[1] = si * ck + cps * ci * sk;
    rot[2] = sps * sk;
    rot[3] = -ci * sk - cps * si * ck;
    rot[4] = -si * sk + cps * ci * ck;
    rot[5] = sps * ck;
    rot[6] = sps * si;
    rot[7] = -sps * ci;
    rot[8] = cps;
    return DDAFFT_OK;
}

int ddafft_rotate(const float* rot, int n, const float* in, float* out) {
    // main.cpp:895-921: v' = T v, row by row in float
    if (!rot || n < 0 || (n > 0 && (!in || !out))) return DDAFFT_EINVAL;
    for (int k = 0; k < n; ++k) {
        const float x = in[3 * k], y = in[3 * k + 1], z = in[3 * k + 2];
        out[3 * k] = rot[0] * x + rot[1] * y + rot[2] * z;
        out[3 * k + 1] = rot[3] * x + rot[4] * y + rot[5] * z;
        out[3 * k + 2] = rot[6] * x + rot[7] * y + rot[8] * z;
    }
    return DDAFFT_OK;
}

int ddafft_target_file_read(const char* path, int ni, int nj, int nl, signed char* occ) {
    // main.cpp:645-677: the count N, then N lines "X Y Z T"; cells not listed are 0
    if (!path || !occ || ni < 1 || nj < 1 || nl < 1) return -2;
    FILE* f = std::fopen(path, "r");
    if (!f) return -2;
    std::memset(occ, 0, (size_t)ni * nj * nl);
    char tok[256];
    if (std::fscanf(f, "%255s", tok) != 1) {
        std::fclose(f);
        return -4;
    }
    const int n = std::atoi(tok);
    int v[4];
    for (int k = 0; k < n; ++k) {
        for (int q = 0; q < 4; ++q) {
            if (std::fscanf(f, "%255s", tok) != 1) {
                std::fclose(f);
                return -4;
            }
            v[q] = std::atoi(tok);
        }
        if (v[0] >= ni || v[1] >= nj || v[2] >= nl || v[0] < 0 || v[1] < 0 || v[2] < 0) {
            std::fclose(f);
            return -3;   // the reference exits with status 3 here
        }
        occ[at(v[0], v[1], v[2], nj, nl)] = (signed char)v[3];
    }
    std::fclose(f);
    return n;
}

int ddafft_target_file_write(const char* path, int ni, int nj, int nl, const signed char* occ) {
    if (!path || !occ || ni < 1 || nj < 1 || nl < 1) return -2;
    int n = 0;
    for (size_t k = 0; k < (size_t)ni * nj * nl; ++k) n += occ[k] != 0;
    FILE* f = std::fopen(path, "w");
    if (!f) return -2;
    std::fprintf(f, "%d\n", n);
    for (int i = 0; i < ni; ++i)
        for (int j = 0; j < nj; ++j)
            for (int l = 0; l < nl; ++l) {
                const int t = occ[at(i, j, l, nj, nl)];
                if (t) std::fprintf(f, "%d %d %d %d\n", i, j, l, t);
            }
    std::fclose(f);
    return n;
}

// ---- generators of the benchmark workloads that the reference has no target for.
// They draw from the same xorshift64* stream; the arithmetic is double.

int ddafft_target_mixture(int n, const signed char* base, int material, double fraction,
                          unsigned long long* state, signed char* occ) {
    if (n < 0 || !base || !occ || !state || material < 1 || material > 64) return -2;
    if (*state == 0) *state = 0x9E3779B97F4A7C15ULL;
    Stream rng{state};
    int changed = 0;
    for (int k = 0; k < n; ++k) {
        occ[k] = base[k];
        if (base[k] == 1) {
            const double u = (double)rng.next() / 2147483647.0;
            if (u < fraction) {
                occ[k] = (signed char)material;
                ++changed;
            }
        }
    }
    return changed;
}

// consecutive rejected centre draws after which a packing starts over
const long long kAggStall = 100000;

int ddafft_agglomerate_count(int ni, int nj, int nl, double r_mono, double packing) {
    if (ni < 1 || nj < 1 || nl < 1 || !(r_mono > 0.) || !(packing > 0.)) return -2;
    const double R = 0.5 * (double)std::min(ni, std::min(nj, nl));
    const double ratio = R / r_mono;
    return (int)std::floor(packing * ratio * ratio * ratio + 0.5);
}

int ddafft_target_agglomerate(int ni, int nj, int nl, double r_mono, double packing, long long max_attempts,
                              unsigned long long* state, signed char* occ, double* centres) {
    const int K = ddafft_agglomerate_count(ni, nj, nl, r_mono, packing);
    if (K < 0 || !state || !occ) return -2;
    if (*state == 0) *state = 0x9E3779B97F4A7C15ULL;
    Stream rng{state};
    const double R = 0.5 * (double)std::min(ni, std::min(nj, nl));
    const double Rc = R - r_mono;
    if (!(Rc > 0.)) return -2;
    const double cx = 0.5 * (ni - 1), cy = 0.5 * (nj - 1), cz = 0.5 * (nl - 1);
    std::vector<double> c((size_t)3 * K);
    long long attempts = 0, stall = 0;
    const double dmin2 = 4. * r_mono * r_mono;
    for (int k = 0; k < K;) {
        if (max_attempts > 0 && ++attempts > max_attempts) return -3;
        // a packing jammed short of K (no free room left for a centre) starts over
        if (++stall > kAggStall) {
            k = 0;
            stall = 0;
        }
        double v[3];
        for (double& q : v) q = Rc * (2. * ((double)rng.next() / 2147483647.0) - 1.);
        if (v[0] * v[0] + v[1] * v[1] + v[2] * v[2] > Rc * Rc) continue;
        const double x = cx + v[0], y = cy + v[1], z = cz + v[2];
        bool ok = true;
        for (int q = 0; q < k && ok; ++q) {
            const double dx = x - c[3 * q], dy = y - c[3 * q + 1], dz = z - c[3 * q + 2];
            ok = dx * dx + dy * dy + dz * dz >= dmin2;
        }
        if (!ok) continue;
        c[3 * k] = x;
        c[3 * k + 1] = y;
        c[3 * k + 2] = z;
        ++k;
        stall = 0;
    }
    const double r2 = r_mono * r_mono;
    std::memset(occ, 0, (size_t)ni * nj * nl);
    int count = 0;
    for (int k = 0; k < K; ++k) {
        const double x = c[3 * k], y = c[3 * k + 1], z = c[3 * k + 2];
        const int i0 = std::max(0, (int)std::floor(x - r_mono)), i1 = std::min(ni - 1, (int)std::ceil(x + r_mono));
        const int j0 = std::max(0, (int)std::floor(y - r_mono)), j1 = std::min(nj - 1, (int)std::ceil(y + r_mono));
        const int l0 = std::max(0, (int)std::floor(z - r_mono)), l1 = std::min(nl - 1, (int)std::ceil(z + r_mono));
        for (int i = i0; i <= i1; ++i)
            for (int j = j0; j <= j1; ++j)
                for (int l = l0; l <= l1; ++l) {
                    const double dx = i - x, dy = j - y, dz = l - z;
                    signed char& o = occ[at(i, j, l, nj, nl)];
                    if (dx * dx + dy * dy + dz * dz <= r2 && o == 0) {
                        o = 1;
                        ++count;
                    }
                }
    }
    if (centres) std::memcpy(centres, c.data(), c.size() * sizeof(double));
    return count;
}

}  // extern "C"
