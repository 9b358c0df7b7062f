// libddafft_amd: the DDA-FFT solve path on MI355X behind the C ABI of
// include/ddafft.h.  Host code here only stages tiny inputs, builds the
// constant tables and sequences kernels on the context's stream; every
// pass over lattice-sized data runs on the GPU.
#include <hip/hip_runtime.h>

#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ddafft.h"
#include "ahat_kernels.hpp"
#include "device_types.hpp"
#include "../host/host_internal.hpp"
#include "fft_common.hpp"
#include "gradient_kernels.hpp"
#include "matvec_kernels.hpp"
#include "solver_kernels.hpp"
#include "tuned_table.hpp"

using namespace ddafft;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define HIPCHECK(expr)                                                                        \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return fail(DDAFFT_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_));      \
    } while (0)

#define TRY(x)                  \
    do {                        \
        int rc_ = (x);          \
        if (rc_) return rc_;    \
    } while (0)

const double kPi = 3.1415926535897932384626433832795;

bool pow2_ok(int n) { return n >= 4 && n <= 512 && (n & (n - 1)) == 0; }

// Cells of a lattice at most 2^24: the passes address Â (6 complex values
// per padded cell, 8 N cells) with 32-bit offsets.
constexpr long long kMaxCells = 1LL << 24;

// FourierPhase (fft_func.cpp:77-88): W_2n^k for k < n by a float recurrence.
std::vector<float2> fourier_phase(int n) {
    std::vector<float2> w(n > 2 ? n : 2);
    const float a = (float)(-2. * kPi / (2 * n));
    const float2 cw = make_float2(std::cos(a), std::sin(a));
    w[0] = make_float2(1.f, 0.f);
    w[1] = cw;
    for (int i = 2; i < n; ++i)
        w[i] = make_float2(w[i - 1].x * cw.x - w[i - 1].y * cw.y, w[i - 1].x * cw.y + w[i - 1].y * cw.x);
    w.resize(n);
    return w;
}

// Exact float twiddles W_N^q (fft_kernel.hpp:238-243).
std::vector<float2> plan_twiddles(int n) {
    std::vector<float2> t(n);
    const double two_pi = 6.283185307179586476925286766559;
    for (int q = 0; q < n; ++q) {
        const double a = -two_pi * (double)q / (double)n;
        t[q] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
    return t;
}

// One octant of the Green tensor (dda_func.cpp:92-147) at r = (a d, b d, c d).
void green_octant(int ni, int nj, int nl, float d0, std::vector<float2>& oct) {
    oct.assign((size_t)(ni + 1) * (nj + 1) * (nl + 1) * 6, make_float2(0.f, 0.f));
    for (int a = 0; a <= ni; ++a) {
        const float rx = (float)a * d0;
        for (int b = 0; b <= nj; ++b) {
            const float ry = (float)b * d0;
            for (int c = 0; c <= nl; ++c) {
                const float rz = (float)c * d0;
                float2* o = &oct[(((size_t)a * (nj + 1) + b) * (nl + 1) + c) * 6];
                const float e00 = rx * rx, e11 = ry * ry, e22 = rz * rz;
                const float e01 = rx * ry, e02 = rx * rz, e12 = ry * rz;
                const float r2 = e00 + e11 + e22;
                if ((double)r2 < 0.0000000001) continue;
                const float r1 = std::sqrt(r2);
                const float2 c0 = make_float2(r2 - 1.f, r1);
                const float2 c1 = make_float2(3.f - r2, -3.f * r1);
                const float r3 = r1 * r2;
                float2 c3 = make_float2(std::cos(r1), std::sin(r1));
                const float2 c2 = make_float2(c3.x / r3, c3.y / r3);
                c3 = make_float2(c0.x * c2.x - c0.y * c2.y, c0.x * c2.y + c0.y * c2.x);
                const float2 c4 = make_float2(c1.x * c2.x - c1.y * c2.y, c1.x * c2.y + c1.y * c2.x);
                const float2 c5 = make_float2(c4.x / r2, c4.y / r2);
                const float e[6] = {e00, e01, e02, e11, e12, e22};
                for (int q = 0; q < 6; ++q) o[q] = make_float2(e[q] * c5.x, e[q] * c5.y);
                o[0] = make_float2(o[0].x + c3.x, o[0].y + c3.y);
                o[3] = make_float2(o[3].x + c3.x, o[3].y + c3.y);
                o[5] = make_float2(o[5].x + c3.x, o[5].y + c3.y);
            }
        }
    }
}

// Set while an allocation-free call (ddafft_solve_noalloc) runs on this host
// thread: any path that would allocate fails loudly instead.
thread_local bool t_noalloc = false;

template <class T>
int dev_alloc(T** p, size_t count) {
    *p = nullptr;
    if (count == 0) return DDAFFT_OK;
    if (t_noalloc) return fail(DDAFFT_ESTATE, "internal: device allocation inside the allocation-free solve");
    hipError_t e = hipMalloc((void**)p, count * sizeof(T));
    if (e != hipSuccess) return fail(DDAFFT_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    return DDAFFT_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// Device scratch of the allocation-free solve (ddafft_solve_noalloc): a
// file-scope array of fixed size, one copy per GPU, held until the process
// ends.  Host code never passes `scratch` itself to a kernel (its host-side
// name is only a stand-in); it uses the device address that
// hipGetSymbolAddress returns, looked up once per device when a context is
// created there.  Every allocation-free call on one GPU shares that GPU's
// array: such calls must not overlap.  Nothing in it needs to start at zero
// (every kernel that reads a part of it reads what an earlier kernel of the
// same call wrote).
// ---------------------------------------------------------------------------

constexpr size_t kScratchBytes = (size_t)32 << 20;
static __device__ __align__(256) char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
char* g_scratch_dev[kMaxDevices] = {};
bool g_scratch_done[kMaxDevices] = {};

// Device address of this device's scratch (the calling thread's current
// device), looked up once; nullptr when the lookup failed.
char* scratch_address(int dev) {
    if (dev < 0 || dev >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_done[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        g_scratch_dev[dev] = e == hipSuccess ? static_cast<char*>(p) : nullptr;
        g_scratch_done[dev] = true;
    }
    return g_scratch_dev[dev];
}

}  // namespace

// ---------------------------------------------------------------------------
// Context
// ---------------------------------------------------------------------------

struct ddafft_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    Dims d{};
    int nrhs_max = 1;
    float kd = 0.f;

    float2* ahat = nullptr;            // [2nj][2nl][2ni][6]
    float2* tables = nullptr;
    const float2 *wpre_i = nullptr, *wpre_j = nullptr, *wpre_l = nullptr;
    const float2 *tw_i = nullptr, *tw_j = nullptr, *tw_l = nullptr;      // n-point plans
    const float2 *tw_2i = nullptr, *tw_2j = nullptr, *tw_2l = nullptr;   // 2n-point plans

    // target
    int8_t* occ = nullptr;
    int8_t* dipf = nullptr;            // [n] 1 where the cell holds a dipole (mv[occ] != 0)
    int* mv = nullptr;
    int nmat = 0;
    int* cells = nullptr;   // dipole cell indices (far field)
    int ncells = 0;
    int2* ffseg = nullptr;  // far-field column segments {first dipole, (i nj + j) << 6 | count}
    int nseg = 0;
    long long ndip = 0;
    std::vector<long long> mat_count;  // dipoles per material index
    bool have_target = false;

    // matvec scratch
    float2 *b1 = nullptr, *b2 = nullptr, *b3 = nullptr;
    // solver fields [nrhs][3][n]
    float2 *P = nullptr, *Pmin = nullptr, *R = nullptr, *S = nullptr, *G = nullptr, *H = nullptr, *W = nullptr;
    float2 *X = nullptr, *Y = nullptr;
    float2* alpha = nullptr;           // [nrhs][nmat]
    float2* einc = nullptr;            // [nrhs][3]
    float* kinc = nullptr;             // [nrhs][3] incident direction per right-hand side
    int* inc_axis = nullptr;           // [nrhs] axis of an axis-aligned incidence, else -1
    float2* inc_tab = nullptr;         // [nrhs][2][max(ni, nj, nl)] tabulated incident phases
    std::vector<int> inc_axis_h;       // host staging of the two (kept alive for async copies)
    std::vector<float2> inc_tab_h;
    float2* inc_cell = nullptr;        // [nrhs][2][n] host-evaluated phases of any incidence (exact_phases)
    size_t inc_cell_cap = 0;
    std::vector<float2> inc_cell_h;
    std::vector<int8_t> dipf_h;        // host copy of the dipole flags
    SolveState* st = nullptr;          // [nrhs]
    double* hist = nullptr;
    int hist_cap = 0;
    double* partials = nullptr;
    size_t partials_cap = 0;
    float* dirs = nullptr;
    int ndirs = 0, dirs_cap = 0;       // resident directions (dirs_cap in directions)
    std::vector<char> dir_pairs;       // per-RHS directions: do RHS 2q and 2q+1 share them?
    // gradients: forward dipoles kept for the adjoint, per-axis phase table,
    // small per-call arrays
    float2* Pf = nullptr;
    size_t Pf_cap = 0;
    float2* phase = nullptr;
    size_t phase_cap = 0;
    char* gaux = nullptr;
    size_t gaux_cap = 0;
    float2* esc = nullptr;
    int esc_cap = 0;
    float2* fftab = nullptr;           // far-field per-axis phases [ndirs][ni + nj + nl]
    size_t fftab_cap = 0;
    int sum_mode = 0;                  // observables: 0 FP64 sums, 1 the reference's sequential float sums
    int ahat_mode = 0;                 // 0: full Â as the reference builds it; 1: one octant (Nyquist planes zeroed)
    int ff_mfma = -1;                  // far field on the matrix cores: -1 default, 0 off, 1 on
    int ff_group = -1;                 // far field by groups of directions sharing z: -1 default (on), 0 off, 1 on
    int ffg_n = 0, ffg_S = 0;          // the resident direction set as ffg_n groups of ffg_S (0: no such split)
    bool ffg_off_call = false;         // set by an allocation-free solve whose scratch cannot hold the grouped form
    float2* ffpc = nullptr;            // dipole moments in dipole-list order for the grouped far field
    size_t ffpc_cap = 0;
    // launch schedule, 0 / -1: from the tuned table, else the built-in rule
    int x_blocks = 0;                  // X pass: blocks per launch up to which right-hand-side groups are spread
    int x_rb = 0;                      // X pass: right-hand sides per tile
    int fuse_zy = -1;                  // z/y forward fused per i-plane where the plane allows it (-1: schedule)
    // one solver iteration captured as a HIP graph and replayed (use_graph:
    // 1 yes, 0 no, -1 the tuned choice); recaptured when anything its
    // launches bake in changes
    int use_graph = -1;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    std::vector<char> gkey;
    float* seqbuf = nullptr;           // terms of the sequential float sums (sum_mode 1)
    size_t seqbuf_cap = 0;
    SolveState* state_ring = nullptr;  // pinned [4][nrhs_max]: solver states after each queued iteration
    SolveState* ring_dev = nullptr;    // its device address: the iteration's last scalar step writes it
    // per-solve inputs in one device block with a page-locked mirror of the
    // same layout (st | alpha | einc | kinc | inc_axis | inc_tab): staged on
    // the host, then one copy of the range written
    char* small = nullptr;
    char* small_h = nullptr;
    size_t small_bytes = 0, dirty_lo = 0, dirty_hi = 0;
    bool pending_first_approx = false; // job staged: the first approximation runs after the copy
    char* dn_dev = nullptr;            // device address of the pinned result buffer dn
    char* scratch = nullptr;           // this device's allocation-free scratch (hipGetSymbolAddress), or nullptr
    alignas(256) char noalloc_dn[1 << 20];   // host side of the results of an allocation-free solve
    hipEvent_t fin_ev = nullptr;       // marks the results of a batched solve as copied back
    // pinned staging of the per-solve transfers, so that they are truly
    // asynchronous: uploads are packed at up_off (reset when a solve starts,
    // with the stream idle), results come back into dn
    char* up = nullptr;
    size_t up_cap = 0, up_off = 0;
    char* dn = nullptr;
    size_t dn_cap = 0;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};

    // opt-in per-pass timing of every matvec (ddafft_profile)
    bool prof = false;
    std::vector<hipEvent_t> prof_ev;   // 6 events per profiled matvec
    std::vector<int> prof_iter;        // solver iteration of each recorded matvec
    int prof_used = 0;                 // matvecs recorded since the last collect
    double prof_ms[5] = {0, 0, 0, 0, 0};
    long long prof_n = 0;
};

namespace {

int flush_small(ddafft_ctx* c);
int inc_tab_len(const ddafft_ctx* c);

// The launch schedule of one matvec / iteration: the tuned table's entry for
// this lattice and right-hand-side count (tuned_table.hpp, written by
// dda-fft_amd/tune.py on an MI355X), else the built-in rule; a value set
// through ddafft_set_tuning wins over both.
struct Schedule {
    int x_rb, x_blocks, graph, fuse;
};

Schedule schedule(const ddafft_ctx* c, int nrhs) {
    // built-in rule: four right-hand sides per X tile when they come in fours
    // on lines of at most 32 points (32^3, 16 RHS: 41 us against 44.5 in
    // pairs; 64^3: 425 against 315), every group its own block, graphs on;
    // z/y forward fused when its launch (one block per (component, i)-plane)
    // has at least 192 blocks: below that the split passes' wider grids win
    // (32^3, 1 RHS: 38.0 us a matvec split, 39.3 fused; 64^3, 16 RHS: 709
    // against 675)
    Schedule s{(nrhs % 4 == 0 && c->d.ni <= 32) ? 4 : 2, 1 << 30, 1, c->d.ni * 3 * nrhs >= 192 ? 1 : 0};
    for (const TunedEntry& e : kTuned)
        if (e.ni == c->d.ni && e.nj == c->d.nj && e.nl == c->d.nl && e.nrhs == nrhs) {
            s = Schedule{e.x_rb, e.x_blocks, e.graph, e.fuse};
            break;
        }
    if (c->x_rb > 0) s.x_rb = c->x_rb;
    if (c->x_blocks > 0) s.x_blocks = c->x_blocks;
    if (c->use_graph >= 0) s.graph = c->use_graph;
    if (c->fuse_zy >= 0) s.fuse = c->fuse_zy;
    return s;
}

int ensure_partials(ddafft_ctx* c, size_t doubles) {
    if (doubles <= c->partials_cap) return DDAFFT_OK;
    if (c->stream) HIPCHECK(hipStreamSynchronize(c->stream));
    if (c->partials) HIPCHECK(hipFree(c->partials));
    c->partials = nullptr;
    c->partials_cap = 0;
    int rc = dev_alloc(&c->partials, doubles);
    if (rc) return rc;
    c->partials_cap = doubles;
    return DDAFFT_OK;
}

// grid sizes of the five passes (tile widths as in matvec_kernels.hpp)
int tile_zf(int n) {
    switch (n) {
        case 4: return TileZF<4>::T;
        case 8: return TileZF<8>::T;
        case 16: return TileZF<16>::T;
        case 32: return TileZF<32>::T;
        case 64: return TileZF<64>::T;
        case 128: return TileZF<128>::T;
        case 256: return TileZF<256>::T;
        case 512: return TileZF<512>::T;
    }
    return 1;
}
template <int RB>
int tile_x_rb(int n) {
    switch (n) {
        case 4: return TileX<4, RB>::T;
        case 8: return TileX<8, RB>::T;
        case 16: return TileX<16, RB>::T;
        case 32: return TileX<32, RB>::T;
        case 64: return TileX<64, RB>::T;
        case 128: return TileX<128, RB>::T;
        case 256: return TileX<256, RB>::T;
        case 512: return TileX<512, 1>::T;
    }
    return 1;
}
int tile_x(int n, int rb) { return rb == 4 ? tile_x_rb<4>(n) : rb == 2 ? tile_x_rb<2>(n) : tile_x_rb<1>(n); }
int tile_yi(int n2) {
    switch (n2) {
        case 8: return TileYI<8>::T;
        case 16: return TileYI<16>::T;
        case 32: return TileYI<32>::T;
        case 64: return TileYI<64>::T;
        case 128: return TileYI<128>::T;
        case 256: return TileYI<256>::T;
        case 512: return TileYI<512>::T;
        case 1024: return TileYI<1024>::T;
    }
    return 1;
}
int tile_zi(int n2) {
    switch (n2) {
        case 8: return TileZI<8>::T;
        case 16: return TileZI<16>::T;
        case 32: return TileZI<32>::T;
        case 64: return TileZI<64>::T;
        case 128: return TileZI<128>::T;
        case 256: return TileZI<256>::T;
        case 512: return TileZI<512>::T;
        case 1024: return TileZI<1024>::T;
    }
    return 1;
}

int zi_blocks(const Dims& d) {
    const int tj = std::min(tile_zi(2 * d.nl), d.nj);
    return d.ni * (d.nj / tj);
}

#define DDA_CASE_N(N, ...)    \
    case N: {                 \
        constexpr int NN = N; \
        __VA_ARGS__;          \
    } break;

// line lengths of the forward passes (input lines of n points)
#define DDA_SWITCH_SMALL(n, ...)                                                                        \
    switch (n) {                                                                                         \
        DDA_CASE_N(4, __VA_ARGS__) DDA_CASE_N(8, __VA_ARGS__) DDA_CASE_N(16, __VA_ARGS__)                \
        DDA_CASE_N(32, __VA_ARGS__) DDA_CASE_N(64, __VA_ARGS__) DDA_CASE_N(128, __VA_ARGS__)             \
        DDA_CASE_N(256, __VA_ARGS__) DDA_CASE_N(512, __VA_ARGS__)                                        \
        default: return fail(DDAFFT_EINVAL, "unsupported line length");                                 \
    }

// X-pass tiles of two or four right-hand sides: lines up to 256 points (a
// 512-point axis takes one right-hand side per tile, whose LDS fits)
#define DDA_SWITCH_X(n, ...)                                                                            \
    switch (n) {                                                                                         \
        DDA_CASE_N(4, __VA_ARGS__) DDA_CASE_N(8, __VA_ARGS__) DDA_CASE_N(16, __VA_ARGS__)                \
        DDA_CASE_N(32, __VA_ARGS__) DDA_CASE_N(64, __VA_ARGS__) DDA_CASE_N(128, __VA_ARGS__)             \
        DDA_CASE_N(256, __VA_ARGS__)                                                                     \
        default: return fail(DDAFFT_EINVAL, "unsupported line length");                                 \
    }

// line lengths of the inverse passes and of the A-hat transform (2 n points)
#define DDA_SWITCH_LARGE(n, ...)                                                                        \
    switch (n) {                                                                                         \
        DDA_CASE_N(8, __VA_ARGS__) DDA_CASE_N(16, __VA_ARGS__) DDA_CASE_N(32, __VA_ARGS__)               \
        DDA_CASE_N(64, __VA_ARGS__) DDA_CASE_N(128, __VA_ARGS__) DDA_CASE_N(256, __VA_ARGS__)            \
        DDA_CASE_N(512, __VA_ARGS__) DDA_CASE_N(1024, __VA_ARGS__)                                       \
        default: return fail(DDAFFT_EINVAL, "unsupported line length");                                 \
    }

struct PassEvents {
    hipEvent_t* ev;  // 6 events bracketing the 5 passes, or nullptr
};

// Host-side check that every pointer a pass will dereference is set, so that
// a sequencing mistake returns an error instead of faulting the GPU.
int check_args(const ProArgs& pa, const EpiArgs& ea, bool solver) {
    bool ok = pa.A != nullptr;
    if (pa.mode == PRO_LC1 || pa.mode == PRO_LC2) ok = ok && pa.B;
    if (pa.mode == PRO_LC4) ok = ok && pa.B && pa.D;
    if ((pa.P == nullptr) != (pa.Pmin == nullptr)) ok = false;
    switch (ea.kind) {
        case EK_CONV: ok = ok && ea.out; break;
        case EK_MV: case EK_BICG_W: case EK_AA_S0: case EK_AA_W: ok = ok && ea.x && ea.out; break;
        case EK_BICG_H: case EK_CG_W: ok = ok && ea.x && ea.out && ea.y; break;
        case EK_BICG_R0: ok = ok && ea.out && ea.out2 && ea.out3; break;
        case EK_CG_R0: ok = ok && ea.out && ea.out2; break;
        case EK_AA_R0: ok = ok && ea.out; break;
        case EK_SIM_R0: ok = ok && ea.x && ea.out && ea.out2 && ea.out3; break;
        case EK_AA_SUPD: ok = ok && ea.x && ea.out2; break;
        case EK_SIM: ok = ok && ea.x && ea.out && ea.out2 && ea.y; break;
        default: ok = false;
    }
    if (ea.kind != EK_CONV) ok = ok && ea.alpha && ea.nmat > 0;
    if (solver && ea.kind != EK_CONV && ea.kind != EK_MV && ea.kind != EK_AA_SUPD) ok = ok && ea.partials;
    if (!ok) return fail(DDAFFT_ESTATE, "internal: matvec launched with a missing buffer");
    return DDAFFT_OK;
}

// The fused z/y-forward pass for the plane shapes it is built for (n_j, n_l
// in {8, 16, 32, 64}); false when the plane needs ZF + YF.
bool zyf_launch(ddafft_ctx* c, const Dims& d, const ProArgs& pa, SolveState* st, int C, hipStream_t s) {
    const dim3 grid(d.ni, C);
#define DDA_ZYF(NJ, NL)                                                                                        \
    case NJ * 1024 + NL:                                                                                       \
        hipLaunchKernelGGL((k_zyf<NJ, NL>), grid, (ZyfShape<NJ, NL>::BT), 0, s, d, pa, st, c->dipf, c->b2, c->wpre_l, c->tw_l, \
                           c->wpre_j, c->tw_j);                                                                \
        return true;
    switch (d.nj * 1024 + d.nl) {
        DDA_ZYF(8, 8) DDA_ZYF(8, 16) DDA_ZYF(8, 32) DDA_ZYF(8, 64)
        DDA_ZYF(16, 8) DDA_ZYF(16, 16) DDA_ZYF(16, 32) DDA_ZYF(16, 64)
        DDA_ZYF(32, 8) DDA_ZYF(32, 16) DDA_ZYF(32, 32) DDA_ZYF(32, 64)
        DDA_ZYF(64, 8) DDA_ZYF(64, 16) DDA_ZYF(64, 32) DDA_ZYF(64, 64)
        default:
            return false;
    }
#undef DDA_ZYF
}

// One matvec of nrhs right-hand sides: ZF -> YF (or the fused ZYF) -> X -> YI -> ZI.
int launch_matvec(ddafft_ctx* c, int nrhs, const ProArgs& pa, SolveState* st, int iter, const EpiArgs& ea,
                  int conj, hipEvent_t* ev = nullptr) {
    if (!ev && c->prof) {
        const size_t need = (size_t)6 * (c->prof_used + 1);
        while (c->prof_ev.size() < need) {
            hipEvent_t e;
            HIPCHECK(hipEventCreate(&e));
            c->prof_ev.push_back(e);
        }
        ev = &c->prof_ev[(size_t)6 * c->prof_used];
        if ((int)c->prof_iter.size() <= c->prof_used) c->prof_iter.resize(c->prof_used + 1);
        c->prof_iter[c->prof_used] = st ? iter : -1;
        c->prof_used++;
    }
    {
        const int rc = check_args(pa, ea, st != nullptr);
        if (rc) return rc;
        if (!c->occ || !c->mv || !c->b1 || !c->b2 || !c->b3 || !c->ahat || !c->tables)
            return fail(DDAFFT_ESTATE, "internal: context buffers missing");
    }
    TRY(flush_small(c));
    Dims d = c->d;
    d.nrhs = nrhs;
    const int C = 3 * nrhs;
    hipStream_t s = c->stream;
    const Schedule sch = schedule(c, nrhs);
    if (ev) HIPCHECK(hipEventRecord(ev[0], s));
    // z and y forward: fused per i-plane for planes up to 64 x 64, else ZF then YF
    if (sch.fuse && zyf_launch(c, d, pa, st, C, s)) {
        if (ev) HIPCHECK(hipEventRecord(ev[1], s));
    } else {
    {
        const int tj = std::min(tile_zf(d.nl), d.nj);
        dim3 grid(d.ni * (d.nj / tj), C);
        DDA_SWITCH_SMALL(d.nl, hipLaunchKernelGGL((k_zf<NN>), grid, kBlock, 0, s, d, pa, st, c->dipf, c->b1, c->wpre_l,
                                                                  c->tw_l));
    }
    if (ev) HIPCHECK(hipEventRecord(ev[1], s));
    {
        const int tl = std::min(tile_zf(d.nj), 2 * d.nl);   // L' lines per block
        dim3 grid(d.ni * (2 * d.nl / tl), C);
        DDA_SWITCH_SMALL(d.nj, hipLaunchKernelGGL((k_yf<NN>), grid, kBlock, 0, s, d, st, c->b1, c->b2, c->wpre_j, c->tw_j));
    }
    }
    if (ev) HIPCHECK(hipEventRecord(ev[2], s));
    {
        // right-hand sides go through the X pass in groups of rb that share
        // each Â value; a block takes `gpb` consecutive groups and reads its
        // tile's Â once for all of them, as many groups as keep about
        // x_blocks blocks in the launch
        int rb = sch.x_rb;
        while (rb > 1 && rb > nrhs) rb >>= 1;
        if (d.ni > 256) rb = 1;
        const int tj = std::min(tile_x(d.ni, rb), 2 * d.nl);   // L' columns per tile
        const int ntile = 2 * d.nj * (2 * d.nl / tj), ngroups = (nrhs + rb - 1) / rb;
        const int ysets = std::max(1, std::min(ngroups, (int)std::min<long long>(((long long)sch.x_blocks + ntile - 1) / ntile, 1 << 20)));
        const int gpb = (ngroups + ysets - 1) / ysets;
        dim3 grid(ntile, (ngroups + gpb - 1) / gpb);
#define DDA_KX(SW, RB, RUN, CP)                                                                           \
    SW(d.ni, hipLaunchKernelGGL((k_x<NN, RB, RUN, CP>), grid, kBlock, 0, s, d, st, c->b2, c->b3, c->ahat, conj, \
                                c->wpre_i, c->tw_i, c->tw_2i, gpb))
#define DDA_KX_RB(RUN, CP)                              \
    {                                                   \
        if (rb == 4) DDA_KX(DDA_SWITCH_X, 4, RUN, CP)      \
        else if (rb == 2) DDA_KX(DDA_SWITCH_X, 2, RUN, CP) \
        else DDA_KX(DDA_SWITCH_SMALL, 1, RUN, CP)          \
    }
        const bool compact = c->ahat_mode == 1;
        if (gpb > 1) {
            if (compact) DDA_KX_RB(true, true)
            else DDA_KX_RB(true, false)
        } else {
            if (compact) DDA_KX_RB(false, true)
            else DDA_KX_RB(false, false)
        }
#undef DDA_KX_RB
#undef DDA_KX
    }
    if (ev) HIPCHECK(hipEventRecord(ev[3], s));
    {
        const int tl = std::min(tile_yi(2 * d.nj), 2 * d.nl);
        dim3 grid(d.ni * (2 * d.nl / tl), C);
        DDA_SWITCH_LARGE(2 * d.nj, hipLaunchKernelGGL((k_yi<NN>), grid, kBlock, 0, s, d, st, c->b3, c->b1, c->tw_2j));
    }
    if (ev) HIPCHECK(hipEventRecord(ev[4], s));
    {
        dim3 grid(zi_blocks(d), nrhs);
        DDA_SWITCH_LARGE(2 * d.nl, hipLaunchKernelGGL((k_zi<NN>), grid, kBlock, 0, s, d, st, c->b1, ea, c->occ, c->dipf,
                                                                      c->tw_2l));
    }
    if (ev) HIPCHECK(hipEventRecord(ev[5], s));
    HIPCHECK(hipGetLastError());
    return DDAFFT_OK;
}

// Â into c->ahat: the full padded grid [2nj][2nl][2ni][6] (mode 0), or one
// octant [nj+1][nl+1][ni+1][6] of the transform of A with its Nyquist planes
// zeroed (mode 1), which the X pass mirrors.
int build_ahat(ddafft_ctx* c) {
    const int ni = c->d.ni, nj = c->d.nj, nl = c->d.nl;
    const bool compact = c->ahat_mode == 1;
    std::vector<float2> oct;
    green_octant(ni, nj, nl, c->kd, oct);
    float2 *doct = nullptr, *grid = nullptr, *tw = nullptr, *full = nullptr;
    const size_t P = (size_t)8 * ni * nj * nl;
    int rc = dev_alloc(&doct, oct.size());
    if (!rc) rc = dev_alloc(&grid, P * 6);
    if (!rc) rc = dev_alloc(&tw, (size_t)(ni + nj + nl));
    if (!rc && compact) rc = dev_alloc(&full, P * 6);
    if (rc) {
        hipFree(doct);
        hipFree(grid);
        hipFree(tw);
        hipFree(full);
        return rc;
    }
    float2* dst = compact ? full : c->ahat;
    std::vector<float2> wi = fourier_phase(ni), wj = fourier_phase(nj), wl = fourier_phase(nl);
    hipStream_t s = c->stream;
    hipError_t e = hipMemcpyAsync(doct, oct.data(), oct.size() * sizeof(float2), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(tw, wi.data(), ni * sizeof(float2), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(tw + ni, wj.data(), nj * sizeof(float2), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(tw + ni + nj, wl.data(), nl * sizeof(float2), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) {
        const long long total = (long long)P;
        const int blocks = (int)std::min<long long>((total + kBlock - 1) / kBlock, 65536);
        hipLaunchKernelGGL(k_ahat_expand, blocks, kBlock, 0, s, ni, nj, nl, doct, grid, compact ? 1 : 0);
        const float norma = (float)(1. / ((2 * ni) * (2 * nj) * (2 * nl)));
        auto axis_pass = [&](int axis, int na, const float2* w) -> int {
            int ext = axis == 0 ? 2 * nj : 2 * nl;
            int outer = axis == 2 ? 2 * nj : 2 * ni;
            int tt = 1;
            switch (na) {
                DDA_CASE_N(8, tt = std::min(clamp16(pow2floor(8192 / (6 * (NN + kLpad)))), ext))
                DDA_CASE_N(16, tt = std::min(clamp16(pow2floor(8192 / (6 * (NN + kLpad)))), ext))
                DDA_CASE_N(32, tt = std::min(clamp16(pow2floor(8192 / (6 * (NN + kLpad)))), ext))
                DDA_CASE_N(64, tt = std::min(clamp16(pow2floor(8192 / (6 * (NN + kLpad)))), ext))
                DDA_CASE_N(128, tt = std::min(clamp16(pow2floor(8192 / (6 * (NN + kLpad)))), ext))
                DDA_CASE_N(256, tt = std::min(clamp16(pow2floor(8192 / (6 * (NN + kLpad)))), ext))
                DDA_CASE_N(512, tt = std::min(clamp16(pow2floor(8192 / (6 * (NN + kLpad)))), ext))
                DDA_CASE_N(1024, tt = std::min(clamp16(pow2floor(8192 / (6 * (NN + kLpad)))), ext))
                default: return fail(DDAFFT_EINVAL, "unsupported axis length");
            }
            dim3 g(outer * (ext / tt));
            DDA_SWITCH_LARGE(na, hipLaunchKernelGGL((k_legacy_axis<NN>), g, kBlock, 0, s, ni, nj, nl, axis, grid, dst, w, norma));
            return DDAFFT_OK;
        };
        rc = axis_pass(0, 2 * nl, tw + ni + nj);
        if (!rc) rc = axis_pass(1, 2 * nj, tw + ni);
        if (!rc) rc = axis_pass(2, 2 * ni, tw);
        if (!rc && compact) {
            const long long noct = (long long)(nj + 1) * (nl + 1) * (ni + 1);
            const int ob = (int)std::min<long long>((noct + kBlock - 1) / kBlock, 65536);
            hipLaunchKernelGGL(k_ahat_octant, ob, kBlock, 0, s, ni, nj, nl, full, c->ahat);
        }
        e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(s);
    }
    hipFree(doct);
    hipFree(grid);
    hipFree(tw);
    hipFree(full);
    if (rc) return rc;
    if (e != hipSuccess) return fail(DDAFFT_EHIP, std::string("Â build: ") + hipGetErrorString(e));
    return DDAFFT_OK;
}

// Adds the pass durations of every matvec recorded since the last call (the
// stream must be idle).  Matvecs of solver iterations past `last_iter` were
// queued ahead of the stopping decision and exited at once; they are dropped.
int prof_collect(ddafft_ctx* c, int last_iter = 1 << 30) {
    for (int m = 0; m < c->prof_used; ++m) {
        hipEvent_t* e = &c->prof_ev[(size_t)6 * m];
        if (c->prof_iter[m] > last_iter) continue;
        for (int k = 0; k < 5; ++k) {
            float ms = 0.f;
            HIPCHECK(hipEventElapsedTime(&ms, e[k], e[k + 1]));
            c->prof_ms[k] += ms;
        }
        c->prof_n++;
    }
    c->prof_used = 0;
    return DDAFFT_OK;
}

// Host-to-device copy on the context's stream.  Packed into the pinned
// staging buffer while it has room (a pinned source keeps the copy
// asynchronous), otherwise from the caller's memory.
int upload(ddafft_ctx* c, void* dst, const void* src, size_t bytes) {
    char* d = static_cast<char*>(dst);
    if (c->small && d >= c->small && d + bytes <= c->small + c->small_bytes) {
        const size_t o = (size_t)(d - c->small);
        std::memcpy(c->small_h + o, src, bytes);
        if (c->dirty_hi <= c->dirty_lo) {
            c->dirty_lo = o;
            c->dirty_hi = o + bytes;
        } else {
            c->dirty_lo = std::min(c->dirty_lo, o);
            c->dirty_hi = std::max(c->dirty_hi, o + bytes);
        }
        return DDAFFT_OK;
    }
    const size_t need = (bytes + 255) & ~(size_t)255;
    if (c->up && c->up_off + need <= c->up_cap) {
        char* q = c->up + c->up_off;
        std::memcpy(q, src, bytes);
        c->up_off += need;
        HIPCHECK(hipMemcpyAsync(dst, q, bytes, hipMemcpyHostToDevice, c->stream));
        return DDAFFT_OK;
    }
    // unstaged: complete the copy before the caller may reuse its memory
    HIPCHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));
    return DDAFFT_OK;
}

// Copies what was staged into the small-input mirror since the last copy
// (one copy of the range written).  Called before every launch that may read
// the small inputs; a range is never restaged while its copy is pending
// (each solve stages once, after the stream has drained).
int flush_small(ddafft_ctx* c) {
    if (c->dirty_hi <= c->dirty_lo) return DDAFFT_OK;
    HIPCHECK(hipMemcpyAsync(c->small + c->dirty_lo, c->small_h + c->dirty_lo, c->dirty_hi - c->dirty_lo,
                            hipMemcpyHostToDevice, c->stream));
    c->dirty_lo = c->dirty_hi = 0;
    return DDAFFT_OK;
}

// (Re)allocates pinned host memory of at least `bytes`; waits for the stream
// first, since pending copies may use the old buffer.  dev (optional)
// receives the device address of the buffer.
int pinned_reserve(ddafft_ctx* c, char** p, size_t* cap, size_t bytes, char** dev = nullptr) {
    if (*p && *cap >= bytes) return DDAFFT_OK;
    if (t_noalloc) return fail(DDAFFT_ESTATE, "internal: pinned allocation inside the allocation-free solve");
    HIPCHECK(hipStreamSynchronize(c->stream));
    if (*p) hipHostFree(*p);
    *p = nullptr;
    *cap = 0;
    const size_t n = std::max<size_t>(bytes, 1 << 20);
    if (hipHostMalloc((void**)p, n, hipHostMallocDefault) != hipSuccess) {
        *p = nullptr;
        return fail(DDAFFT_ENOMEM, "hipHostMalloc of pinned staging failed");
    }
    *cap = n;
    if (dev && hipHostGetDevicePointer((void**)dev, *p, 0) != hipSuccess) {
        hipHostFree(*p);
        *p = nullptr;
        *cap = 0;
        return fail(DDAFFT_EHIP, "hipHostGetDevicePointer of pinned staging failed");
    }
    return DDAFFT_OK;
}

// Carves the small per-solve inputs out of one device block and allocates
// its page-locked mirror.
int small_layout(ddafft_ctx* c) {
    const int R = c->nrhs_max;
    size_t off = 0;
    auto take = [&](size_t b) {
        const size_t o = off;
        off += (b + 255) & ~(size_t)255;
        return o;
    };
    const size_t o_st = take((size_t)R * sizeof(SolveState));
    const size_t o_al = take((size_t)R * 128 * sizeof(float2));   // set_target allows 128 materials
    const size_t o_ei = take((size_t)R * 3 * sizeof(float2));
    const size_t o_ki = take((size_t)R * 3 * sizeof(float));
    const size_t o_ax = take((size_t)R * sizeof(int));
    const size_t o_tb = take((size_t)R * 2 * inc_tab_len(c) * sizeof(float2));
    TRY(dev_alloc(&c->small, off));
    if (hipHostMalloc((void**)&c->small_h, off, hipHostMallocDefault) != hipSuccess) {
        c->small_h = nullptr;
        return fail(DDAFFT_ENOMEM, "hipHostMalloc of the input mirror failed");
    }
    c->small_bytes = off;
    c->st = reinterpret_cast<SolveState*>(c->small + o_st);
    c->alpha = reinterpret_cast<float2*>(c->small + o_al);
    c->einc = reinterpret_cast<float2*>(c->small + o_ei);
    c->kinc = reinterpret_cast<float*>(c->small + o_ki);
    c->inc_axis = reinterpret_cast<int*>(c->small + o_ax);
    c->inc_tab = reinterpret_cast<float2*>(c->small + o_tb);
    return DDAFFT_OK;
}

// Starts the uploads of one solve: the stream is idle then (every solve ends
// with a synchronization), so the staging buffer is free to reuse.
int upload_begin(ddafft_ctx* c, size_t bytes) {
    HIPCHECK(hipStreamSynchronize(c->stream));
    c->up_off = 0;
    // allocation-free: stage only into what exists (upload() copies the rest
    // from the caller's memory and waits)
    if (t_noalloc) return DDAFFT_OK;
    return pinned_reserve(c, &c->up, &c->up_cap, bytes);
}

int field_in(ddafft_ctx* c, const float* host_aos, float2* dev_soa, float2* staging) {
    const long long n = c->d.n;
    HIPCHECK(hipMemcpyAsync(staging, host_aos, (size_t)n * 3 * sizeof(float2), hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_aos_to_soa, (int)std::min<long long>((n + kBlock - 1) / kBlock, 4096), kBlock, 0, c->stream, n, staging,
                                                                                                     dev_soa);
    HIPCHECK(hipGetLastError());
    return DDAFFT_OK;
}

int field_out(ddafft_ctx* c, const float2* dev_soa, float* host_aos, float2* staging) {
    const long long n = c->d.n;
    hipLaunchKernelGGL(k_soa_to_aos, (int)std::min<long long>((n + kBlock - 1) / kBlock, 4096), kBlock, 0, c->stream, n, dev_soa,
                                                                                                     staging);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipMemcpyAsync(host_aos, staging, (size_t)n * 3 * sizeof(float2), hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));
    return DDAFFT_OK;
}

Geometry make_geometry(const ddafft_ctx* c, const float* kinc) {
    Geometry g;
    const float d0 = c->kd;
    g.d0 = d0;
    g.kx = kinc ? kinc[0] : 0.f;
    g.ky = kinc ? kinc[1] : 0.f;
    g.kz = kinc ? kinc[2] : 1.f;
    g.cx = (float)((c->d.ni - 1) * d0 / 2.);   // dda_func.cpp:659, 1581
    g.cy = (float)((c->d.nj - 1) * d0 / 2.);
    g.cz = (float)((c->d.nl - 1) * d0 / 2.);
    g.ex = (float)(c->d.ni * d0 / 2.);         // dda_func.cpp:1685
    g.ey = (float)(c->d.nj * d0 / 2.);
    g.ez = (float)(c->d.nl * d0 / 2.);
    return g;
}

// Uploads the incidence of nrhs right-hand sides (kinc [nrhs][3]) and, for
// each axis-aligned one, the tables of its FirstApproximation and CabsCext
// phases evaluated as the reference evaluates them: the same float argument
// (dda_func.cpp:663-675, 1694-1700) and the C library's cosf/sinf.
int inc_tab_len(const ddafft_ctx* c) { return std::max(c->d.ni, std::max(c->d.nj, c->d.nl)); }

IncPhase inc_phase(const ddafft_ctx* c) { return IncPhase{c->inc_axis, c->inc_tab, inc_tab_len(c), c->inc_cell}; }

// Per-cell phases of right-hand side r (exact_phases): the reference's float
// arguments (dda_func.cpp:663-675 for FirstApproximation, 1694-1700 for
// CabsCext) through the C library's cosf/sinf, on the dipole cells.
void cell_phases(const ddafft_ctx* c, const Geometry& g, const float* k, float2* t1, float2* t2) {
    const int ni = c->d.ni, nj = c->d.nj, nl = c->d.nl;
    const float kx = k[0], ky = k[1], kz = k[2];
    size_t q = 0;
    for (int i = 0; i < ni; ++i)
        for (int j = 0; j < nj; ++j)
            for (int l = 0; l < nl; ++l, ++q) {
                if (!c->dipf_h[q]) continue;
                const float rx = (float)i * g.d0, ry = (float)j * g.d0, rz = (float)l * g.d0;
                const float dx = rx - g.cx, dy = ry - g.cy, dz = rz - g.cz;
                const float a1 = kx * dx + ky * dy + kz * dz;
                const float a2 = -(kx * (rx - g.ex) + ky * (ry - g.ey) + kz * (rz - g.ez));
                t1[q] = make_float2(std::cos(a1), std::sin(a1));
                t2[q] = make_float2(std::cos(a2), std::sin(a2));
            }
}

int upload_incidence(ddafft_ctx* c, int nrhs, const float* kinc, bool exact = false) {
    if (!c->up) HIPCHECK(hipStreamSynchronize(c->stream));   // unstaged copies read the host vectors below
    const Geometry g = make_geometry(c, nullptr);
    c->inc_axis_h.assign(nrhs, -1);
    const int tn = inc_tab_len(c);
    c->inc_tab_h.assign((size_t)nrhs * 2 * tn, make_float2(1.f, 0.f));
    const int n[3] = {c->d.ni, c->d.nj, c->d.nl};
    const float cen1[3] = {g.cx, g.cy, g.cz}, cen2[3] = {g.ex, g.ey, g.ez};
    const long long ncell = c->d.n;
    bool cells = false;
    for (int r = 0; r < nrhs; ++r) {
        const float* k = kinc + 3 * r;
        const int nz = (k[0] != 0.f) + (k[1] != 0.f) + (k[2] != 0.f);
        if (nz != 1) {
            if (!exact) continue;
            if (!cells) {
                c->inc_cell_h.assign((size_t)nrhs * 2 * ncell, make_float2(0.f, 0.f));
                cells = true;
            }
            c->inc_axis_h[r] = -2;
            cell_phases(c, g, k, &c->inc_cell_h[(size_t)r * 2 * ncell], &c->inc_cell_h[((size_t)r * 2 + 1) * ncell]);
            continue;
        }
        const int a = k[0] != 0.f ? 0 : (k[1] != 0.f ? 1 : 2);
        c->inc_axis_h[r] = a;
        float2* t1 = &c->inc_tab_h[(size_t)r * 2 * tn];
        float2* t2 = t1 + tn;
        for (int q = 0; q < n[a]; ++q) {
            const float rq = (float)q * g.d0;
            const float arg1 = k[a] * (rq - cen1[a]);
            const float arg2 = -(k[a] * (rq - cen2[a]));
            t1[q] = make_float2(std::cos(arg1), std::sin(arg1));
            t2[q] = make_float2(std::cos(arg2), std::sin(arg2));
        }
    }
    TRY(upload(c, c->kinc, kinc, (size_t)nrhs * 3 * sizeof(float)));
    TRY(upload(c, c->inc_axis, c->inc_axis_h.data(), (size_t)nrhs * sizeof(int)));
    TRY(upload(c, c->inc_tab, c->inc_tab_h.data(), c->inc_tab_h.size() * sizeof(float2)));
    if (cells) {
        if (c->inc_cell_cap < c->inc_cell_h.size()) {
            HIPCHECK(hipStreamSynchronize(c->stream));
            if (c->inc_cell) hipFree(c->inc_cell);
            c->inc_cell = nullptr;
            c->inc_cell_cap = 0;
            TRY(dev_alloc(&c->inc_cell, c->inc_cell_h.size()));
            c->inc_cell_cap = c->inc_cell_h.size();
        }
        TRY(upload(c, c->inc_cell, c->inc_cell_h.data(), c->inc_cell_h.size() * sizeof(float2)));
    }
    return DDAFFT_OK;
}

// 4 pi / Re(E . conj(E)) rounded to float (dda_func.cpp:1683)
float cabs_prefactor(const float* e) {
    auto cm = [](float ar, float ai, float br, float bi, float& cr, float& ci) {
        cr = ar * br - ai * bi;
        ci = ar * bi + ai * br;
    };
    float xr, xi, yr, yi, zr, zi;
    cm(e[0], e[1], e[0], -e[1], xr, xi);
    cm(e[2], e[3], e[2], -e[3], yr, yi);
    cm(e[4], e[5], e[4], -e[5], zr, zi);
    const float re = (xr + yr) + zr;
    return (float)((4. * kPi) / re);
}

int update_blocks(const Dims& d) { return (int)std::min<long long>((d.n + kBlock - 1) / kBlock, 1024); }

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

extern "C" {

const char* ddafft_last_error(void) { return g_err.c_str(); }

int ddafft_version(void) { return 100; }

int ddafft_host_alloc(size_t bytes, void** out) {
    if (!out || bytes == 0) return fail(DDAFFT_EINVAL, "bad argument");
    *out = nullptr;
    if (hipHostMalloc(out, bytes, hipHostMallocDefault) != hipSuccess) {
        *out = nullptr;
        return fail(DDAFFT_ENOMEM, "hipHostMalloc failed");
    }
    return DDAFFT_OK;
}

int ddafft_host_free(void* p) {
    if (p) HIPCHECK(hipHostFree(p));
    return DDAFFT_OK;
}

int ddafft_destroy(ddafft_ctx* c) {
    if (!c) return DDAFFT_OK;
    if (c->stream) hipStreamSynchronize(c->stream);
    void* bufs[] = {c->ahat, c->tables, c->occ, c->dipf, c->mv, c->cells, c->b1, c->b2, c->b3, c->P, c->Pmin, c->R,
                    c->S, c->G, c->H, c->W, c->X, c->Y, c->small, c->hist, c->partials,
                    c->dirs, c->esc, c->fftab, c->Pf, c->phase, c->gaux, c->inc_cell, c->seqbuf,
                    c->ffseg, c->ffpc};
    for (void* b : bufs)
        if (b) hipFree(b);
    if (c->state_ring) hipHostFree(c->state_ring);
    if (c->small_h) hipHostFree(c->small_h);
    if (c->up) hipHostFree(c->up);
    if (c->dn) hipHostFree(c->dn);
    for (hipEvent_t e : c->prof_ev) hipEventDestroy(e);
    for (hipEvent_t& e : c->ev)
        if (e) hipEventDestroy(e);
    if (c->fin_ev) hipEventDestroy(c->fin_ev);
    if (c->gexec) hipGraphExecDestroy(c->gexec);
    if (c->graph) hipGraphDestroy(c->graph);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
    return DDAFFT_OK;
}

int ddafft_create(int ni, int nj, int nl, float kd, int nrhs_max, ddafft_ctx** out) {
    if (!out) return fail(DDAFFT_EINVAL, "out is NULL");
    *out = nullptr;
    if (!pow2_ok(ni) || !pow2_ok(nj) || !pow2_ok(nl))
        return fail(DDAFFT_EINVAL, "lattice dimensions must be powers of two in [4, 512]");
    if ((long long)ni * nj * nl > kMaxCells)
        return fail(DDAFFT_EINVAL, "lattice of more than 2^24 cells (e.g. 256^3 or 512x256x128)");
    if (!(kd > 0.f) || nrhs_max < 1 || nrhs_max > 64) return fail(DDAFFT_EINVAL, "bad kd or nrhs_max");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(DDAFFT_ENODEV, "no HIP device");
    ddafft_ctx* c = new ddafft_ctx();
    c->d.ni = ni;
    c->d.nj = nj;
    c->d.nl = nl;
    c->d.nrhs = nrhs_max;
    c->d.n = (long long)ni * nj * nl;
    c->nrhs_max = nrhs_max;
    c->kd = kd;
    int rc = DDAFFT_OK;
    auto chk = [&](hipError_t e, const char* what) {
        if (e != hipSuccess && rc == DDAFFT_OK) rc = fail(DDAFFT_EHIP, std::string(what) + ": " + hipGetErrorString(e));
    };
    chk(hipGetDevice(&c->device), "hipGetDevice");
    if (rc == DDAFFT_OK) c->scratch = scratch_address(c->device);   // set-up of the allocation-free solve
    chk(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate");
    for (hipEvent_t& e : c->ev) chk(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
    chk(hipEventCreateWithFlags(&c->fin_ev, hipEventDisableTiming), "hipEventCreate");
    if (rc) {
        ddafft_destroy(c);
        return rc;
    }
    const long long n = c->d.n, P = 8 * n;
    const long long C = 3LL * nrhs_max;
    // tables: wpre(ni,nj,nl), tw(ni,nj,nl), tw(2ni,2nj,2nl)
    const size_t tab = (size_t)4 * (ni + nj + nl);
    rc = dev_alloc(&c->tables, tab);
    if (!rc) rc = dev_alloc(&c->ahat, (size_t)P * 6);
    if (!rc) rc = dev_alloc(&c->b1, (size_t)(2 * n * C));
    if (!rc) rc = dev_alloc(&c->b2, (size_t)(4 * n * C));
    if (!rc) rc = dev_alloc(&c->b3, (size_t)(4 * n * C));
    float2** fields[] = {&c->P, &c->Pmin, &c->R, &c->S, &c->G, &c->H, &c->W, &c->X, &c->Y};
    for (float2** f : fields)
        if (!rc) rc = dev_alloc(f, (size_t)(n * C));
    if (!rc) rc = small_layout(c);
    if (!rc) rc = dev_alloc(&c->occ, (size_t)n);
    if (!rc) rc = dev_alloc(&c->dipf, (size_t)n);
    if (!rc && hipHostMalloc((void**)&c->state_ring, (size_t)4 * nrhs_max * sizeof(SolveState),
                             hipHostMallocDefault) != hipSuccess)
        rc = fail(DDAFFT_ENOMEM, "hipHostMalloc");
    if (!rc && hipHostGetDevicePointer((void**)&c->ring_dev, c->state_ring, 0) != hipSuccess)
        rc = fail(DDAFFT_EHIP, "hipHostGetDevicePointer of the state ring");
    if (!rc) rc = ensure_partials(c, (size_t)nrhs_max * std::max<long long>(zi_blocks(c->d), 1024) * kSlots);
    if (rc) {
        ddafft_destroy(c);
        return rc;
    }
    {
        std::vector<float2> t;
        auto app = [&](const std::vector<float2>& v) { t.insert(t.end(), v.begin(), v.end()); };
        app(fourier_phase(ni));
        app(fourier_phase(nj));
        app(fourier_phase(nl));
        app(plan_twiddles(ni));
        app(plan_twiddles(nj));
        app(plan_twiddles(nl));
        app(plan_twiddles(2 * ni));
        app(plan_twiddles(2 * nj));
        app(plan_twiddles(2 * nl));
        const float2* b = c->tables;
        c->wpre_i = b;
        c->wpre_j = b + ni;
        c->wpre_l = b + ni + nj;
        c->tw_i = b + ni + nj + nl;
        c->tw_j = c->tw_i + ni;
        c->tw_l = c->tw_j + nj;
        c->tw_2i = c->tw_l + nl;
        c->tw_2j = c->tw_2i + 2 * ni;
        c->tw_2l = c->tw_2j + 2 * nj;
        hipError_t e = hipMemcpy(c->tables, t.data(), t.size() * sizeof(float2), hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            ddafft_destroy(c);
            return fail(DDAFFT_EHIP, std::string("table upload: ") + hipGetErrorString(e));
        }
    }
    rc = build_ahat(c);
    if (rc) {
        ddafft_destroy(c);
        return rc;
    }
    *out = c;
    return DDAFFT_OK;
}

int ddafft_get_ahat(ddafft_ctx* c, float* out) {
    if (!c || !out) return fail(DDAFFT_EINVAL, "NULL argument");
    const int ni = c->d.ni, nj = c->d.nj, nl = c->d.nl;
    const int NI2 = 2 * ni, NJ2 = 2 * nj, NL2 = 2 * nl;
    const size_t P = (size_t)NI2 * NJ2 * NL2;
    const bool compact = c->ahat_mode == 1;
    std::vector<float2> h((compact ? (size_t)(ni + 1) * (nj + 1) * (nl + 1) : P) * 6);
    HIPCHECK(hipMemcpyAsync(h.data(), c->ahat, h.size() * sizeof(float2), hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));
    float2* o = reinterpret_cast<float2*>(out);
    for (int J = 0; J < NJ2; ++J)
        for (int L = 0; L < NL2; ++L)
            for (int I = 0; I < NI2; ++I) {
                float2* dst = &o[(((size_t)I * NJ2 + J) * NL2 + L) * 6];
                if (!compact) {
                    memcpy(dst, &h[(((size_t)J * NL2 + L) * NI2 + I) * 6], 6 * sizeof(float2));
                    continue;
                }
                // the grid the X pass reads: the octant mirrored, odd components negated
                const bool fx = I > ni, fy = J > nj, fz = L > nl;
                const int Ic = fx ? NI2 - I : I, Jc = fy ? NJ2 - J : J, Lc = fz ? NL2 - L : L;
                memcpy(dst, &h[(((size_t)Jc * (nl + 1) + Lc) * (ni + 1) + Ic) * 6], 6 * sizeof(float2));
                auto neg = [](float2& v) { v.x = -v.x; v.y = -v.y; };
                if (fx != fy) neg(dst[1]);
                if (fx != fz) neg(dst[2]);
                if (fy != fz) neg(dst[4]);
            }
    return DDAFFT_OK;
}

int ddafft_set_target(ddafft_ctx* c, const signed char* occ, int nmat, const int* mv) {
    if (!c || !occ || !mv || nmat < 1 || nmat > 128) return fail(DDAFFT_EINVAL, "bad target arguments");
    const long long n = c->d.n;
    std::vector<int> cells;
    cells.reserve((size_t)n);
    for (long long k = 0; k < n; ++k) {
        const int m = occ[k];
        if (m < 0 || m >= nmat) return fail(DDAFFT_EINVAL, "occupancy index outside the material table");
        if (mv[m]) cells.push_back((int)k);
    }
    if (cells.empty()) return fail(DDAFFT_EINVAL, "target has no dipoles");
    // far-field segments: each run of dipoles in one (i, j) column, cut into
    // near-equal pieces of at most kFFSeg
    std::vector<int2> segs;
    {
        const int lnl = __builtin_ctz(c->d.nl);
        size_t e = 0;
        while (e < cells.size()) {
            const int ij = cells[e] >> lnl;
            size_t e1 = e + 1;
            while (e1 < cells.size() && (cells[e1] >> lnl) == ij) ++e1;
            const int len = (int)(e1 - e), np = (len + kFFSeg - 1) / kFFSeg;
            for (int q = 0, at = 0; q < np; ++q) {
                const int cnt = (len - at) / (np - q);
                segs.push_back(make_int2((int)e + at, (ij << 6) | cnt));
                at += cnt;
            }
            e = e1;
        }
    }
    HIPCHECK(hipStreamSynchronize(c->stream));
    if (c->mv) hipFree(c->mv);
    if (c->cells) hipFree(c->cells);
    if (c->ffseg) hipFree(c->ffseg);
    c->mv = nullptr;
    c->cells = nullptr;
    c->ffseg = nullptr;
    c->nseg = 0;
    int rc = dev_alloc(&c->mv, (size_t)nmat);
    if (!rc) rc = dev_alloc(&c->cells, cells.size());
    if (!rc) rc = dev_alloc(&c->ffseg, segs.size());
    if (rc) return rc;
    c->nmat = nmat;
    c->mat_count.assign(nmat, 0);
    for (long long k = 0; k < n; ++k) c->mat_count[occ[k]] += mv[occ[k]] ? 1 : 0;
    c->ncells = (int)cells.size();
    c->ndip = (long long)cells.size();
    if ((rc = upload(c, c->occ, occ, (size_t)n))) return rc;
    c->dipf_h.assign((size_t)n, 0);
    for (long long k = 0; k < n; ++k) c->dipf_h[k] = mv[occ[k]] ? 1 : 0;
    if ((rc = upload(c, c->dipf, c->dipf_h.data(), (size_t)n))) return rc;
    if ((rc = upload(c, c->mv, mv, (size_t)nmat * sizeof(int)))) return rc;
    if ((rc = upload(c, c->cells, cells.data(), cells.size() * sizeof(int)))) return rc;
    if ((rc = upload(c, c->ffseg, segs.data(), segs.size() * sizeof(int2)))) return rc;
    c->nseg = (int)segs.size();
    HIPCHECK(hipStreamSynchronize(c->stream));
    c->have_target = true;
    return DDAFFT_OK;
}

int ddafft_set_tuning(ddafft_ctx* c, int key, int value) {
    if (!c) return fail(DDAFFT_EINVAL, "NULL context");
    switch (key) {
        case DDAFFT_TUNE_X_BLOCKS:
            if (value < 0) return fail(DDAFFT_EINVAL, "x_blocks must be 0 (tuned) or positive");
            c->x_blocks = value;
            return DDAFFT_OK;
        case DDAFFT_TUNE_X_RB:
            if (value != 0 && value != 1 && value != 2 && value != 4)
                return fail(DDAFFT_EINVAL, "x_rb must be 0 (automatic), 1, 2 or 4");
            c->x_rb = value;
            return DDAFFT_OK;
        case DDAFFT_TUNE_FUSE_ZY:
            if (value < -1 || value > 1) return fail(DDAFFT_EINVAL, "fuse_zy must be -1 (tuned), 0 or 1");
            c->fuse_zy = value;
            return DDAFFT_OK;
        case DDAFFT_TUNE_FF_MFMA:
            if (value < -1 || value > 1) return fail(DDAFFT_EINVAL, "ff_mfma must be -1 (default), 0 or 1");
            c->ff_mfma = value;
            return DDAFFT_OK;
        case DDAFFT_TUNE_FF_GROUP:
            if (value < -1 || value > 1) return fail(DDAFFT_EINVAL, "ff_group must be -1 (default), 0 or 1");
            c->ff_group = value;
            return DDAFFT_OK;
        case DDAFFT_TUNE_GRAPH:
            if (value < -1 || value > 1) return fail(DDAFFT_EINVAL, "graph must be -1 (tuned), 0 or 1");
            c->use_graph = value;
            return DDAFFT_OK;
        default:
            return fail(DDAFFT_EINVAL, "unknown tuning key");
    }
}

int ddafft_set_ahat_mode(ddafft_ctx* c, int mode) {
    if (!c || mode < 0 || mode > 1) return fail(DDAFFT_EINVAL, "A-hat mode must be 0 or 1");
    if (mode == c->ahat_mode) return DDAFFT_OK;
    HIPCHECK(hipStreamSynchronize(c->stream));   // a captured iteration is keyed by the mode and is captured again
    if (c->ahat) HIPCHECK(hipFree(c->ahat));
    c->ahat = nullptr;
    c->ahat_mode = mode;
    const size_t cells = mode == 1 ? (size_t)(c->d.ni + 1) * (c->d.nj + 1) * (c->d.nl + 1) : (size_t)8 * c->d.n;
    TRY(dev_alloc(&c->ahat, cells * 6));
    return build_ahat(c);
}

int ddafft_set_sum_mode(ddafft_ctx* c, int mode) {
    if (!c || mode < 0 || mode > 1) return fail(DDAFFT_EINVAL, "sum mode must be 0 or 1");
    c->sum_mode = mode;
    return DDAFFT_OK;
}

namespace {

// A set of ng * S directions whose z components repeat with period ng
// (directions d and d + ng share z bit for bit) splits into ng groups of S
// (group g: g, g + ng, ..., g + (S - 1) ng).  Returns ng of the smallest
// such period with 2 <= S <= kFFDirs (S in *S), else 0.
int dir_groups(const float* dirs, int ndirs, int* S) {
    *S = 0;
    if (!dirs || ndirs < 2) return 0;
    for (int ng = std::max(1, (ndirs + kFFDirs - 1) / kFFDirs); ng <= ndirs / 2; ++ng) {
        if (ndirs % ng) continue;
        bool ok = true;
        for (int dd = ng; dd < ndirs && ok; ++dd)
            ok = std::memcmp(&dirs[3 * dd + 2], &dirs[3 * (dd - ng) + 2], sizeof(float)) == 0;
        if (ok) {
            *S = ndirs / ng;
            return ng;
        }
    }
    return 0;
}

}  // namespace

int ddafft_set_directions(ddafft_ctx* c, int ndirs, const float* dirs) {
    if (!c || ndirs < 0 || (ndirs > 0 && !dirs)) return fail(DDAFFT_EINVAL, "bad directions");
    HIPCHECK(hipStreamSynchronize(c->stream));
    if (ndirs > c->dirs_cap) {
        if (c->dirs) hipFree(c->dirs);
        c->dirs = nullptr;
        c->dirs_cap = 0;
        int rc = dev_alloc(&c->dirs, (size_t)ndirs * 3);
        if (rc) return rc;
        c->dirs_cap = ndirs;
    }
    c->ndirs = ndirs;
    c->ffg_n = dir_groups(dirs, ndirs, &c->ffg_S);
    if (ndirs) {
        HIPCHECK(hipMemcpy(c->dirs, dirs, (size_t)ndirs * 3 * sizeof(float), hipMemcpyHostToDevice));
    }
    return DDAFFT_OK;
}

int ddafft_convolution(ddafft_ctx* c, const float* x, float* y, int conj) {
    if (!c || !x || !y) return fail(DDAFFT_EINVAL, "NULL argument");
    if (!c->have_target) return fail(DDAFFT_ESTATE, "set_target first");
    int rc = field_in(c, x, c->X, c->Y);
    if (rc) return rc;
    ProArgs pa{PRO_COPY, c->X, nullptr, nullptr, nullptr, nullptr};
    EpiArgs ea{};
    ea.kind = EK_CONV;
    ea.out = c->W;
    ea.alpha = c->alpha;
    ea.nmat = c->nmat;
    rc = launch_matvec(c, 1, pa, nullptr, 0, ea, conj);
    if (rc) return rc;
    return field_out(c, c->W, y, c->Y);
}

int ddafft_matvec(ddafft_ctx* c, const float* alpha, const float* x, float* y, int conj) {
    if (!c || !alpha || !x || !y) return fail(DDAFFT_EINVAL, "NULL argument");
    if (!c->have_target) return fail(DDAFFT_ESTATE, "set_target first");
    int rc = upload(c, c->alpha, alpha, (size_t)c->nmat * sizeof(float2));
    if (!rc) rc = field_in(c, x, c->X, c->Y);
    if (rc) return rc;
    ProArgs pa{PRO_COPY, c->X, nullptr, nullptr, nullptr, nullptr};
    EpiArgs ea{};
    ea.kind = EK_MV;
    ea.conj = conj;
    ea.x = c->X;
    ea.out = c->W;
    ea.alpha = c->alpha;
    ea.nmat = c->nmat;
    rc = launch_matvec(c, 1, pa, nullptr, 0, ea, conj);
    if (rc) return rc;
    return field_out(c, c->W, y, c->Y);
}

// --- the solve ---------------------------------------------------------------

namespace {

struct Seq {
    ddafft_ctx* c;
    int nrhs;
    FinArgs fa;

    // the scalar step `kind` over nblk block partials per right-hand side
    int fin(int kind, int iter, int nblk) {
        FinArgs f = fa;
        f.kind = kind;
        f.nblk = nblk;
        f.partials = c->partials;
        f.hist = c->hist;
        // the step that closes an iteration publishes the states to the ring
        // (slot: the iteration & 3, read on the device)
        f.ring = ends_step(kind) ? c->ring_dev : nullptr;
        f.ring_stride = c->nrhs_max;
        if (!f.partials || !f.hist || (ends_step(kind) && !c->ring_dev))
            return fail(DDAFFT_ESTATE, "solver scratch not allocated");
        hipLaunchKernelGGL(k_finalize, nrhs, kBlock, 0, c->stream, c->st, f);
        HIPCHECK(hipGetLastError());
        return DDAFFT_OK;
    }
    // a matvec, then (fin_kind >= 0) the scalar step over its Z-inverse partials
    int mv(int iter, ProArgs pa, EpiArgs ea, int conj, int fin_kind = -1) {
        ea.alpha = c->alpha;
        ea.nmat = c->nmat;
        ea.partials = c->partials;
        ea.conj = conj;
        TRY(launch_matvec(c, nrhs, pa, c->st, iter, ea, conj));
        return fin_kind >= 0 ? fin(fin_kind, iter, zi_blocks(c->d)) : DDAFFT_OK;
    }
    // the solver update, then (fin_kind >= 0) the scalar step over its partials
    int upd(int iter, UpdArgs ua, int fin_kind = -1) {
        Dims d = c->d;
        d.nrhs = nrhs;
        ua.partials = c->partials;
        if (!ua.P || !ua.R || !ua.S || !ua.W || !ua.partials || (ua.kind == UK_BICG && !ua.G))
            return fail(DDAFFT_ESTATE, "internal: update launched with a missing buffer");
        dim3 g(update_blocks(d), nrhs);
        hipLaunchKernelGGL(k_update, g, kBlock, 0, c->stream, d, c->st, ua, c->dipf);
        HIPCHECK(hipGetLastError());
        return fin_kind >= 0 ? fin(fin_kind, iter, update_blocks(d)) : DDAFFT_OK;
    }
};

ProArgs pro(int mode, float2* A, float2* B = nullptr, const float2* D = nullptr) {
    return ProArgs{mode, A, B, D, nullptr, nullptr};
}

EpiArgs epi(int kind, const float2* x, float2* out, float2* out2 = nullptr, float2* out3 = nullptr,
            const float2* y = nullptr, float tau = 0.f) {
    EpiArgs e{};
    e.kind = kind;
    e.x = x;
    e.out = out;
    e.out2 = out2;
    e.out3 = out3;
    e.y = y;
    e.tau = tau;
    return e;
}

int step0(Seq& q, int solver) {
    ddafft_ctx* c = q.c;
    const int zb = zi_blocks(c->d);
    const int ub = update_blocks(c->d);
    switch (solver) {
        case DDAFFT_SOLVER_CG:
            TRY(q.mv(0, pro(PRO_COPY, c->P), epi(EK_CG_R0, nullptr, c->R, c->S), 0, FK_CG_E0));
            break;
        case DDAFFT_SOLVER_AA:
            TRY(q.mv(0, pro(PRO_COPY, c->P), epi(EK_AA_R0, nullptr, c->R), 0, FK_AA_0));
            TRY(q.mv(0, pro(PRO_COPY, c->R), epi(EK_AA_S0, c->R, c->S), 1, FK_AA_E0));
            break;
        case DDAFFT_SOLVER_BICGSTAB:
            TRY(q.mv(0, pro(PRO_COPY, c->P), epi(EK_BICG_R0, nullptr, c->R, c->S, c->G), 0, FK_BICG_0));
            TRY(q.mv(0, pro(PRO_COPY, c->S), epi(EK_BICG_H, c->S, c->H, nullptr, nullptr, c->G), 0, FK_BICG_A));
            TRY(q.mv(0, pro(PRO_LC1, c->R, c->H), epi(EK_BICG_W, c->R, c->W), 0, FK_BICG_W));
            TRY(q.upd(0, UpdArgs{UK_BICG, c->P, c->R, c->S, c->W, c->G, nullptr}, FK_BICG_E));
            break;
        case DDAFFT_SOLVER_SIMPLE:
            TRY(q.mv(0, pro(PRO_COPY, c->P), epi(EK_SIM_R0, c->P, c->R, c->S, c->P), 0, FK_SIM_E));
            break;
    }
    return DDAFFT_OK;
}

// part 0: the whole iteration; BiCGStab also in two parts, 1 (the first
// matvec and its scalar step) and 2 (the rest)
int iteration(Seq& q, int solver, int o, float tau, int part = 0) {
    ddafft_ctx* c = q.c;
    const int zb = zi_blocks(c->d);
    const int ub = update_blocks(c->d);
    auto snap = [&](ProArgs p) {
        p.P = c->P;
        p.Pmin = c->Pmin;
        return p;
    };
    switch (solver) {
        case DDAFFT_SOLVER_CG: {
            ProArgs p = (o == 1) ? pro(PRO_COPY, c->S) : pro(PRO_LC2, c->R, c->S);
            TRY(q.mv(o, snap(p), epi(EK_CG_W, c->S, c->W, nullptr, nullptr, c->R), 0, FK_CG_A));
            TRY(q.upd(o, UpdArgs{UK_CG, c->P, c->R, c->S, c->W, nullptr, nullptr}, FK_CG_E));
            break;
        }
        case DDAFFT_SOLVER_AA:
            TRY(q.mv(o, snap(pro(PRO_COPY, c->S)), epi(EK_AA_W, c->S, c->W), 0, FK_AA_A));
            TRY(q.upd(o, UpdArgs{UK_AA, c->P, c->R, c->S, c->W, nullptr, nullptr}, FK_AA_E));
            TRY(q.mv(o, pro(PRO_COPY, c->R), epi(EK_AA_SUPD, c->R, nullptr, c->S), 1));
            break;
        case DDAFFT_SOLVER_BICGSTAB:
            if (part != 2)
                TRY(q.mv(o, snap(pro(PRO_LC4, c->R, c->S, c->H)), epi(EK_BICG_H, c->S, c->H, nullptr, nullptr, c->G), 0,
                         FK_BICG_A));
            if (part != 1) {
                TRY(q.mv(o, pro(PRO_LC1, c->R, c->H), epi(EK_BICG_W, c->R, c->W), 0, FK_BICG_W));
                TRY(q.upd(o, UpdArgs{UK_BICG, c->P, c->R, c->S, c->W, c->G, nullptr}, FK_BICG_E));
            }
            break;
        case DDAFFT_SOLVER_SIMPLE:
            TRY(q.mv(o, snap(pro(PRO_COPY, c->P)), epi(EK_SIM, c->P, c->R, c->P, nullptr, c->S, tau), 0, FK_SIM_E));
            break;
    }
    return DDAFFT_OK;
}

// Results of a solve on their way back to the host: offsets into the
// context's pinned download buffer, filled by stream-ordered copies, read
// after the stream (or an event after the copies) has completed.
struct Fin {
    int nrhs = 0, ndirs = 0, cb = 0, fchunk = 0, max_step = 0;
    bool state = false, hist = false, grouped = false;
    size_t o_st = 0, o_hist = 0, o_part = 0, o_esc = 0, total = 0;
};

// Block counts of the observables: cross-section partials per right-hand
// side (block sums, or the one sequential float sum of the reference-order
// mode) and dipole chunks per far-field direction tile (enough blocks to
// fill the chip).
// The FP64-summed far field on the matrix cores (k_far_field_mfma): when the
// context asks for it (or by default) and the phase rows of 16 directions fit
// LDS (ni + nj + nl <= 384).
constexpr int kFFMfmaDefault = 0;
bool ff_mfma_on(const ddafft_ctx* c) {
    const int mode = c->ff_mfma >= 0 ? c->ff_mfma : kFFMfmaDefault;
    return mode == 1 && !c->sum_mode && c->d.ni + c->d.nj + c->d.nl <= 384;
}

// The far field by groups of directions sharing z (k_far_field_grp): when a
// set of ndirs directions splits into ng groups of S = 2..16 (dir_groups)
// and the lattice's phase rows fit the kernel's LDS (n_i + n_j <= 512,
// n_l <= 256).
bool ff_group_ok(const ddafft_ctx* c, int ndirs, int ng, int S) {
    const int mode = c->ff_group >= 0 ? c->ff_group : 1;
    return mode == 1 && !c->sum_mode && !ff_mfma_on(c) && ndirs > 0 && ng > 0 && S >= 2 && S <= kFFDirs &&
           ng * S == ndirs && c->nseg > 0 && c->d.ni + c->d.nj <= 512 && c->d.nl <= 256;
}
// ... for the resident set
bool ff_group_on(const ddafft_ctx* c, int ndirs) {
    return !c->ffg_off_call && ndirs == c->ndirs && ff_group_ok(c, ndirs, c->ffg_n, c->ffg_S);
}

void obs_blocks(const ddafft_ctx* c, int ndirs, bool grouped, int* cb, int* fchunk) {
    *cb = c->sum_mode ? 1 : (int)std::min<long long>((c->d.n + kBlock - 1) / kBlock, 512);
    if (grouped) {   // one chunk per 256 segments, at most 64
        *fchunk = (int)std::max(1, std::min(64, (c->nseg + kBlock - 1) / kBlock));
        return;
    }
    const bool mf = ff_mfma_on(c);
    const int fd = mf ? kFDM : kFD;
    const int dtiles = (ndirs + fd - 1) / fd;
    // the MFMA form walks its chunk with dependent loads: shorter chunks, more blocks
    const int per = mf ? 1024 : 4096, want = mf ? 4096 : 2048;
    *fchunk = ndirs > 0 ? (int)std::max(1, std::min({64, (c->ncells + per - 1) / per,
                                                     (want + dtiles - 1) / std::max(dtiles, 1)})) : 0;
}

// Sizes the downloads and reserves the pinned buffer (before anything of this
// solve is queued to it).
int fin_layout(ddafft_ctx* c, int nrhs, bool state, bool hist, int max_step, int ndirs, Fin& f) {
    f = Fin{};
    f.nrhs = nrhs;
    f.ndirs = ndirs;
    f.state = state;
    f.hist = state && hist;
    f.max_step = max_step;
    f.grouped = ff_group_on(c, ndirs);
    obs_blocks(c, ndirs, f.grouped, &f.cb, &f.fchunk);
    size_t off = 0;
    auto take = [&](size_t b) {
        const size_t o = off;
        off += (b + 255) & ~(size_t)255;
        return o;
    };
    if (f.state) f.o_st = take((size_t)nrhs * sizeof(SolveState));
    if (f.hist) f.o_hist = take((size_t)nrhs * (max_step + 1) * sizeof(double));
    f.o_part = take((size_t)nrhs * f.cb * 2 * sizeof(double));
    if (ndirs > 0) f.o_esc = take((size_t)nrhs * ndirs * 3 * sizeof(float2));
    f.total = off;
    TRY(ensure_partials(c, std::max<size_t>((size_t)nrhs * f.cb * 2, (size_t)nrhs * f.fchunk * std::max(ndirs, 1) * 6)));
    if (ndirs > 0 && nrhs * ndirs * 3 > c->esc_cap) {
        HIPCHECK(hipStreamSynchronize(c->stream));
        if (c->esc) hipFree(c->esc);
        c->esc = nullptr;
        c->esc_cap = 0;
        TRY(dev_alloc(&c->esc, (size_t)nrhs * ndirs * 3));
        c->esc_cap = nrhs * ndirs * 3;
    }
    if (f.grouped && (size_t)c->ncells * 6 > c->ffpc_cap) {
        HIPCHECK(hipStreamSynchronize(c->stream));
        if (c->ffpc) hipFree(c->ffpc);
        c->ffpc = nullptr;
        c->ffpc_cap = 0;
        TRY(dev_alloc(&c->ffpc, (size_t)c->ncells * 6));
        c->ffpc_cap = (size_t)c->ncells * 6;
    }
    const size_t ntab = (size_t)ndirs * (c->d.ni + c->d.nj + c->d.nl);
    if (ndirs > 0 && ntab > c->fftab_cap) {
        HIPCHECK(hipStreamSynchronize(c->stream));
        if (c->fftab) hipFree(c->fftab);
        c->fftab = nullptr;
        c->fftab_cap = 0;
        TRY(dev_alloc(&c->fftab, ntab));
        c->fftab_cap = ntab;
    }
    return pinned_reserve(c, &c->dn, &c->dn_cap, off, &c->dn_dev);
}

// Cross sections and far fields of the nrhs solutions in c->P; c->kinc and
// c->einc hold each right-hand side's incidence.  Right-hand side r uses the
// directions ddirs + r * dir_stride.  Queues the kernels and the copies of
// their results into the pinned buffer.
int obs_enqueue_seq(ddafft_ctx* c, const Fin& f, const float* ddirs, long long dir_stride);

// The grouped far field of right-hand sides r0 .. r0 + R - 1 (R = 1, 2):
// their dipole moments gathered in list order, then k_far_field_grp with as
// many groups per block as fit 16 directions and 24 T values per thread.
int ff_grouped_enqueue(ddafft_ctx* c, const Fin& f, int r0, int R, int ndirs) {
    Dims d = c->d;
    const int S = c->ffg_S, ng = c->ffg_n;
    if (!c->ffpc || !c->ffseg || (size_t)c->ncells * 3 * R > c->ffpc_cap)
        return fail(DDAFFT_ESTATE, "internal: grouped far-field buffers not allocated");
    const int gb = std::min((c->ncells + kBlock - 1) / kBlock, 2048);
    if (R == 2)
        hipLaunchKernelGGL(k_ff_gather<2>, gb, kBlock, 0, c->stream, d, c->cells, c->ncells, r0, c->P, c->ffpc);
    else
        hipLaunchKernelGGL(k_ff_gather<1>, gb, kBlock, 0, c->stream, d, c->cells, c->ncells, r0, c->P, c->ffpc);
    HIPCHECK(hipGetLastError());
    int G = 1;
    while (2 * G * S <= kFFDirs && 2 * G * 3 * R <= 24) G *= 2;
    const dim3 grid(f.fchunk, (ng + G - 1) / G);
    const bool small = d.ni + d.nj <= 128 && d.nl <= 64;
#define DDA_FFG(RR, GG)                                                                                        \
    do {                                                                                                        \
        if (small)                                                                                              \
            hipLaunchKernelGGL((k_far_field_grp<RR, GG, 128, 64>), grid, kBlock, 0, c->stream, d, c->fftab,      \
                               c->ffpc, c->cells, c->ffseg, c->nseg, ng, S, ndirs, r0, c->partials);            \
        else                                                                                                    \
            hipLaunchKernelGGL((k_far_field_grp<RR, GG, 512, 256>), grid, kBlock, 0, c->stream, d, c->fftab,     \
                               c->ffpc, c->cells, c->ffseg, c->nseg, ng, S, ndirs, r0, c->partials);            \
    } while (0)
    if (R == 2) {
        if (G == 4) DDA_FFG(2, 4);
        else if (G == 2) DDA_FFG(2, 2);
        else DDA_FFG(2, 1);
    } else {
        if (G == 8) DDA_FFG(1, 8);
        else if (G == 4) DDA_FFG(1, 4);
        else if (G == 2) DDA_FFG(1, 2);
        else DDA_FFG(1, 1);
    }
#undef DDA_FFG
    HIPCHECK(hipGetLastError());
    return DDAFFT_OK;
}

int obs_enqueue(ddafft_ctx* c, const Fin& f, const float* ddirs, long long dir_stride) {
    if (c->sum_mode) return obs_enqueue_seq(c, f, ddirs, dir_stride);
    const int nrhs = f.nrhs, ndirs = f.ndirs;
    Dims d = c->d;
    d.nrhs = nrhs;
    const Geometry g = make_geometry(c, nullptr);
    if (!c->dn_dev) return fail(DDAFFT_ESTATE, "internal: result buffer not reserved");
    TRY(flush_small(c));
    dim3 gc(f.cb, nrhs);
    // the block partials go straight to the pinned result buffer
    hipLaunchKernelGGL(k_cabs_cext, gc, kBlock, 0, c->stream, d, g, c->kinc, inc_phase(c), c->occ,
                       c->alpha, c->nmat, c->einc, c->P, reinterpret_cast<double*>(c->dn_dev + f.o_part));
    HIPCHECK(hipGetLastError());
    if (ndirs > 0) {
        // right-hand sides in pairs that share one direction set share the
        // phases; the phase table is rebuilt (in stream order) only when the
        // direction set changes
        if (!c->fftab || (size_t)ndirs * (d.ni + d.nj + d.nl) > c->fftab_cap)
            return fail(DDAFFT_ESTATE, "internal: far-field phase table not allocated");
        const int ntab = ndirs * (d.ni + d.nj + d.nl);
        const float* tab_dirs = nullptr;
        for (int r0 = 0; r0 < nrhs;) {
            const bool pair = r0 + 1 < nrhs && (dir_stride == 0 || c->dir_pairs[r0 / 2]);
            dim3 gf(f.fchunk, (ndirs + kFD - 1) / kFD);
            const float* dd = ddirs + r0 * dir_stride;
            if (dd != tab_dirs) {
                hipLaunchKernelGGL(k_ff_phase, std::min((ntab + kBlock - 1) / kBlock, 1024), kBlock, 0, c->stream, d,
                                   g, dd, ndirs, c->fftab);
                HIPCHECK(hipGetLastError());
                tab_dirs = dd;
            }
            // phase rows in LDS sized for lattices up to 256 per axis, or the
            // larger form for 512-point axes
            const bool big = d.ni + d.nj + d.nl > kFFAxis;
            if (f.grouped) {
                TRY(ff_grouped_enqueue(c, f, r0, pair ? 2 : 1, ndirs));
            } else if (ff_mfma_on(c)) {
                const dim3 gm(f.fchunk, (ndirs + kFDM - 1) / kFDM);
                const bool small = d.ni + d.nj + d.nl <= 192;
#define DDA_FFM(R, AXM)                                                                                 \
    hipLaunchKernelGGL((k_far_field_mfma<R, AXM>), gm, kBlock, 0, c->stream, d, c->fftab, c->cells, c->ncells, \
                       ndirs, r0, c->P, c->partials)
                if (pair && small) DDA_FFM(2, 192);
                else if (pair) DDA_FFM(2, 384);
                else if (small) DDA_FFM(1, 192);
                else DDA_FFM(1, 384);
#undef DDA_FFM
            } else if (pair && !big)
                hipLaunchKernelGGL((k_far_field<2, kFFAxis>), gf, kBlock, 0, c->stream, d, c->fftab, c->cells,
                                   c->ncells, ndirs, r0, c->P, c->partials);
            else if (pair)
                hipLaunchKernelGGL((k_far_field<2, kFFAxisMax>), gf, kBlock, 0, c->stream, d, c->fftab, c->cells,
                                   c->ncells, ndirs, r0, c->P, c->partials);
            else if (!big)
                hipLaunchKernelGGL((k_far_field<1, kFFAxis>), gf, kBlock, 0, c->stream, d, c->fftab, c->cells,
                                   c->ncells, ndirs, r0, c->P, c->partials);
            else
                hipLaunchKernelGGL((k_far_field<1, kFFAxisMax>), gf, kBlock, 0, c->stream, d, c->fftab, c->cells,
                                   c->ncells, ndirs, r0, c->P, c->partials);
            HIPCHECK(hipGetLastError());
            r0 += pair ? 2 : 1;
        }
        // one thread per (right-hand side, direction), in blocks of 64 to spread over the CUs
        hipLaunchKernelGGL(k_far_field_sum, (nrhs * ndirs + 63) / 64, 64, 0, c->stream, c->partials,
                           f.fchunk, ndirs, nrhs, ddirs, dir_stride, reinterpret_cast<float2*>(c->dn_dev + f.o_esc));
        HIPCHECK(hipGetLastError());
    }
    return DDAFFT_OK;
}

// The same observables with the reference's sequential float sums
// (sum_mode 1): the cross-section terms of every dipole, then one thread per
// sum adds them in the reference's cell order; the far field goes through
// the dipole list in chunks (terms of a chunk formed in parallel, then each
// of the nrhs * ndirs * 6 float sums continued over the chunk in order).
// Results land where obs_enqueue puts them: partials [nrhs][1][2] (the float
// sums, widened exactly) and esc.
int obs_enqueue_seq(ddafft_ctx* c, const Fin& f, const float* ddirs, long long dir_stride) {
    const int nrhs = f.nrhs, ndirs = f.ndirs, nc = c->ncells;
    Dims d = c->d;
    d.nrhs = nrhs;
    const Geometry g = make_geometry(c, nullptr);
    // chunk of dipoles whose far-field terms fit 256 MB, or the space the
    // allocation-free call has left (at least one dipole)
    const long long per = (long long)nrhs * std::max(ndirs, 1) * 6;
    const long long budget = t_noalloc ? (long long)c->seqbuf_cap : (64LL << 20);
    const int chunk = (int)std::max<long long>(1, std::min<long long>(nc, budget / per));
    const size_t need = std::max<size_t>((size_t)2 * nrhs * nc, ndirs > 0 ? (size_t)chunk * per : 0);
    if (need > c->seqbuf_cap) {
        HIPCHECK(hipStreamSynchronize(c->stream));
        if (c->seqbuf) hipFree(c->seqbuf);
        c->seqbuf = nullptr;
        c->seqbuf_cap = 0;
        TRY(dev_alloc(&c->seqbuf, need));
        c->seqbuf_cap = need;
    }
    if (f.cb != 1) return fail(DDAFFT_ESTATE, "internal: sequential sums need one partial per right-hand side");
    if (!c->dn_dev) return fail(DDAFFT_ESTATE, "internal: result buffer not reserved");
    TRY(flush_small(c));
    {
        dim3 gt((int)std::min<long long>((nc + kBlock - 1) / kBlock, 1024), nrhs);
        hipLaunchKernelGGL(k_cc_terms, gt, kBlock, 0, c->stream, d, g, c->kinc, inc_phase(c), c->cells, nc, c->occ,
                           c->alpha, c->nmat, c->einc, c->P, c->seqbuf);
        HIPCHECK(hipGetLastError());
        hipLaunchKernelGGL(k_seq_sum, (2 * nrhs + 63) / 64, 64, 0, c->stream, c->seqbuf, nc, 2 * nrhs,
                           reinterpret_cast<double*>(c->dn_dev + f.o_part));
        HIPCHECK(hipGetLastError());
    }
    if (ndirs > 0) {
        float* acc = reinterpret_cast<float*>(c->esc);   // [nrhs][ndirs][6]
        const int nsums = (int)per;
        for (int e0 = 0; e0 < nc; e0 += chunk) {
            const int cnt = std::min(chunk, nc - e0);
            const long long tot = (long long)cnt * nrhs * ndirs;
            hipLaunchKernelGGL(k_ff_terms, (int)std::min<long long>((tot + kBlock - 1) / kBlock, 8192), kBlock, 0,
                               c->stream, d, g, c->cells, e0, cnt, ndirs, nrhs, ddirs, dir_stride, c->P, c->seqbuf);
            HIPCHECK(hipGetLastError());
            float* out = e0 + cnt >= nc ? reinterpret_cast<float*>(c->dn_dev + f.o_esc) : nullptr;
            hipLaunchKernelGGL(k_ff_seq_sum, (nsums + kBlock - 1) / kBlock, kBlock, 0, c->stream, c->seqbuf, cnt, nsums,
                               e0 == 0 ? 1 : 0, acc, out);
            HIPCHECK(hipGetLastError());
        }
    }
    return DDAFFT_OK;
}

// After the copies have completed: Cabs, Cext (fixed-order FP64 sums of the
// block partials) and the far fields.
void obs_complete(ddafft_ctx* c, const Fin& f, const float* einc, float* esc, float* cabs, float* cext) {
    const double* part = reinterpret_cast<const double*>(c->dn + f.o_part);
    for (int r = 0; r < f.nrhs; ++r) {
        double s1 = 0, s2 = 0;
        for (int b = 0; b < f.cb; ++b) {
            s1 += part[((size_t)r * f.cb + b) * 2];
            s2 += part[((size_t)r * f.cb + b) * 2 + 1];
        }
        const float r0 = cabs_prefactor(einc + 6 * r);
        if (cext) cext[r] = r0 * (float)s1;
        if (cabs) cabs[r] = r0 * (float)s2;
    }
    if (f.ndirs > 0 && esc) std::memcpy(esc, c->dn + f.o_esc, (size_t)f.nrhs * f.ndirs * 3 * sizeof(float2));
}

// Standalone cross sections / far fields (one synchronization).
int observables(ddafft_ctx* c, int nrhs, const float* einc, int ndirs, const float* ddirs, long long dir_stride,
                float* esc, float* cabs, float* cext) {
    Fin f;
    TRY(fin_layout(c, nrhs, false, false, 0, ndirs, f));
    TRY(obs_enqueue(c, f, ddirs, dir_stride));
    HIPCHECK(hipStreamSynchronize(c->stream));
    obs_complete(c, f, einc, esc, cabs, cext);
    return DDAFFT_OK;
}

}  // namespace

namespace {

// One context's solve in flight.  The sources (P_0) are in c->P and the
// polarizabilities in c->alpha when it begins; then the host enqueues one
// iteration at a time, keeping one iteration queued ahead of the stopping
// decision the device makes for each right-hand side.
struct Run {
    ddafft_ctx* c = nullptr;
    int nrhs = 0, solver = 0, max_step = 0;
    float tau = 0.f, err_max = 0.f;
    Seq q{};
    int o = 1;            // next iteration to enqueue
    int known_stop = 0;   // first iteration known not to be needed
};

int run_begin(Run& r) {
    ddafft_ctx* c = r.c;
    Dims d = c->d;
    d.nrhs = r.nrhs;
    if (r.max_step + 1 > c->hist_cap || c->hist == nullptr) {
        HIPCHECK(hipStreamSynchronize(c->stream));
        if (c->hist) hipFree(c->hist);
        c->hist = nullptr;
        TRY(dev_alloc(&c->hist, (size_t)c->nrhs_max * (r.max_step + 1)));
        c->hist_cap = r.max_step + 1;
    }
    std::vector<SolveState> s0(r.nrhs);
    for (auto& s : s0) {
        memset(&s, 0, sizeof(s));
        s.stop_at = r.max_step + 1;
    }
    TRY(upload(c, c->st, s0.data(), r.nrhs * sizeof(SolveState)));
    TRY(ensure_partials(c, (size_t)r.nrhs * std::max(zi_blocks(d), update_blocks(d)) * kSlots));
    TRY(flush_small(c));   // every staged input of the solve, in one copy
    if (c->pending_first_approx) {
        c->pending_first_approx = false;
        const Geometry g = make_geometry(c, nullptr);
        dim3 grid((int)std::min<long long>((d.n + kBlock - 1) / kBlock, 2048), r.nrhs);
        hipLaunchKernelGGL(k_first_approx, grid, kBlock, 0, c->stream, d, g, c->kinc, inc_phase(c), c->occ, c->mv,
                           c->alpha, c->nmat, c->einc, c->P);
        HIPCHECK(hipGetLastError());
    }
    r.q = Seq{c, r.nrhs, FinArgs{}};
    r.q.fa.hist = c->hist;
    r.q.fa.max_step = r.max_step;
    r.q.fa.err_max = (double)r.err_max;
    r.q.fa.inv_ndip = 1.0 / (double)c->ndip;
    TRY(step0(r.q, r.solver));   // its last scalar step publishes the states to ring slot 0
    HIPCHECK(hipEventRecord(c->ev[0], c->stream));
    r.o = 1;
    r.known_stop = r.max_step + 1;
    return DDAFFT_OK;
}

bool run_more(const Run& r) { return r.o <= r.max_step && r.o < r.known_stop; }

// Everything the launches of one captured iteration bake in: if any of it
// changes, the graph is captured again.
void graph_key(const Run& r, std::vector<char>& out) {
    const ddafft_ctx* c = r.c;
    struct {
        int solver, nrhs, max_step, x_rb, x_blocks, nmat, fuse, ahat_mode;
        float tau;
        double err_max, inv_ndip;
        const void *hist, *partials, *alpha, *ring, *mv, *ahat;
    } k;
    std::memset(&k, 0, sizeof(k));
    k.solver = r.solver;
    k.nrhs = r.nrhs;
    k.max_step = r.max_step;
    k.x_rb = c->x_rb;
    k.x_blocks = c->x_blocks;
    k.fuse = c->fuse_zy;
    k.ahat_mode = c->ahat_mode;
    k.nmat = c->nmat;
    k.tau = r.tau;
    k.err_max = r.q.fa.err_max;
    k.inv_ndip = r.q.fa.inv_ndip;
    k.hist = c->hist;
    k.partials = c->partials;
    k.alpha = c->alpha;
    k.ring = c->ring_dev;
    k.mv = c->mv;
    k.ahat = c->ahat;
    const char* b = reinterpret_cast<const char*>(&k);
    out.assign(b, b + sizeof(k));
}

// Enqueues one steady-state iteration as a replay of its captured graph
// (captured on first use: the kernels take their iteration from the solver
// states, so the same launches serve every iteration).
int graph_iteration(Run& r) {
    ddafft_ctx* c = r.c;
    std::vector<char> key;
    graph_key(r, key);
    if (!c->gexec || key != c->gkey) {
        if (c->gexec) HIPCHECK(hipGraphExecDestroy(c->gexec));
        if (c->graph) HIPCHECK(hipGraphDestroy(c->graph));
        c->gexec = nullptr;
        c->graph = nullptr;
        TRY(flush_small(c));   // nothing staged may be copied from inside the graph
        HIPCHECK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        const int rc = iteration(r.q, r.solver, r.o, r.tau);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (rc) {
            if (g) hipGraphDestroy(g);
            return rc;
        }
        HIPCHECK(e);
        hipGraphExec_t x = nullptr;
        const hipError_t ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
        if (ei != hipSuccess) {
            hipGraphDestroy(g);
            return fail(DDAFFT_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
        }
        c->graph = g;
        c->gexec = x;
        c->gkey.swap(key);
    }
    HIPCHECK(hipGraphLaunch(c->gexec, c->stream));
    return DDAFFT_OK;
}

// enqueues iteration r.o; its last scalar step publishes the states (the
// stopping decisions) to ring slot o & 3, which the event marks as written.
// Iterations of one form run as graph replays; CG's first iteration has a
// form of its own, profiling records events per pass, and the
// allocation-free solve does not instantiate graphs: those run eagerly.
int run_enqueue(Run& r) {
    ddafft_ctx* c = r.c;
    const int o = r.o;
    const bool steady = !(r.solver == DDAFFT_SOLVER_CG && o == 1);
    if (schedule(c, r.nrhs).graph && steady && !c->prof && !t_noalloc)
        TRY(graph_iteration(r));
    else
        TRY(iteration(r.q, r.solver, o, r.tau));
    HIPCHECK(hipEventRecord(c->ev[o & 3], c->stream));
    r.o = o + 1;
    return DDAFFT_OK;
}

// waits for the stop flag of the iteration before the last one enqueued
int run_wait(Run& r) {
    const int w = r.o - 2;
    if (w < 0) return DDAFFT_OK;
    HIPCHECK(hipEventSynchronize(r.c->ev[w & 3]));
    const SolveState* sr = r.c->state_ring + (size_t)(w & 3) * r.c->nrhs_max;
    int m = 0;
    for (int q = 0; q < r.nrhs; ++q) m = std::max(m, sr[q].stop_at);
    r.known_stop = std::min(r.known_stop, m);
    return DDAFFT_OK;
}

// One more iteration of a single solve.  BiCGStab iterations launched
// eagerly go in two parts: the first matvec and its scalar step are queued
// before the host reads the stopping decisions of the previous iteration,
// the rest only if the solve goes on.  After the last iteration only that
// first part runs idle (its kernels exit at once) instead of a whole
// iteration, and the GPU has the first part to run while the host decides.
// Other solvers and graph replays queue whole iterations one ahead.
int run_step(Run& r) {
    ddafft_ctx* c = r.c;
    const int o = r.o;
    const bool graph = schedule(c, r.nrhs).graph && !c->prof && !t_noalloc;
    if (r.solver != DDAFFT_SOLVER_BICGSTAB || graph) {
        TRY(run_enqueue(r));
        return run_wait(r);
    }
    TRY(iteration(r.q, r.solver, o, r.tau, 1));
    // the decisions of iteration o - 1 (step 0 for o = 1), published to its ring slot
    HIPCHECK(hipEventSynchronize(c->ev[(o - 1) & 3]));
    const SolveState* sr = c->state_ring + (size_t)((o - 1) & 3) * c->nrhs_max;
    int m = 0;
    for (int q = 0; q < r.nrhs; ++q) m = std::max(m, sr[q].stop_at);
    r.known_stop = std::min(r.known_stop, m);
    if (o >= r.known_stop) return DDAFFT_OK;   // iteration o does not run: run_more ends the loop
    TRY(iteration(r.q, r.solver, o, r.tau, 2));
    HIPCHECK(hipEventRecord(c->ev[o & 3], c->stream));
    r.o = o + 1;
    return DDAFFT_OK;
}

// best-iterate restore, which also writes the final states into the pinned
// result buffer; then the copy of the history, when asked for
int fin_state_enqueue(Run& r, const Fin& f) {
    ddafft_ctx* c = r.c;
    if (!c->dn_dev) return fail(DDAFFT_ESTATE, "internal: result buffer not reserved");
    Dims d = c->d;
    d.nrhs = r.nrhs;
    {
        dim3 grid((int)std::min<long long>((3 * d.n + kBlock - 1) / kBlock, 2048), r.nrhs);
        hipLaunchKernelGGL(k_restore, grid, kBlock, 0, c->stream, d, c->st, c->P, c->Pmin,
                           reinterpret_cast<SolveState*>(c->dn_dev + f.o_st));
        HIPCHECK(hipGetLastError());
    }
    if (f.hist)   // into the result buffer's device address (page-locked memory, or the scratch)
        HIPCHECK(hipMemcpyAsync(c->dn_dev + f.o_hist, c->hist, (size_t)r.nrhs * (r.max_step + 1) * sizeof(double),
                                hipMemcpyDefault, c->stream));
    return DDAFFT_OK;
}

// final states, history and counters, once the copies have completed
int fin_state_complete(Run& r, const Fin& f, double* err_hist, int* info, std::vector<SolveState>& sf) {
    ddafft_ctx* c = r.c;
    const int nrhs = r.nrhs, max_step = r.max_step;
    sf.assign(reinterpret_cast<const SolveState*>(c->dn + f.o_st),
              reinterpret_cast<const SolveState*>(c->dn + f.o_st) + nrhs);
    if (c->prof) {
        int last = 0;
        for (int q = 0; q < nrhs; ++q) last = std::max(last, sf[q].last);
        TRY(prof_collect(c, last));
    }
    if (err_hist && f.hist) {
        const double* h = reinterpret_cast<const double*>(c->dn + f.o_hist);
        for (int q = 0; q < nrhs; ++q)
            for (int o = 0; o <= max_step; ++o)
                err_hist[(size_t)q * (max_step + 1) + o] = o <= sf[q].last ? h[(size_t)q * (max_step + 1) + o] : NAN;
    }
    if (info) {
        for (int q = 0; q < nrhs; ++q) {
            const int L = sf[q].last;
            int mvs = 0;
            switch (r.solver) {
                case DDAFFT_SOLVER_CG: mvs = 1 + L; break;
                case DDAFFT_SOLVER_AA: mvs = 2 + 2 * L; break;
                case DDAFFT_SOLVER_BICGSTAB: mvs = 3 + 2 * L; break;
                default: mvs = 1 + L; break;
            }
            info[4 * q + 0] = L;
            info[4 * q + 1] = sf[q].step_min;
            info[4 * q + 2] = sf[q].restored;
            info[4 * q + 3] = mvs;
        }
    }
    return DDAFFT_OK;
}

int run_finish(Run& r, double* err_hist, int* info, std::vector<SolveState>& sf) {
    Fin f;
    TRY(fin_layout(r.c, r.nrhs, true, err_hist != nullptr, r.max_step, 0, f));
    TRY(fin_state_enqueue(r, f));
    HIPCHECK(hipStreamSynchronize(r.c->stream));
    return fin_state_complete(r, f, err_hist, info, sf);
}

int solve_core(ddafft_ctx* c, int nrhs, int solver, float tau, int max_step, float err_max, double* err_hist,
               int* info, std::vector<SolveState>& sf) {
    Run r;
    r.c = c;
    r.nrhs = nrhs;
    r.solver = solver;
    r.tau = tau;
    r.max_step = max_step;
    r.err_max = err_max;
    TRY(run_begin(r));
    while (run_more(r)) TRY(run_step(r));
    return run_finish(r, err_hist, info, sf);
}

// Checks one job against its context and stages its inputs: polarizabilities,
// incidence, directions; run_begin then copies them and computes the first
// approximation in c->P.
int job_prepare(ddafft_ctx* c, const ddafft_job* j, int solver, int max_step, const float** ddirs,
                long long* dir_stride) {
    if (!c || !j || !j->alpha || !j->kinc || !j->einc) return fail(DDAFFT_EINVAL, "NULL argument");
    if (!c->have_target) return fail(DDAFFT_ESTATE, "set_target first");
    const int nrhs = j->nrhs;
    if (nrhs < 1 || nrhs > c->nrhs_max) return fail(DDAFFT_EINVAL, "nrhs outside [1, nrhs_max]");
    if (solver < 0 || solver > 3 || max_step < 0) return fail(DDAFFT_EINVAL, "bad solver or max_step");
    const int ndirs = j->ndirs;
    if (ndirs < 0 || (ndirs > 0 && !j->esc)) return fail(DDAFFT_EINVAL, "bad direction arguments");
    {
        const size_t dir_bytes = (ndirs > 0 && j->dirs) ? (size_t)ndirs * 3 * sizeof(float) * (j->dirs_per_rhs ? nrhs : 1) : 0;
        const size_t cell_bytes = j->exact_phases ? (size_t)nrhs * 2 * c->d.n * sizeof(float2) : 0;
        const size_t bytes = dir_bytes + cell_bytes +
                             (size_t)nrhs * (c->nmat * sizeof(float2) + 3 * sizeof(float2) + 3 * sizeof(float) +
                                             sizeof(int) + 2 * (size_t)inc_tab_len(c) * sizeof(float2) +
                                             sizeof(SolveState));
        TRY(upload_begin(c, bytes + 16 * 256));
    }
    *ddirs = nullptr;
    *dir_stride = 0;
    if (ndirs > 0) {
        if (j->dirs) {
            const size_t per = (size_t)ndirs * 3;
            const int sets = j->dirs_per_rhs ? nrhs : 1;
            if ((long long)ndirs * sets > c->dirs_cap) {
                HIPCHECK(hipStreamSynchronize(c->stream));
                if (c->dirs) hipFree(c->dirs);
                c->dirs = nullptr;
                c->dirs_cap = 0;
                TRY(dev_alloc(&c->dirs, per * sets));
                c->dirs_cap = ndirs * sets;
            }
            TRY(upload(c, c->dirs, j->dirs, per * sets * sizeof(float)));
            c->ndirs = j->dirs_per_rhs ? 0 : ndirs;   // per-RHS sets are not resident for later calls
            c->ffg_n = j->dirs_per_rhs ? 0 : dir_groups(j->dirs, ndirs, &c->ffg_S);
            if (j->dirs_per_rhs) {
                *dir_stride = (long long)per;
                c->dir_pairs.assign((size_t)(nrhs + 1) / 2, 0);
                for (int q = 0; q + 1 < nrhs; q += 2)
                    c->dir_pairs[q / 2] =
                        std::memcmp(j->dirs + q * per, j->dirs + (q + 1) * per, per * sizeof(float)) == 0;
            }
        } else if (ndirs != c->ndirs) {
            return fail(DDAFFT_EINVAL, "dirs == NULL but ndirs differs from the resident directions");
        }
        *ddirs = c->dirs;
    }
    Dims d = c->d;
    d.nrhs = nrhs;
    TRY(upload(c, c->alpha, j->alpha, (size_t)nrhs * c->nmat * sizeof(float2)));
    TRY(upload(c, c->einc, j->einc, (size_t)nrhs * 3 * sizeof(float2)));
    TRY(upload_incidence(c, nrhs, j->kinc, j->exact_phases != 0));
    // the first approximation (the sources in c->P) runs once the staged
    // inputs have been copied, at the start of the run
    c->pending_first_approx = true;
    return DDAFFT_OK;
}

}  // namespace

int ddafft_solve_ex(ddafft_ctx* c, const ddafft_job* job, int solver, float tau, int max_step, float err_max,
                    double* err_hist) {
    const float* ddirs = nullptr;
    long long stride = 0;
    TRY(job_prepare(c, job, solver, max_step, &ddirs, &stride));
    Run r;
    r.c = c;
    r.nrhs = job->nrhs;
    r.solver = solver;
    r.tau = tau;
    r.max_step = max_step;
    r.err_max = err_max;
    TRY(run_begin(r));
    while (run_more(r)) TRY(run_step(r));
    Fin f;
    TRY(fin_layout(c, job->nrhs, true, err_hist != nullptr, max_step, job->ndirs, f));
    TRY(fin_state_enqueue(r, f));
    TRY(obs_enqueue(c, f, ddirs, stride));
    HIPCHECK(hipStreamSynchronize(c->stream));
    if (t_noalloc)   // the results sit in the device scratch: one copy to the context's host buffer
        HIPCHECK(hipMemcpy(c->dn, c->dn_dev, f.total, hipMemcpyDeviceToHost));
    std::vector<SolveState> sf;
    TRY(fin_state_complete(r, f, err_hist, job->info, sf));
    obs_complete(c, f, job->einc, job->esc, job->cabs, job->cext);
    return DDAFFT_OK;
}

int ddafft_solve(ddafft_ctx* c, int nrhs, int solver, float tau, int max_step, float err_max, const float* alpha,
                 const float* kinc, const float* einc, int ndirs, const float* dirs, float* esc, float* cabs,
                 float* cext, double* err_hist, int* info) {
    if (!kinc) return fail(DDAFFT_EINVAL, "NULL argument");
    if (nrhs < 1 || nrhs > 1 << 20) return fail(DDAFFT_EINVAL, "nrhs outside [1, nrhs_max]");
    std::vector<float> k3((size_t)3 * nrhs);
    for (int q = 0; q < nrhs; ++q) std::memcpy(&k3[3 * (size_t)q], kinc, 3 * sizeof(float));
    ddafft_job j{};
    j.nrhs = nrhs;
    j.alpha = alpha;
    j.kinc = k3.data();
    j.einc = einc;
    j.ndirs = ndirs;
    j.dirs = dirs;
    j.dirs_per_rhs = 0;
    j.esc = esc;
    j.cabs = cabs;
    j.cext = cext;
    j.info = info;
    return ddafft_solve_ex(c, &j, solver, tau, max_step, err_max, err_hist);
}

int ddafft_solve_noalloc(ddafft_ctx* c, int nrhs, int solver, float tau, int max_step, float err_max,
                         const float* alpha, const float* kinc, const float* einc, int ndirs, const float* dirs,
                         float* esc, float* cabs, float* cext, double* err_hist, int* info) {
    if (!c) return fail(DDAFFT_EINVAL, "NULL context");
    if (!c->scratch) return fail(DDAFFT_ESTATE, "the allocation-free scratch is not available on this device");
    if (!c->have_target) return fail(DDAFFT_ESTATE, "set_target first");
    if (nrhs < 1 || nrhs > c->nrhs_max) return fail(DDAFFT_EINVAL, "nrhs outside [1, nrhs_max]");
    if (max_step < 0 || ndirs < 0) return fail(DDAFFT_EINVAL, "bad max_step or ndirs");
    if (ndirs > 0 && !dirs && ndirs != c->ndirs)
        return fail(DDAFFT_EINVAL, "dirs == NULL but ndirs differs from the resident directions");
    if (c->prof) return fail(DDAFFT_ESTATE, "profiling records events: disable it for the allocation-free solve");
    const Dims& d = c->d;
    // this call's pieces of the scratch
    // the far field's form, as the solve will choose it for these directions;
    // the grouped form needs the dipole moments gathered in the scratch
    bool grouped = false;
    {
        int S = c->ffg_S;
        const int ng = dirs ? dir_groups(dirs, ndirs, &S) : c->ffg_n;
        grouped = ff_group_ok(c, ndirs, ng, S);
    }
    int cb = 0, fchunk = 0;
    size_t off = 0, res = 0, n_part = 0, n_hist = 0, n_esc = 0, n_tab = 0, n_dirs = 0, n_pc = 0;
    size_t o_part = 0, o_hist = 0, o_esc = 0, o_tab = 0, o_dirs = 0, o_res = 0, o_pc = 0;
    auto take = [&](size_t b) {
        const size_t o = off;
        off += (b + 255) & ~(size_t)255;
        return o;
    };
    // the plan of the scratch; the grouped far field only where its pieces fit
    for (;;) {
        off = 0;
        res = 0;   // results, laid out as fin_layout lays them out
        obs_blocks(c, ndirs, grouped, &cb, &fchunk);
        n_part = std::max({(size_t)nrhs * std::max(zi_blocks(d), update_blocks(d)) * kSlots, (size_t)nrhs * cb * 2,
                           (size_t)nrhs * fchunk * std::max(ndirs, 1) * 6});
        n_hist = (size_t)c->nrhs_max * (max_step + 1);
        n_esc = (size_t)nrhs * ndirs * 3;
        n_tab = (size_t)ndirs * (d.ni + d.nj + d.nl);
        n_dirs = dirs ? (size_t)ndirs * 3 : 0;
        n_pc = grouped ? (size_t)c->ncells * 6 : 0;
        {
            auto t = [&](size_t b) { res += (b + 255) & ~(size_t)255; };
            t((size_t)nrhs * sizeof(SolveState));
            if (err_hist) t((size_t)nrhs * (max_step + 1) * sizeof(double));
            t((size_t)nrhs * cb * 2 * sizeof(double));
            if (ndirs > 0) t((size_t)nrhs * ndirs * 3 * sizeof(float2));
        }
        o_part = take(n_part * sizeof(double));
        o_hist = take(n_hist * sizeof(double));
        o_esc = take(n_esc * sizeof(float2));
        o_tab = take(n_tab * sizeof(float2));
        o_dirs = take(n_dirs * sizeof(float));
        o_res = take(res);
        o_pc = take(n_pc * sizeof(float2));
        if (!grouped || off <= kScratchBytes) break;
        grouped = false;
    }
    size_t o_seq = 0, n_seq = 0;
    if (c->sum_mode) {   // the sequential-sum terms get the rest
        o_seq = take(0);
        n_seq = off < kScratchBytes ? (kScratchBytes - off) / sizeof(float) : 0;
        if (n_seq < (size_t)2 * nrhs * c->ncells || n_seq < (size_t)nrhs * std::max(ndirs, 1) * 6)
            return fail(DDAFFT_EINVAL, "the reference-order sums of this solve exceed the allocation-free scratch");
    }
    if (off > kScratchBytes || res > sizeof(c->noalloc_dn))
        return fail(DDAFFT_EINVAL, "this solve needs " + std::to_string(off) + " bytes of scratch (" +
                                       std::to_string(res) + " of results); the allocation-free solve has " +
                                       std::to_string(kScratchBytes) + " (" + std::to_string(sizeof(c->noalloc_dn)) +
                                       ")");
    // bind the scratch for the duration of the call (every capacity check
    // then passes without allocating), then put the context back
    struct Saved {
        double* partials; size_t partials_cap;
        double* hist; int hist_cap;
        float2* esc; int esc_cap;
        float2* fftab; size_t fftab_cap;
        float* dirs; int dirs_cap, ndirs, ffg_n, ffg_S;
        float2* ffpc; size_t ffpc_cap;
        float* seqbuf; size_t seqbuf_cap;
        char *dn, *dn_dev; size_t dn_cap;
    } sv{c->partials, c->partials_cap, c->hist, c->hist_cap, c->esc, c->esc_cap, c->fftab, c->fftab_cap, c->dirs,
         c->dirs_cap, c->ndirs, c->ffg_n, c->ffg_S, c->ffpc, c->ffpc_cap, c->seqbuf, c->seqbuf_cap, c->dn, c->dn_dev,
         c->dn_cap};
    HIPCHECK(hipStreamSynchronize(c->stream));   // nothing of an earlier call still uses the old buffers
    char* sc = c->scratch;
    c->partials = reinterpret_cast<double*>(sc + o_part);
    c->partials_cap = n_part;
    c->hist = reinterpret_cast<double*>(sc + o_hist);
    c->hist_cap = max_step + 1;
    c->esc = reinterpret_cast<float2*>(sc + o_esc);
    c->esc_cap = (int)n_esc;
    c->fftab = reinterpret_cast<float2*>(sc + o_tab);
    c->fftab_cap = n_tab;
    if (dirs) {
        c->dirs = reinterpret_cast<float*>(sc + o_dirs);
        c->dirs_cap = ndirs;
    }
    if (grouped) {
        c->ffpc = reinterpret_cast<float2*>(sc + o_pc);
        c->ffpc_cap = n_pc;
    }
    c->ffg_off_call = !grouped;   // the solve's own choice then agrees with this plan
    if (c->sum_mode) {
        c->seqbuf = reinterpret_cast<float*>(sc + o_seq);
        c->seqbuf_cap = n_seq;
    }
    c->dn = c->noalloc_dn;
    c->dn_dev = sc + o_res;
    c->dn_cap = sizeof(c->noalloc_dn);
    t_noalloc = true;
    const int rc = ddafft_solve(c, nrhs, solver, tau, max_step, err_max, alpha, kinc, einc, ndirs, dirs, esc, cabs,
                                cext, err_hist, info);
    t_noalloc = false;
    c->ffg_off_call = false;
    if (rc) hipStreamSynchronize(c->stream);   // a failed call may have left work queued on the scratch
    c->partials = sv.partials;
    c->partials_cap = sv.partials_cap;
    c->hist = sv.hist;
    c->hist_cap = sv.hist_cap;
    c->esc = sv.esc;
    c->esc_cap = sv.esc_cap;
    c->fftab = sv.fftab;
    c->fftab_cap = sv.fftab_cap;
    c->dirs = sv.dirs;
    c->dirs_cap = sv.dirs_cap;
    c->ndirs = sv.ndirs;
    c->ffg_n = sv.ffg_n;
    c->ffg_S = sv.ffg_S;
    c->ffpc = sv.ffpc;
    c->ffpc_cap = sv.ffpc_cap;
    c->seqbuf = sv.seqbuf;
    c->seqbuf_cap = sv.seqbuf_cap;
    c->dn = sv.dn;
    c->dn_dev = sv.dn_dev;
    c->dn_cap = sv.dn_cap;
    return rc;
}

int ddafft_solve_batch(int n, ddafft_ctx* const* ctxs, const ddafft_job* jobs, int solver, float tau, int max_step,
                       float err_max, int width) {
    if (n < 1 || !ctxs || !jobs) return fail(DDAFFT_EINVAL, "bad batch");
    for (int k = 0; k < n; ++k)
        for (int m = 0; m < k; ++m)
            if (ctxs[k] == ctxs[m]) return fail(DDAFFT_EINVAL, "a context appears twice in one batch");
    if (width < 1) width = kTunedBatchWidth;   // the measured choice
    width = std::min(width, n);
    struct Slot {
        Run r;
        int job = -1;
        const float* ddirs = nullptr;
        long long stride = 0;
        Fin f;
        int next_read = 0;        // next iteration whose solver states the host reads
        bool finishing = false;   // results queued back, event recorded
    };
    std::vector<Slot> act;
    int next = 0;
    auto start = [&](int k) -> int {
        Slot s;
        s.job = k;
        TRY(job_prepare(ctxs[k], &jobs[k], solver, max_step, &s.ddirs, &s.stride));
        s.r.c = ctxs[k];
        s.r.nrhs = jobs[k].nrhs;
        s.r.solver = solver;
        s.r.tau = tau;
        s.r.max_step = max_step;
        s.r.err_max = err_max;
        TRY(run_begin(s.r));
        act.push_back(s);
        return DDAFFT_OK;
    };
    while (next < n && (int)act.size() < width) TRY(start(next++));
    // Event-driven: each instance keeps up to two iterations queued beyond the
    // last one whose stopping decisions the host has read, and reads them as
    // they arrive, so fast instances run ahead of slow ones; the host blocks
    // only when no instance can make progress.
    auto read_states = [&](Slot& s, bool block) -> int {
        const int it = s.next_read;
        hipEvent_t e = s.r.c->ev[it & 3];
        if (block) {
            HIPCHECK(hipEventSynchronize(e));
        } else {
            const hipError_t q = hipEventQuery(e);
            if (q == hipErrorNotReady) return 0;
            HIPCHECK(q);
        }
        const SolveState* sr = s.r.c->state_ring + (size_t)(it & 3) * s.r.c->nrhs_max;
        int m = 0;
        for (int q = 0; q < s.r.nrhs; ++q) m = std::max(m, sr[q].stop_at);
        s.r.known_stop = std::min(s.r.known_stop, m);
        ++s.next_read;
        return 1;
    };
    while (!act.empty()) {
        bool progress = false;
        for (Slot& s : act) {
            if (s.finishing) continue;
            while (s.next_read < s.r.o) {   // iterations queued and not yet read
                const int got = read_states(s, false);
                if (got < 0) return got;
                if (!got) break;
                progress = true;
            }
            while (run_more(s.r) && s.r.o - s.next_read < 2) {
                TRY(run_enqueue(s.r));
                progress = true;
            }
        }
        // instances that are done queue their results back without waiting
        for (Slot& s : act) {
            if (s.finishing || run_more(s.r)) continue;
            progress = true;
            const ddafft_job& j = jobs[s.job];
            TRY(fin_layout(s.r.c, j.nrhs, true, false, max_step, j.ndirs, s.f));
            TRY(fin_state_enqueue(s.r, s.f));
            TRY(obs_enqueue(s.r.c, s.f, s.ddirs, s.stride));
            HIPCHECK(hipEventRecord(s.r.c->fin_ev, s.r.c->stream));
            s.finishing = true;
        }
        // collect what has arrived
        for (size_t k = 0; k < act.size();) {
            Slot& s = act[k];
            if (!s.finishing) {
                ++k;
                continue;
            }
            const hipError_t q = hipEventQuery(s.r.c->fin_ev);
            if (q == hipErrorNotReady) {
                ++k;
                continue;
            }
            HIPCHECK(q);
            progress = true;
            const ddafft_job& j = jobs[s.job];
            std::vector<SolveState> sf;
            TRY(fin_state_complete(s.r, s.f, nullptr, j.info, sf));
            obs_complete(s.r.c, s.f, j.einc, j.esc, j.cabs, j.cext);
            act.erase(act.begin() + (long)k);
            if (next < n) TRY(start(next++));
        }
        if (progress || act.empty()) continue;
        // nothing could move: wait for the oldest outstanding event of the
        // first instance (its next stopping decision, or its results)
        Slot& s = act.front();
        if (s.finishing) {
            HIPCHECK(hipEventSynchronize(s.r.c->fin_ev));
        } else if (s.next_read < s.r.o) {
            const int got = read_states(s, true);
            if (got < 0) return got;
        }
    }
    return DDAFFT_OK;
}

}  // extern "C"

namespace {

template <class T>
int ensure_buf(ddafft_ctx* c, T** p, size_t& cap, size_t count) {
    if (count <= cap && *p) return DDAFFT_OK;
    HIPCHECK(hipStreamSynchronize(c->stream));
    if (*p) HIPCHECK(hipFree(*p));
    *p = nullptr;
    cap = 0;
    TRY(dev_alloc(p, count));
    cap = count;
    return DDAFFT_OK;
}

}  // namespace

extern "C" {

int ddafft_gradient(ddafft_ctx* c, int npol, int nobj, int solver, float tau, int max_step, float err_max,
                    const float* alpha, const float* dalpha_dm, const float* m, const float* kinc, const float* einc,
                    int ndirs, const float* dirs, const float* normals, const float* weights, float* f, float* esc,
                    float* cabs, float* cext, float* s11, float* grad_m, float* grad_chi, int* info) {
    if (!c || !alpha || !dalpha_dm || !m || !kinc || !einc || !weights) return fail(DDAFFT_EINVAL, "NULL argument");
    if (!c->have_target) return fail(DDAFFT_ESTATE, "set_target first");
    if (npol < 1 || npol > 2 || nobj < 1) return fail(DDAFFT_EINVAL, "npol must be 1 or 2 and nobj >= 1");
    const int nadj = npol * nobj;
    if (nadj > c->nrhs_max) return fail(DDAFFT_EINVAL, "npol * nobj exceeds nrhs_max");
    if (ndirs < 0 || (ndirs > 0 && !dirs && ndirs != c->ndirs)) return fail(DDAFFT_EINVAL, "bad directions");
    const int nmat = c->nmat;
    const int nw = 2 + ndirs;
    bool any_s11 = false;
    for (int o = 0; o < nobj; ++o)
        for (int a = 0; a < ndirs; ++a) any_s11 = any_s11 || weights[o * nw + 2 + a] != 0.f;
    if (any_s11 && npol != 2) return fail(DDAFFT_EINVAL, "S11 weights need both polarizations (npol = 2)");
    if ((any_s11 || s11) && ndirs > 0 && (!normals || npol != 2))
        return fail(DDAFFT_EINVAL, "S11 needs npol = 2 and the plane normal of every direction");
    if ((size_t)nadj * nmat * 2 > 4096) return fail(DDAFFT_EINVAL, "too many objectives x materials");
    Dims d = c->d;
    const long long n = d.n;

    // forward solve of the npol polarizations; its dipoles are kept
    std::vector<float> esc_h((size_t)npol * std::max(ndirs, 1) * 6), cabs_h(npol), cext_h(npol);
    std::vector<int> info_f(4 * npol);
    TRY(ddafft_solve(c, npol, solver, tau, max_step, err_max, alpha, kinc, einc, ndirs, dirs,
                     ndirs > 0 ? esc_h.data() : nullptr, cabs_h.data(), cext_h.data(), nullptr, info_f.data()));
    std::vector<float> hdirs((size_t)ndirs * 3);
    if (ndirs > 0) {
        if (dirs)
            std::memcpy(hdirs.data(), dirs, hdirs.size() * sizeof(float));
        else
            HIPCHECK(hipMemcpy(hdirs.data(), c->dirs, hdirs.size() * sizeof(float), hipMemcpyDeviceToHost));
    }
    TRY(ensure_buf(c, &c->Pf, c->Pf_cap, (size_t)npol * 3 * n));
    HIPCHECK(hipMemcpyAsync(c->Pf, c->P, (size_t)npol * 3 * n * sizeof(float2), hipMemcpyDeviceToDevice, c->stream));

    // objective pieces on the host (double)
    using Cd = std::complex<double>;
    std::vector<double> m00(ndirs, 0.);
    std::vector<Cd> v((size_t)2 * ndirs * 3, Cd(0., 0.));   // [p][a][3] dM00/dEsc_p
    if (npol == 2 && ndirs > 0 && normals)
        for (int a = 0; a < ndirs; ++a)
            m00[a] = ddafft_host::m00_grad(kinc, &hdirs[3 * a], normals + 3 * a, einc, einc + 6,
                                           &esc_h[(size_t)a * 6], &esc_h[((size_t)ndirs + a) * 6], &v[(size_t)a * 3],
                                           &v[((size_t)ndirs + a) * 3]);
    std::vector<float> r0(npol);
    for (int p = 0; p < npol; ++p) r0[p] = cabs_prefactor(einc + 6 * p);
    // aux layout (16-byte aligned pieces)
    auto al16 = [](size_t b) { return (b + 15) & ~(size_t)15; };
    const size_t off_u = 0;
    const size_t off_ce = al16(off_u + (size_t)nadj * std::max(ndirs, 1) * 3 * sizeof(float2));
    const size_t off_ca = al16(off_ce + nadj * sizeof(float2));
    const size_t off_ex = al16(off_ca + nadj * sizeof(float));
    const size_t off_s = al16(off_ex + nadj * sizeof(float2));
    const size_t off_is = al16(off_s + nadj * sizeof(float));
    const size_t off_dc = al16(off_is + nadj * sizeof(float));
    const size_t off_na = al16(off_dc + (size_t)npol * nmat * sizeof(float2));
    const size_t aux_bytes = al16(off_na + nadj * sizeof(double));
    std::vector<char> aux(aux_bytes, 0);
    float2* hu = reinterpret_cast<float2*>(aux.data() + off_u);
    float2* hce = reinterpret_cast<float2*>(aux.data() + off_ce);
    float* hca = reinterpret_cast<float*>(aux.data() + off_ca);
    float2* hex = reinterpret_cast<float2*>(aux.data() + off_ex);
    float2* hdc = reinterpret_cast<float2*>(aux.data() + off_dc);
    double* hna = reinterpret_cast<double*>(aux.data() + off_na);
    // size of each forward source alpha_j E_j (summed over the dipoles): the
    // adjoint source is scaled to it on the device, so that err_max means the
    // same relative residual for both solves
    for (int r = 0; r < nadj; ++r) {
        const int p = r % npol;
        double sa = 0.;
        for (int q = 0; q < nmat; ++q)
            sa += (double)c->mat_count[q] *
                  std::hypot((double)alpha[((size_t)p * nmat + q) * 2], (double)alpha[((size_t)p * nmat + q) * 2 + 1]);
        double e2 = 0.;
        for (int q = 0; q < 6; ++q) e2 += (double)einc[6 * p + q] * einc[6 * p + q];
        hna[r] = sa * std::sqrt(e2);
    }
    std::vector<float2> alpha_adj((size_t)nadj * nmat);
    for (int o = 0; o < nobj; ++o) {
        const float* w = weights + (size_t)o * nw;
        double fo = 0., ce = 0., ca = 0.;
        for (int p = 0; p < npol; ++p) {
            ce += cext_h[p] / npol;
            ca += cabs_h[p] / npol;
        }
        fo = w[0] * ce + w[1] * ca;
        for (int a = 0; a < ndirs; ++a) fo += w[2 + a] * m00[a];
        if (f) f[o] = (float)fo;
        for (int p = 0; p < npol; ++p) {
            const int r = o * npol + p;
            hce[r] = make_float2(0.f, (float)(-(double)w[0] / npol * r0[p] / 2.));
            hca[r] = (float)((double)w[1] / npol * r0[p]);
            hex[r] = make_float2(0.f, (float)(-(double)w[1] / npol * r0[p] / 2.));
            for (int a = 0; a < ndirs; ++a) {
                const double* nn = nullptr;
                const double nd[3] = {hdirs[3 * a], hdirs[3 * a + 1], hdirs[3 * a + 2]};
                nn = nd;
                const Cd* vv = &v[((size_t)p * ndirs + a) * 3];
                const Cd nv = nn[0] * vv[0] + nn[1] * vv[1] + nn[2] * vv[2];
                for (int q = 0; q < 3; ++q) {
                    const Cd t = (double)w[2 + a] * (vv[q] - nn[q] * nv);
                    hu[((size_t)r * ndirs + a) * 3 + q] = make_float2((float)t.real(), (float)t.imag());
                }
            }
            for (int q = 0; q < nmat; ++q)
                alpha_adj[(size_t)r * nmat + q] = make_float2(alpha[((size_t)p * nmat + q) * 2],
                                                              alpha[((size_t)p * nmat + q) * 2 + 1]);
        }
    }
    for (int p = 0; p < npol; ++p)
        for (int q = 0; q < nmat; ++q) {
            const Cd mq(m[2 * q], m[2 * q + 1]);
            const Cd da(dalpha_dm[((size_t)p * nmat + q) * 2], dalpha_dm[((size_t)p * nmat + q) * 2 + 1]);
            const Cd dc = std::abs(mq) > 0. ? da / (2. * mq) : Cd(0., 0.);
            hdc[(size_t)p * nmat + q] = make_float2((float)dc.real(), (float)dc.imag());
        }
    TRY(ensure_buf(c, &c->gaux, c->gaux_cap, aux_bytes));
    TRY(upload(c, c->alpha, alpha_adj.data(), alpha_adj.size() * sizeof(float2)));
    const int naxis = d.ni + d.nj + d.nl;
    if (ndirs > 0 && any_s11) {
        TRY(ensure_buf(c, &c->phase, c->phase_cap, (size_t)ndirs * naxis));
        const Geometry g = make_geometry(c, kinc);
        hipLaunchKernelGGL(k_axis_phase, (int)std::min<long long>(((long long)ndirs * naxis + kBlock - 1) / kBlock, 1024),
                           kBlock, 0, c->stream, d, g, c->dirs, ndirs, c->phase);
        HIPCHECK(hipGetLastError());
    }
    // the weights and coefficients (the scales are computed on the device)
    TRY(upload(c, c->gaux, aux.data(), aux_bytes));
    TRY(flush_small(c));   // the adjoint polarizabilities, before the kernels below read them

    // adjoint sources b = diag(alpha) g and their mean norms
    d.nrhs = nadj;
    const int nblk = (int)std::min<long long>((n + kBlock - 1) / kBlock, 1024);
    TRY(ensure_partials(c, (size_t)nadj * nblk));
    {
        AdjArgs aa{};
        aa.Pf = c->Pf;
        aa.alpha = c->alpha;
        aa.einc = c->einc;
        aa.cext_c = reinterpret_cast<const float2*>(c->gaux + off_ce);
        aa.cabs_c = reinterpret_cast<const float*>(c->gaux + off_ca);
        aa.u = reinterpret_cast<const float2*>(c->gaux + off_u);
        aa.tab = c->phase;
        aa.ndirs = any_s11 ? ndirs : 0;   // no S11 weight: the far-field term of g is zero
        aa.npol = npol;
        aa.nmat = nmat;
        aa.tdirs = ndirs > 0 ? std::max(1, std::min(ndirs, kAdjLds / (naxis + 3))) : 1;
        aa.b = c->P;
        aa.partials = c->partials;
        if (aa.ndirs > 0 && !aa.tab) return fail(DDAFFT_ESTATE, "internal: phase table missing");
        const Geometry g = make_geometry(c, kinc);
        hipLaunchKernelGGL(k_adj_source, dim3(nblk, nadj), kBlock, 0, c->stream, d, g, c->occ, c->mv, aa);
        HIPCHECK(hipGetLastError());
    }
    hipLaunchKernelGGL(k_adj_scale, 1, 64, 0, c->stream, c->partials, nblk, nadj,
                       reinterpret_cast<const double*>(c->gaux + off_na), reinterpret_cast<float*>(c->gaux + off_s),
                       reinterpret_cast<float*>(c->gaux + off_is));
    HIPCHECK(hipGetLastError());
    {
        dim3 grid((int)std::min<long long>((3 * n + kBlock - 1) / kBlock, 2048), nadj);
        hipLaunchKernelGGL(k_scale_rhs, grid, kBlock, 0, c->stream, d, reinterpret_cast<const float*>(c->gaux + off_is),
                           c->P);
        HIPCHECK(hipGetLastError());
    }

    // adjoint solve: the same iteration from the sources in c->P; its final
    // states come back with the gradients (one synchronization at the end)
    std::vector<SolveState> sfa;
    std::vector<int> info_a(4 * nadj);
    Run ra;
    ra.c = c;
    ra.nrhs = nadj;
    ra.solver = solver;
    ra.tau = tau;
    ra.max_step = max_step;
    ra.err_max = err_max;
    TRY(run_begin(ra));
    while (run_more(ra)) TRY(run_step(ra));
    Fin fa;
    TRY(fin_layout(c, nadj, true, false, max_step, 0, fa));
    TRY(fin_state_enqueue(ra, fa));

    // df/dalpha per dipole, the chi map, and the per-material sums
    {
        GradArgs ga{};
        ga.lam = c->P;
        ga.s = reinterpret_cast<const float*>(c->gaux + off_s);
        ga.Pf = c->Pf;
        ga.alpha = c->alpha;
        ga.dadchi = reinterpret_cast<const float2*>(c->gaux + off_dc);
        ga.expl = reinterpret_cast<const float2*>(c->gaux + off_ex);
        ga.npol = npol;
        ga.nmat = nmat;
        ga.H = c->R;
        ga.gchi = c->S;
        dim3 grid((int)std::min<long long>((n + kBlock - 1) / kBlock, 2048), nobj);
        hipLaunchKernelGGL(k_grad, grid, kBlock, 0, c->stream, d, c->occ, c->mv, ga);
        HIPCHECK(hipGetLastError());
    }
    const int nb2 = (int)std::min<long long>((n + kBlock - 1) / kBlock, 256);
    TRY(ensure_partials(c, (size_t)nadj * nb2 * nmat * 2));
    hipLaunchKernelGGL(k_mat_sum, dim3(nb2, nadj), kBlock, 0, c->stream, d, c->occ, c->R, nmat, c->partials);
    HIPCHECK(hipGetLastError());
    std::vector<double> ms((size_t)nadj * nb2 * nmat * 2);
    HIPCHECK(hipMemcpyAsync(ms.data(), c->partials, ms.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (grad_chi)
        HIPCHECK(hipMemcpyAsync(grad_chi, c->S, (size_t)nobj * n * sizeof(float2), hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));
    TRY(fin_state_complete(ra, fa, nullptr, info_a.data(), sfa));
    if (grad_m) {
        for (int o = 0; o < nobj; ++o)
            for (int q = 0; q < nmat; ++q) {
                Cd t(0., 0.);
                for (int p = 0; p < npol && q > 0; ++p) {
                    const int r = o * npol + p;
                    Cd h(0., 0.);
                    for (int b = 0; b < nb2; ++b) {
                        const double* e = &ms[(((size_t)r * nb2 + b) * nmat + q) * 2];
                        h += Cd(e[0], e[1]);
                    }
                    t += h * Cd(dalpha_dm[((size_t)p * nmat + q) * 2], dalpha_dm[((size_t)p * nmat + q) * 2 + 1]);
                }
                grad_m[((size_t)o * nmat + q) * 2] = (float)(2. * t.real());
                grad_m[((size_t)o * nmat + q) * 2 + 1] = (float)(-2. * t.imag());
            }
    }
    if (esc && ndirs > 0) std::memcpy(esc, esc_h.data(), (size_t)npol * ndirs * 6 * sizeof(float));
    for (int p = 0; p < npol; ++p) {
        if (cabs) cabs[p] = cabs_h[p];
        if (cext) cext[p] = cext_h[p];
    }
    if (s11)
        for (int a = 0; a < ndirs; ++a) s11[a] = (float)m00[a];
    if (info) {
        std::memcpy(info, info_f.data(), info_f.size() * sizeof(int));
        std::memcpy(info + 4 * npol, info_a.data(), info_a.size() * sizeof(int));
    }
    return DDAFFT_OK;
}

int ddafft_get_dipoles(ddafft_ctx* c, int rhs, float* p) {
    if (!c || !p || rhs < 0 || rhs >= c->nrhs_max) return fail(DDAFFT_EINVAL, "bad argument");
    return field_out(c, c->P + (long long)rhs * 3 * c->d.n, p, c->Y);
}

int ddafft_cabs_cext(ddafft_ctx* c, const float* alpha, const float* kinc, const float* einc, const float* p,
                     float* cabs, float* cext) {
    if (!c || !alpha || !kinc || !einc || !p) return fail(DDAFFT_EINVAL, "NULL argument");
    if (!c->have_target) return fail(DDAFFT_ESTATE, "set_target first");
    TRY(upload(c, c->alpha, alpha, (size_t)c->nmat * sizeof(float2)));
    TRY(upload(c, c->einc, einc, 3 * sizeof(float2)));
    TRY(upload_incidence(c, 1, kinc));
    TRY(field_in(c, p, c->P, c->Y));
    return observables(c, 1, einc, 0, nullptr, 0, nullptr, cabs, cext);
}

int ddafft_far_field(ddafft_ctx* c, const float* p, int ndirs, const float* dirs, float* esc) {
    if (!c || !p || !dirs || !esc || ndirs < 1) return fail(DDAFFT_EINVAL, "bad argument");
    if (!c->have_target) return fail(DDAFFT_ESTATE, "set_target first");
    TRY(ddafft_set_directions(c, ndirs, dirs));
    TRY(field_in(c, p, c->P, c->Y));
    // cross sections are computed alongside; their values are discarded
    std::vector<float> zero_alpha((size_t)2 * c->nmat, 0.f);
    TRY(upload(c, c->alpha, zero_alpha.data(), zero_alpha.size() * sizeof(float)));
    const float kz[3] = {0.f, 0.f, 1.f};
    const float e[6] = {1.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    TRY(upload(c, c->einc, e, sizeof(e)));
    TRY(upload_incidence(c, 1, kz));
    return observables(c, 1, e, ndirs, c->dirs, 0, esc, nullptr, nullptr);
}

int ddafft_profile(ddafft_ctx* c, int enable, double* pass_ms, long long* matvecs) {
    if (!c) return fail(DDAFFT_EINVAL, "NULL context");
    HIPCHECK(hipStreamSynchronize(c->stream));
    TRY(prof_collect(c));
    if (pass_ms)
        for (int k = 0; k < 5; ++k) pass_ms[k] = c->prof_ms[k];
    if (matvecs) *matvecs = c->prof_n;
    if (enable) {
        for (double& v : c->prof_ms) v = 0;
        c->prof_n = 0;
    }
    c->prof = enable != 0;
    return DDAFFT_OK;
}

int ddafft_time_matvec(ddafft_ctx* c, int nrhs, int reps, float* kernel_ms, float* total_ms) {
    if (!c || nrhs < 1 || nrhs > c->nrhs_max || reps < 1) return fail(DDAFFT_EINVAL, "bad argument");
    if (!c->have_target) return fail(DDAFFT_ESTATE, "set_target first");
    hipEvent_t ev[6];
    for (auto& e : ev) HIPCHECK(hipEventCreate(&e));
    ProArgs pa{PRO_COPY, c->S, nullptr, nullptr, nullptr, nullptr};
    EpiArgs ea{};
    ea.kind = EK_MV;
    ea.x = c->S;
    ea.out = c->W;
    ea.alpha = c->alpha;
    ea.nmat = c->nmat;
    double acc[5] = {0, 0, 0, 0, 0};
    double tot = 0;
    // warm-up
    TRY(launch_matvec(c, nrhs, pa, nullptr, 0, ea, 0));
    HIPCHECK(hipStreamSynchronize(c->stream));
    for (int r = 0; r < reps; ++r) {
        TRY(launch_matvec(c, nrhs, pa, nullptr, 0, ea, 0, ev));
        HIPCHECK(hipEventSynchronize(ev[5]));
        for (int k = 0; k < 5; ++k) {
            float ms = 0;
            HIPCHECK(hipEventElapsedTime(&ms, ev[k], ev[k + 1]));
            acc[k] += ms;
        }
        float ms = 0;
        HIPCHECK(hipEventElapsedTime(&ms, ev[0], ev[5]));
        tot += ms;
    }
    for (auto& e : ev) hipEventDestroy(e);
    if (kernel_ms)
        for (int k = 0; k < 5; ++k) kernel_ms[k] = (float)(acc[k] / reps);
    if (total_ms) *total_ms = (float)(tot / reps);
    return DDAFFT_OK;
}

}  // extern "C"
