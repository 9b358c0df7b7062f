// The five passes of one DDA matvec  y = x - alpha * IFFT(Â ⊙ FFT(pad x))
// (dda_func.cpp:150-187 Convolution, :227-437 Convolution_conj, epilogues
// e.g. :1335-1356), for `nrhs` right-hand sides sharing one Â.
//
// HBM layouts (float2 = complex64, C = 3*nrhs component planes):
//   field x, out : [C][ni][nj][nl]            (one plane per component)
//   B1 / B4      : [C][ni][2nl'][nj] / [C][ni][nj][2nl']
//   B2 / B3      : [C][2nj][ni][2nl']         (an X tile's i rows 2nl apart)
//   Â            : [2nj][2nl][2ni][6]         (e00 e01 e02 e11 e12 e22)
// L' is the z-frequency in parity-split order: L' = L / 2 for even L and
// n_l + L / 2 for odd L, the order in which the pruned z transform produces
// them (two n_l-point transforms, fft_func.cpp:447-462); the Z-inverse pass
// puts them back in natural order in LDS.  Lines along l, j and L' are
// contiguous; the z/y passes write and read whole runs of L' in each row J
// of an i-plane, and the X pass reads and writes runs of its tile's L'
// columns in the rows i of one J, 2 n_l float2 apart.
//
//   ZF  z forward, pruned (fft_func.cpp:442-463)      x  -> B1
//   YF  y forward, pruned (fft_func.cpp:484-504)      B1 -> B2
//   X   x forward (:525-545) · Â (dda_func.cpp:157-182) · x inverse
//       (fft_func.cpp:578-589), keeping i < ni         B2 -> B3
//   YI  y inverse (fft_func.cpp:605-616), j < nj       B3 -> B4
//   ZI  z inverse (fft_func.cpp:632-643), l < nl, fused with the solver
//       epilogue and its FP64 partial sums             B4 -> out
//
// Each inverse line transform is conj(FFT(conj(v))) (ifft_line,
// fft_kernel.hpp:301-322).  The conjugation after one axis and the one
// before the next cancel exactly, so X conjugates on entry only, ZI on exit
// only, and B3 and B4 hold the conjugated partial inverse transforms.
#pragma once

#include <type_traits>

#include "device_types.hpp"
#include "fft_common.hpp"
#include "finalize.hpp"

namespace ddafft {

// Tile widths (lines per block).  The z and y passes keep the LDS footprint
// <= 48 KB.  The fused X pass holds 3 components of RB right-hand sides at 2ni
// points; its tile stays small so that several blocks per CU hide the latency
// of its two global phases.  Two right-hand sides at n_i = 64 and 128 take
// four (J, L) columns (measured on MI355X with compile-time tile widths,
// 2 RHS: 64^3 50.3 us at 4 columns, 56.5 at 2, 58.5 at 8; 128^3 496 us at 4,
// 530 at 2); 32^3 keeps 4 (13.2 us, 13.8 at 8); 256-point lines keep two,
// so that the 384 butterflies of a stage fill the block's threads evenly
// (256^3, 2 RHS: 7.16 ms at one column, 6.66 at two).
template <int N> struct TileZF { static constexpr int T = clamp16(pow2floor(6144 / (2 * (N + kLpad)))); };
// Four right-hand sides per tile take half as many columns as two, so the
// tile's LDS and its transforms stay those of the two-right-hand-side tile
// while each Â value serves twice the right-hand sides.
template <int N, int RB> struct TileX {
    static constexpr int T = N >= 512 ? 1
                             : RB == 4 ? (N <= 16 ? clamp16(pow2floor(2048 / (12 * (N + kLpad)))) : N <= 64 ? 2 : 1)
                             : (RB == 2 && (N == 64 || N == 128)) ? 4
                             : N >= 128                           ? 2
                                                                  : clamp16(pow2floor(2048 / (6 * RB * (N + kLpad))));
};
// Line strides (float2) of the X pass's forward (n_i points) and inverse
// (2 n_i points) lines.  With the lane grouping of each stage (LaneGroup),
// they decide how the lanes of a half-wave spread over the LDS banks.
template <int NI, int RB> struct XStride {
    static constexpr int S1 = NI + kLpad, S2 = 2 * NI + kLpad;
};
// n_i = 64, two right-hand sides (four-column tiles: 48 forward lines of 64
// points, 24 inverse lines of 128): strides 65 / 129, with six waves per SIMD
// (XWaves below).  Against 66 / 131 at five waves: 64^3 X pass 47.8 -> 46.3 us
// at 2 RHS, 320 -> 304 us at 16 RHS.
template <> struct XStride<64, 2> { static constexpr int S1 = 65, S2 = 129; };
template <> struct XStride<64, 4> { static constexpr int S1 = 66, S2 = 131; };   // the same 48 / 24 lines
// Waves per SIMD the X pass is compiled for (the VGPR budget: 6 waves leave 80
// registers; 1 places no bound).  n_i = 64 with pairs of right-hand sides, one
// group per block, the reference's Â: its LDS tile (26.5 KB) lets six blocks
// share a CU, and 83 registers allowed five.  (Runs of groups need 177
// registers and the compact Â's mirror 90: bounding them spills.)
template <int NI, int RB, bool RUN, bool COMPACT> struct XWaves { static constexpr int W = 1; };
// Which X-pass stages run in registers on full tiles: FWD, the forward
// transforms' first stage computed from the loads; INV, the inverse
// transforms' last stage stored straight to B3.  Against both in LDS (X pass,
// MI355X): 128^3, 2 RHS 430 -> 382 us; 32^3, 16 RHS 37.7 -> 34.8 us;
// 32x32x64, 16 RHS 72 -> 68 us.  At n_i = 64 with pairs, the forward stage in
// registers needs more than the 80 VGPRs of six waves (spills; 46 -> 51 us),
// so only the inverse one runs there: 64^3 46-49 -> 42.7 us at 2 RHS, 306 ->
// 295 us at 16 RHS.
template <int NI, int RB> struct XRegStages { static constexpr bool FWD = true, INV = true; };
template <> struct XRegStages<64, 2> { static constexpr bool FWD = false, INV = true; };
// Y-inverse lines whose first and last stages run in registers (full
// tiles).  Against LDS throughout (MI355X): 64^3, 2 RHS 17.2 -> 14.9 us;
// 32^3, 16 RHS 16.5 -> 14.0 us; 128^3 (256-point lines) 122 -> 126 us and
// 256^3 (512-point lines) 1287 -> 1468 us, so the longer lines stay in LDS.
template <int NJ2> struct YRegStages { static constexpr bool ON = NJ2 <= 128; };
#ifndef DDA_WAVES_PER_EU   // the CPU emulation of the kernels (tests/emul) defines it empty
#define DDA_WAVES_PER_EU(...) __attribute__((amdgpu_waves_per_eu(__VA_ARGS__)))
#endif
template <> struct XWaves<64, 2, false, false> { static constexpr int W = 6; };
// 12 lines of 128 points at stride 131 (the Z-inverse tile below): a half-wave
// takes 32 consecutive butterflies of one line in the last stage.
template <> struct LaneGroup<128, 12, 131, 2> { static constexpr int G = 32; };
// The Z-inverse pass at 2 n_l = 128 transforms 12 lines per block (three
// components of four rows): that stride and grouping take it from 23.1 to
// 22.7 us per launch (64^3, 2 RHS).
template <int NL2, int T> struct ZStride { static constexpr int S = NL2 + kLpad; };
template <> struct ZStride<128, 4> { static constexpr int S = 131; };
template <int N> struct TileYI { static constexpr int T = clamp16(pow2floor(6144 / (N + kLpad))); };
// The Z-inverse pass runs the solver epilogue as well; at most four (i, j)
// rows per block keeps more blocks in flight (64^3: 16.0 vs 17.1 us at 2 RHS,
// 78.6 vs 87.1 us at 16; 32^3: 8.0 vs 9.9 us at 2 RHS).
template <int N> struct TileZI {
    static constexpr int T = clamp16(pow2floor(6144 / (3 * (N + kLpad)))) < 4 ? clamp16(pow2floor(6144 / (3 * (N + kLpad))))
                                                                               : 4;
};

// ---------------------------------------------------------------------------
// Prologue of the z-forward pass: builds the matvec input from solver vectors
// (fusing the LinearComb that precedes every CopyingBack in DDA_FFT).
// ---------------------------------------------------------------------------

enum ProMode : int {
    PRO_COPY = 0,  // v = A
    PRO_LC1 = 1,   // dipoles: A += c1*B (dda_func.cpp:884-905), v = A
    PRO_LC4 = 2,   // dipoles: B = c2*D + (A + c1*B), else 0 (:843-882), v = B
    PRO_LC2 = 3,   // dipoles: B = A + c1*B (:906-924), v = B
};

struct ProArgs {
    int mode;
    float2* A;
    float2* B;
    const float2* D;
    const float2* P;    // snapshot source (P)
    float2* Pmin;       // snapshot destination (P_min, dda_func.cpp:1305-1313)
};

// ---------------------------------------------------------------------------
// ZF: lines (c, i, j) along l; n_l inputs -> 2 n_l outputs.
// ---------------------------------------------------------------------------

template <int NL>
__global__ __launch_bounds__(kBlock) void k_zf(Dims d, ProArgs pa, const SolveState* __restrict__ st,
                                               const int8_t* __restrict__ dipf, float2* __restrict__ b1,
                                               const float2* __restrict__ wpre, const float2* __restrict__ tw) {
    constexpr int TJ = TileZF<NL>::T;
    constexpr int LS = NL + kLpad;
    constexpr int LNL = ilog2c(NL);
    constexpr int NLD = (TJ * NL + kBlock - 1) / kBlock;
    __shared__ float2 s[2 * TJ * LS];
    __shared__ float2 stw[NL];
    const int c = blockIdx.y, rhs = c / 3;
    float2 c1 = make_float2(0.f, 0.f), c2 = c1;
    int snap = 0;
    if (st) {
        if (st[rhs].it >= st[rhs].stop_at) return;
        c1 = st[rhs].pro_c1;
        c2 = st[rhs].pro_c2;
        // only the first matvec of an iteration carries the P -> P_min copy
        snap = (pa.Pmin != nullptr && pa.P != nullptr) ? st[rhs].snapshot : 0;
    }
    // the tile body, with the tile width a compile-time constant for full
    // tiles and a run-time one otherwise
    auto tile = [&](auto tile_arg) {
        const int tj = tile_arg, ntile = d.nj / tj;
        const int i = blockIdx.x / ntile, j0 = (blockIdx.x - i * ntile) * tj;
        const long long cell0 = ((long long)i * d.nj + j0) * NL;
        const long long cb = (long long)c * d.n + cell0;   // this block's cells of component c
        const int nld = tj * NL;
        // every load of the thread is issued before the first is used: flags and
        // the prologue's operands (the prologue writes only cells it has read)
        bool dip[NLD];
        float2 va[NLD], vb[NLD], vd[NLD], vp[NLD];
        const int mode = pa.mode;
#pragma unroll
        for (int q = 0; q < NLD; ++q) {
            const int f = threadIdx.x + q * kBlock;
            dip[q] = false;
            va[q] = vb[q] = vd[q] = vp[q] = make_float2(0.f, 0.f);
            if (f < nld) {
                dip[q] = dipf[cell0 + f] != 0;
                va[q] = pa.A[cb + f];
                if (mode != PRO_COPY) vb[q] = pa.B[cb + f];
                if (mode == PRO_LC4) vd[q] = pa.D[cb + f];
                if (snap) vp[q] = pa.P[cb + f];   // P is never a prologue output
            }
        }
#pragma unroll
        for (int q = 0; q < NLD; ++q) {
            const int f = threadIdx.x + q * kBlock;
            if (f < nld) {
                float2 v;
                switch (mode) {
                    case PRO_LC1:   // dda_func.cpp:884-905
                        v = va[q];
                        if (dip[q]) {
                            v = cadd(va[q], cmul(c1, vb[q]));
                            pa.A[cb + f] = v;
                        }
                        break;
                    case PRO_LC4:   // :843-882
                        v = dip[q] ? cadd(cmul(c2, vd[q]), cadd(va[q], cmul(c1, vb[q]))) : make_float2(0.f, 0.f);
                        pa.B[cb + f] = v;
                        break;
                    case PRO_LC2:   // :906-924
                        v = vb[q];
                        if (dip[q]) {
                            v = cadd(va[q], cmul(c1, vb[q]));
                            pa.B[cb + f] = v;
                        }
                        break;
                    default:
                        v = va[q];
                }
                if (snap) pa.Pmin[cb + f] = vp[q];
                const int t = f >> LNL, l = f & (NL - 1);
                s[t * LS + l] = v;
                s[(tj + t) * LS + l] = cmul(wpre[l], v);
            }
        }
        for (int f = threadIdx.x; f < NL; f += kBlock) stw[f] = tw[f];
        __syncthreads();
        float2* dst = b1 + ((long long)c * d.ni + i) * (2 * NL) * d.nj + j0;
        const int lnj = __builtin_ctz(d.nj);
        constexpr bool FULL = !std::is_same_v<decltype(tile_arg), int>;
        if constexpr (FULL && Plan<NL>::NST >= 2 && NL <= 128) {
            // the last stage in registers, stored straight to B1: row L' = k
            // of the line's output k (n_l + k for the pre-twiddled line)
            // (64^3, 2 RHS: 14.6 -> 14.2 us; 128^3: 75.6 -> 72 us)
            fft_lds_fixed<NL, 2 * TJ, LS, kBlock, Plan<NL>::NST - 1>(s, stw);
            fft_last_stage_out<NL, 2 * TJ, LS, NL>(s, stw, [&](int line, int k, float2 v) {
                const bool odd = line >= TJ;
                dst[((odd ? NL + k : k) << lnj) + (odd ? line - TJ : line)] = v;
            });
        } else {
            fft_lds<NL, 2 * TJ, LS>(s, 2 * tj, stw);
            const int ltj = __builtin_ctz(tj);
            constexpr int NST = (2 * NL * TJ + kBlock - 1) / kBlock;
#pragma unroll
            for (int q = 0; q < NST; ++q) {
                const int f = threadIdx.x + q * kBlock;
                if (f < 2 * NL * tj) {   // row L' of B1: the even transform's outputs, then the odd one's
                    const int Lp = f >> ltj, t = f - (Lp << ltj);
                    const float2 v = Lp >= NL ? s[(tj + t) * LS + (Lp - NL)] : s[t * LS + Lp];
                    dst[(Lp << lnj) + t] = v;
                }
            }
        }
    };
    if (d.nj >= TJ)
        tile(std::integral_constant<int, TJ>{});
    else
        tile(d.nj);
}

// ---------------------------------------------------------------------------
// ZYF: the z- and y-forward passes of one (c, i) plane in one block, for
// planes of at most 64 x 64 cells (ZF + YF cover the larger ones).  The block
// loads the plane once (with the prologue of ZF, each cell read and written
// by this block only), then for each parity p of the z frequencies: the n_l
// point transforms of its n_j rows in LDS, the result taken into registers
// transposed, and for each parity q of the y frequencies the n_j point
// transforms of the n_l columns, written to B2 as whole rows J = 2 Jh + q
// (columns L' = p n_l + Lh).  The same butterflies, twiddles and order as
// ZF then YF: the same bits.  B1 is not touched.
// ---------------------------------------------------------------------------

template <int NJ, int NL>
struct ZyfShape {
    // threads per block: four plane cells per thread from 32 x 32 up (a 64 x 64
    // plane takes 16 waves), at least 256
    static constexpr int BT = NJ * NL / 4 < 256 ? 256 : NJ * NL / 4 > 1024 ? 1024 : NJ * NL / 4;
    static constexpr int LSZ = NL + kLpad;                      // z lines: NJ rows of NL points
    static constexpr int LSY = NJ + kLpad;                      // y lines: NL columns of NJ points
    static constexpr int WORDS = NJ * LSZ > NL * LSY ? NJ * LSZ : NL * LSY;
    static constexpr int E = (NJ * NL + BT - 1) / BT;   // plane cells per thread
    static_assert(E <= 4, "the plane stays in registers across both parities");
};

template <int NJ, int NL>
__global__ __launch_bounds__((ZyfShape<NJ, NL>::BT)) void k_zyf(Dims d, ProArgs pa, const SolveState* __restrict__ st,
                                                const int8_t* __restrict__ dipf, float2* __restrict__ b2,
                                                const float2* __restrict__ wpre_l, const float2* __restrict__ tw_l,
                                                const float2* __restrict__ wpre_j, const float2* __restrict__ tw_j) {
    using S = ZyfShape<NJ, NL>;
    constexpr int E = S::E, LSZ = S::LSZ, LSY = S::LSY, BT = S::BT;
    constexpr int NP = NJ * NL;
    constexpr int LNL = ilog2c(NL), LNJ = ilog2c(NJ);
    __shared__ float2 w[S::WORDS];
    __shared__ float2 stl[NL], stj[NJ];
    const int i = blockIdx.x, c = blockIdx.y, rhs = c / 3;
    float2 c1 = make_float2(0.f, 0.f), c2 = c1;
    int snap = 0;
    if (st) {
        if (st[rhs].it >= st[rhs].stop_at) return;
        c1 = st[rhs].pro_c1;
        c2 = st[rhs].pro_c2;
        snap = (pa.Pmin != nullptr && pa.P != nullptr) ? st[rhs].snapshot : 0;
    }
    const long long cell0 = (long long)i * NP;            // first cell of the plane
    const long long cb = (long long)c * d.n + cell0;      // its cells of component c
    // the plane, through the prologue (dda_func.cpp:843-924), in chunks of
    // KC cells per thread: the loads of a chunk are issued before the first is
    // used, and a chunk's operands are dead before the next chunk's arrive
    // (all at once they took the kernel to 198 VGPRs at 64 x 64)
    for (int f = threadIdx.x; f < NL; f += BT) stl[f] = tw_l[f];
    for (int f = threadIdx.x; f < NJ; f += BT) stj[f] = tw_j[f];
    float2 v[E];
    constexpr int KC = E < 4 ? E : 4;
    const int mode = pa.mode;
#pragma unroll
    for (int q0 = 0; q0 < E; q0 += KC) {
        bool dip[KC];
        float2 va[KC], vb[KC], vd[KC], vp[KC];
#pragma unroll
        for (int k = 0; k < KC; ++k) {
            const int f = threadIdx.x + (q0 + k) * BT;
            dip[k] = false;
            va[k] = vb[k] = vd[k] = vp[k] = make_float2(0.f, 0.f);
            if (NP % BT == 0 || f < NP) {
                dip[k] = dipf[cell0 + f] != 0;
                va[k] = pa.A[cb + f];
                if (mode != PRO_COPY) vb[k] = pa.B[cb + f];
                if (mode == PRO_LC4) vd[k] = pa.D[cb + f];
                if (snap) vp[k] = pa.P[cb + f];
            }
        }
#pragma unroll
        for (int k = 0; k < KC; ++k) {
            const int q = q0 + k;
            const int f = threadIdx.x + q * BT;
            if (NP % BT == 0 || f < NP) {
                float2 x;
                switch (mode) {
                    case PRO_LC1:
                        x = va[k];
                        if (dip[k]) {
                            x = cadd(va[k], cmul(c1, vb[k]));
                            pa.A[cb + f] = x;
                        }
                        break;
                    case PRO_LC4:
                        x = dip[k] ? cadd(cmul(c2, vd[k]), cadd(va[k], cmul(c1, vb[k]))) : make_float2(0.f, 0.f);
                        pa.B[cb + f] = x;
                        break;
                    case PRO_LC2:
                        x = vb[k];
                        if (dip[k]) {
                            x = cadd(va[k], cmul(c1, vb[k]));
                            pa.B[cb + f] = x;
                        }
                        break;
                    default:
                        x = va[k];
                }
                if (snap) pa.Pmin[cb + f] = vp[k];
                v[q] = x;
            } else {
                v[q] = make_float2(0.f, 0.f);
            }
        }
        __builtin_amdgcn_sched_barrier(0);   // keep the next chunk's loads behind this chunk's use
    }
    // row J of plane (c, i) of B2 [C][J][i][L'] at dst + J * rowj
    float2* dst = b2 + ((long long)c * (2 * NJ) * d.ni + i) * (2 * NL);
    const int rowj = d.ni * (2 * NL);
#pragma unroll
    for (int p = 0; p < 2; ++p) {
        if (p) __syncthreads();   // the previous parity's last reads of w are done
        // rows j along l: x (p = 0) or W_2nl^l x (p = 1), as ZF forms them
#pragma unroll
        for (int q = 0; q < E; ++q) {
            const int f = threadIdx.x + q * BT;
            if (NP % BT == 0 || f < NP) {
                const int j = f >> LNL, l = f & (NL - 1);
                w[j * LSZ + l] = p ? cmul(wpre_l[l], v[q]) : v[q];
            }
        }
        __syncthreads();
        fft_lds_fixed<NL, NJ, LSZ, BT>(w, stl);
        // Z_p[j][Lh] into registers, transposed: element f = Lh * NJ + j
        float2 u[E];
#pragma unroll
        for (int q = 0; q < E; ++q) {
            const int f = threadIdx.x + q * BT;
            u[q] = make_float2(0.f, 0.f);
            if (NP % BT == 0 || f < NP) u[q] = w[(f & (NJ - 1)) * LSZ + (f >> LNJ)];
        }
#pragma unroll
        for (int qy = 0; qy < 2; ++qy) {
            __syncthreads();   // every read of w above (or of the previous qy's output) is done
            // columns Lh along j: Z (qy = 0) or W_2nj^j Z (qy = 1), as YF forms them
#pragma unroll
            for (int q = 0; q < E; ++q) {
                const int f = threadIdx.x + q * BT;
                if (NP % BT == 0 || f < NP) {
                    const int lh = f >> LNJ, j = f & (NJ - 1);
                    w[lh * LSY + j] = qy ? cmul(wpre_j[j], u[q]) : u[q];
                }
            }
            __syncthreads();
            fft_lds_fixed<NJ, NL, LSY, BT>(w, stj);
            // row J = 2 Jh + qy of the plane, columns p NL .. p NL + NL - 1
#pragma unroll
            for (int q = 0; q < E; ++q) {
                const int f = threadIdx.x + q * BT;
                if (NP % BT == 0 || f < NP) {
                    const int jh = f >> LNL, lh = f & (NL - 1);
                    dst[(2 * jh + qy) * rowj + p * NL + lh] = w[lh * LSY + jh];
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// YF: lines (c, i, L') along j; n_j -> 2 n_j.  A tile: consecutive L' of one
// i-plane, written as runs of L' in each row J of B2.
// ---------------------------------------------------------------------------

template <int NJ>
__global__ __launch_bounds__(kBlock) void k_yf(Dims d, const SolveState* __restrict__ st,
                                               const float2* __restrict__ b1, float2* __restrict__ b2,
                                               const float2* __restrict__ wpre, const float2* __restrict__ tw) {
    constexpr int TI = TileZF<NJ>::T;
    constexpr int LS = NJ + kLpad;
    __shared__ float2 s[2 * TI * LS];
    __shared__ float2 stw[NJ];
    const int c = blockIdx.y, rhs = c / 3;
    if (st && st[rhs].it >= st[rhs].stop_at) return;
    constexpr int LNJ = ilog2c(NJ);
    const int NL2 = 2 * d.nl;
    // the tile body, with the tile width a compile-time constant for full
    // tiles and a run-time one otherwise
    auto tile = [&](auto tile_arg) {
        const int ti = tile_arg, nlt = NL2 / ti;
        const int i = blockIdx.x / nlt, L0 = (blockIdx.x - i * nlt) * ti;
        const int lnl2 = __builtin_ctz(NL2), lti = __builtin_ctz(ti);
        {
            constexpr int NLD = (TI * NJ + kBlock - 1) / kBlock;
            const float2* src = b1 + (((long long)c * d.ni + i) * NL2 + L0) * NJ;   // ti contiguous lines
            float2 v[NLD];
#pragma unroll
            for (int q = 0; q < NLD; ++q) {
                const int f = threadIdx.x + q * kBlock;
                if (f < ti * NJ) v[q] = src[f];
            }
#pragma unroll
            for (int q = 0; q < NLD; ++q) {
                const int f = threadIdx.x + q * kBlock;
                if (f < ti * NJ) {
                    const int t = f >> LNJ, jj = f & (NJ - 1);
                    s[t * LS + jj] = v[q];
                    s[(ti + t) * LS + jj] = cmul(wpre[jj], v[q]);
                }
            }
        }
        for (int f = threadIdx.x; f < NJ; f += kBlock) stw[f] = tw[f];
        __syncthreads();
        // row J of plane (c, i) of B2 [C][J][i][L'], from L0
        float2* dst = b2 + ((long long)c * (2 * NJ) * d.ni + i) * NL2 + L0;
        const int rowj = d.ni << lnl2;
        constexpr bool FULL = !std::is_same_v<decltype(tile_arg), int>;
        if constexpr (FULL && Plan<NJ>::NST >= 2 && NJ <= 128) {
            // the last stage in registers, stored straight to B2: row
            // J = 2 k (+1 for the pre-twiddled line) of output k (64^3, 2 RHS:
            // 16.3 -> 15.9 us; 256-point lines stay in LDS: 1270 -> 1315 us)
            fft_lds_fixed<NJ, 2 * TI, LS, kBlock, Plan<NJ>::NST - 1>(s, stw);
            fft_last_stage_out<NJ, 2 * TI, LS, NJ>(s, stw, [&](int line, int k, float2 v) {
                const bool odd = line >= TI;
                dst[(2 * k + (odd ? 1 : 0)) * rowj + (odd ? line - TI : line)] = v;
            });
        } else {
            fft_lds<NJ, 2 * TI, LS>(s, 2 * ti, stw);
            constexpr int NST = (2 * NJ * TI + kBlock - 1) / kBlock;
#pragma unroll
            for (int q = 0; q < NST; ++q) {
                const int f = threadIdx.x + q * kBlock;
                if (f < 2 * NJ * ti) {
                    const int J = f >> lti, t = f - (J << lti);
                    const float2 v = (J & 1) ? s[(ti + t) * LS + (J >> 1)] : s[t * LS + (J >> 1)];
                    dst[J * rowj + t] = v;
                }
            }
        }
    };
    if (NL2 >= TI)
        tile(std::integral_constant<int, TI>{});
    else
        tile(NL2);
}

// ---------------------------------------------------------------------------
// X: lines (J, L) along i for the 3 components of RB right-hand sides, which
// share each Â value read from HBM.
// ---------------------------------------------------------------------------

// y = Â x for one padded cell (dda_func.cpp:163-181: c3 = c0 + c1, out = c2 + c3)
__device__ __forceinline__ void ahat_apply(const float2* a, float2 x, float2 y, float2 z, float2& ox, float2& oy,
                                           float2& oz) {
    ox = cadd(cmul(a[2], z), cadd(cmul(a[0], x), cmul(a[1], y)));
    oy = cadd(cmul(a[4], z), cadd(cmul(a[1], x), cmul(a[3], y)));
    oz = cadd(cmul(a[5], z), cadd(cmul(a[2], x), cmul(a[4], y)));
}

// Right-hand sides [rhs0, rhs1) that exist and are still running, as bits
// from rhs0 (at most 64 right-hand sides per context).
__device__ __forceinline__ unsigned long long x_active(const Dims& d, const SolveState* __restrict__ st, int rhs0,
                                                       int rhs1) {
    unsigned long long m = 0;
    for (int r = rhs0; r < rhs1 && r < d.nrhs; ++r)
        if (!(st && st[r].it >= st[r].stop_at)) m |= 1ull << (r - rhs0);
    return m;
}

// One block's tile of the X pass, for the right-hand-side groups
// [g0, g1) of RB each (RUN; otherwise the one group g0); bit r of amask:
// right-hand side g0 * RB + r is running.  Â of the tile is read once,
// before the first group, and kept in registers for all of them.  The
// single-group form is straight-line code: a loop over groups keeps the
// per-group address arithmetic live across it (195 against 76 VGPRs at
// n_i = 64), so it is compiled separately.  FULL: the tile holds TJ columns
// (every lattice with 2 n_j >= TJ), so the tile width is a compile-time
// constant and the index arithmetic reduces to shifts and masks; otherwise
// tj = 2 n_j.
template <int NI, int RB, bool FULL, bool RUN, bool COMPACT>
__device__ __forceinline__ void x_tile(const Dims& d, unsigned long long amask, int g0, int g1, int tj_small, const float2* __restrict__ b2, float2* __restrict__ b3,
                                       const float2* __restrict__ ahat, int conj, const float2* __restrict__ wpre,
                                       const float2* __restrict__ tw1, const float2* __restrict__ tw2, float2* s,
                                       float2* stw1, float2* stw2) {
    // conj: the conjugate convolution (Â at -k, conjugated).  COMPACT: Â
    // stored as one octant (the reference's padded A(r) with its Nyquist
    // planes zeroed, so that Â is even or odd along every axis).
    constexpr int CB = 3 * RB;                       // component lines per (J, L)
    constexpr int TJ = TileX<NI, RB>::T;
    constexpr int NI2 = 2 * NI;
    constexpr int LS1 = XStride<NI, RB>::S1, LS2 = XStride<NI, RB>::S2;   // forward / inverse line strides
    constexpr int ITEMS = (TJ * NI2 + kBlock - 1) / kBlock;
    const int NJ2 = 2 * d.nj, NL2 = 2 * d.nl;
    const int tj = FULL ? TJ : tj_small, nlt = NL2 / tj;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    // columns (J, L0 .. L0 + tj - 1) in the parity-split order L'
    const int J = tile / nlt, L0 = (tile - J * nlt) * tj;

    // Â of this thread's points, requested first so that its latency overlaps
    // the field loads and the forward transform
    float2 a[ITEMS][6];
#pragma unroll
    for (int it = 0; it < ITEMS; ++it) {
        const int f = threadIdx.x + it * kBlock;
        if (f < tj * NI2) {
            const int I = f % NI2, t = f / NI2;
            const int Lp = L0 + t;
            unsigned Ja = J, La = Lp < d.nl ? 2 * Lp : 2 * (Lp - d.nl) + 1, Ia = I;   // natural L
            if (conj) {
                Ja = (NJ2 - Ja) & (NJ2 - 1);
                La = (NL2 - La) & (NL2 - 1);
                Ia = (NI2 - Ia) & (NI2 - 1);
            }
            // < 6 * 8 N <= 6 * 2^27: a 32-bit offset from the base
            unsigned cell = (Ja * NL2 + La) * NI2 + Ia;
            bool fx = false, fy = false, fz = false;
            if constexpr (COMPACT) {
                // one octant [J <= nj][L <= nl][I <= NI]: a frequency above the
                // Nyquist index reads its mirror, the components odd along the
                // mirrored axes negated
                const unsigned nj = d.nj, nl = d.nl;
                fx = Ia > (unsigned)NI;
                fy = Ja > nj;
                fz = La > nl;
                const unsigned Jc = fy ? NJ2 - Ja : Ja, Lc = fz ? NL2 - La : La, Ic = fx ? NI2 - Ia : Ia;
                cell = (Jc * (nl + 1) + Lc) * (NI + 1) + Ic;
            }
            const float4* ap = reinterpret_cast<const float4*>(ahat + cell * 6u);
            const float4 q0 = ap[0], q1 = ap[1], q2 = ap[2];
            a[it][0] = make_float2(q0.x, q0.y);
            a[it][1] = make_float2(q0.z, q0.w);
            a[it][2] = make_float2(q1.x, q1.y);
            a[it][3] = make_float2(q1.z, q1.w);
            a[it][4] = make_float2(q2.x, q2.y);
            a[it][5] = make_float2(q2.z, q2.w);
            if constexpr (COMPACT) {
                if (fx != fy) a[it][1] = cneg(a[it][1]);   // xy
                if (fx != fz) a[it][2] = cneg(a[it][2]);   // xz
                if (fy != fz) a[it][4] = cneg(a[it][4]);   // yz
            }
            if (conj)
#pragma unroll
                for (int e = 0; e < 6; ++e) a[it][e].y = -a[it][e].y;
        }
    }
    const int ltj = __builtin_ctz(tj);
    constexpr int LNI = ilog2c(NI);
    bool staged = false;   // twiddles in LDS
    const int gend = RUN ? g1 : g0 + 1;
    for (int grp = g0; grp < gend; ++grp) {
        const int rhs0 = grp * RB;
        const unsigned gm = (unsigned)(amask >> ((grp - g0) * RB)) & ((1u << RB) - 1u);
        if (!gm) continue;
        bool act[RB];
#pragma unroll
        for (int r = 0; r < RB; ++r) act[r] = (gm >> r) & 1u;
        if (grp > g0) __syncthreads();   // the previous group's stores have read the tile
        // twiddles staged after the first field loads are issued, so that the
        // Â and field loads are in flight together; strides of the
        // compile-time block size (reading blockDim made the compiler wait for
        // every load in flight first)
        auto stage_twiddles = [&]() {
            if (!staged) {
                for (int f = threadIdx.x; f < NI; f += kBlock) stw1[f] = tw1[f];
                for (int f = threadIdx.x; f < NI2; f += kBlock) stw2[f] = tw2[f];
                staged = true;
            }
        };
        // runs of tj columns in row J of the (cc, i) planes: B2 [C][J][i][L']
        const float2* src = b2 + ((long long)(rhs0 * 3) * NJ2 + J) * NI * NL2 + L0;
        const int lcomp = __builtin_ctz(NJ2) + LNI + __builtin_ctz(NL2), lnl2 = __builtin_ctz(NL2);
        if constexpr (FULL && XRegStages<NI, RB>::FWD) {
            // Stage 0 of the forward transforms in registers, straight from the
            // loads.  Item f = (column col = cc TJ + t, stage-0 butterfly b)
            // loads the R0 inputs of b (rows q NI/R0 + rest(b), the digit-
            // reversed gather of the LDS transform) from its column, runs b on
            // them for the column's line x and for its pre-twiddled twin
            // W_2ni^i x, and writes both results where stage 1 reads them.  The
            // same butterflies on the same values as a gather from a natural-
            // order tile, without that tile's LDS round trip and barrier.
            constexpr int R0 = plan_radix(Plan<NI>::K, 0), NB0 = NI / R0, NCOL = CB * TJ;
            constexpr int NITEM = NCOL * NB0, IPT = (NITEM + kBlock - 1) / kBlock;
            float2 xv[IPT][R0];
#pragma unroll
            for (int it = 0; it < IPT; ++it) {
                const int f = threadIdx.x + it * kBlock;
#pragma unroll
                for (int q = 0; q < R0; ++q) xv[it][q] = make_float2(0.f, 0.f);
                if (NITEM % kBlock == 0 || f < NITEM) {
                    const int col = f % NCOL, bf = f / NCOL;
                    const int cc = col / TJ, t = col % TJ, i0 = rest_natural<NI>(bf);
                    if (act[cc / 3])
#pragma unroll
                        for (int q = 0; q < R0; ++q) xv[it][q] = src[(cc << lcomp) + ((q * NB0 + i0) << lnl2) + t];
                }
            }
            stage_twiddles();
#pragma unroll
            for (int it = 0; it < IPT; ++it) {
                const int f = threadIdx.x + it * kBlock;
                if (NITEM % kBlock == 0 || f < NITEM) {
                    const int col = f % NCOL, bf = f / NCOL, i0 = rest_natural<NI>(bf);
                    float2 z[R0];
#pragma unroll
                    for (int q = 0; q < R0; ++q) z[q] = xv[it][q];
                    butterfly<R0>(z);
                    float2* v = s + (2 * col) * LS1 + bf * R0;
#pragma unroll
                    for (int q = 0; q < R0; ++q) v[q] = z[q];
#pragma unroll
                    for (int q = 0; q < R0; ++q) z[q] = cmul(wpre[q * NB0 + i0], xv[it][q]);
                    butterfly<R0>(z);
#pragma unroll
                    for (int q = 0; q < R0; ++q) v[LS1 + q] = z[q];
                }
            }
            __syncthreads();
            FixedStage<NI, 2 * NCOL, LS1, 1>::run(s, stw1);
        } else {
            {
                // all loads of this thread are issued before the first is used;
                // the block's lines lie at 32-bit offsets from one 64-bit base
                constexpr int NLD = (CB * TJ * NI + kBlock - 1) / kBlock;
                const int nld = CB * tj * NI;
                float2 v[NLD];
#pragma unroll
                for (int q = 0; q < NLD; ++q) {
                    const int f = threadIdx.x + q * kBlock;
                    v[q] = make_float2(0.f, 0.f);
                    if (f < nld) {
                        const int t = f & (tj - 1), r = f >> ltj, cc = r >> LNI;   // r = cc * NI + i
                        if (act[cc / 3]) v[q] = src[(cc << lcomp) + ((r & (NI - 1)) << lnl2) + t];
                    }
                }
                stage_twiddles();
#pragma unroll
                for (int q = 0; q < NLD; ++q) {
                    const int f = threadIdx.x + q * kBlock;
                    if (f < nld) {
                        const int t = f & (tj - 1), r = f >> ltj, ii = r & (NI - 1), cc = r >> LNI;
                        const int line = (cc * tj + t) * 2;
                        s[line * LS1 + ii] = v[q];
                        s[(line + 1) * LS1 + ii] = cmul(wpre[ii], v[q]);
                    }
                }
            }
            __syncthreads();
            fft_lds<NI, 2 * CB * TJ, LS1>(s, 2 * CB * tj, stw1);
        }

        // Â multiply in natural order of I, one Â read for every right-hand
        // side; results land conjugated, in natural order, for the inverse
        // transform (ifft_line, fft_kernel.hpp:301-322).
        float2 y[ITEMS][CB];
#pragma unroll
        for (int it = 0; it < ITEMS; ++it) {
            const int f = threadIdx.x + it * kBlock;
            if (f < tj * NI2) {
                const int I = f % NI2, t = f / NI2;
#pragma unroll
                for (int r = 0; r < RB; ++r) {
                    float2 xv[3];
#pragma unroll
                    for (int comp = 0; comp < 3; ++comp)
                        xv[comp] = s[(((r * 3 + comp) * tj + t) * 2 + (I & 1)) * LS1 + (I >> 1)];
                    ahat_apply(a[it], xv[0], xv[1], xv[2], y[it][r * 3], y[it][r * 3 + 1], y[it][r * 3 + 2]);
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int it = 0; it < ITEMS; ++it) {
            const int f = threadIdx.x + it * kBlock;
            if (f < tj * NI2) {
                const int I = f % NI2, t = f / NI2;
#pragma unroll
                for (int cc = 0; cc < CB; ++cc) s[(cc * tj + t) * LS2 + I] = cconj(y[it][cc]);
            }
        }
        __syncthreads();
        // B3 [C][J][i][L'], as B2
        float2* dst = b3 + ((long long)(rhs0 * 3) * NJ2 + J) * NI * NL2 + L0;
        constexpr int NSTI = Plan<NI2>::NST;
        if constexpr (FULL && NSTI >= 2 && XRegStages<NI, RB>::INV) {
            // every stage but the last in LDS; the last in registers, its
            // outputs at positions below n_i (the ones the pruned inverse
            // keeps) stored straight to B3: the same butterflies and twiddles
            // as the LDS stage, without its LDS writes, barrier and read-back
            fft_lds_fixed<NI2, CB * TJ, LS2, kBlock, NSTI - 1>(s, stw2);
            // (inline rather than fft_last_stage_out: through the helper the
            // compiler takes 83 VGPRs instead of 73 at n_i = 32)
            constexpr int RL = plan_radix(Plan<NI2>::K, NSTI - 1), HL = NI2 / RL, NLI = CB * TJ;
            constexpr int WL = NLI * HL, ITL = (WL + kBlock - 1) / kBlock;
#pragma unroll
            for (int it = 0; it < ITL; ++it) {
                const int f = threadIdx.x + it * kBlock;
                if (WL % kBlock == 0 || f < WL) {
                    const int line = f % NLI, bf = f / NLI;   // butterfly bf of the last stage: span NI2, step 1
                    const int cc = line / TJ, t = line % TJ;
                    const float2* v = s + line * LS2 + bf;
                    float2 z[RL];
#pragma unroll
                    for (int q = 0; q < RL; ++q) z[q] = v[q * HL];
                    if (RL > 2 && bf > 0) {
                        int tq = 0;
#pragma unroll
                        for (int q = 1; q < RL; ++q) {
                            tq += bf;   // q * bf * step
                            const float2 tv = stw2[tq];
                            const float tr = z[q].x * tv.x - z[q].y * tv.y;
                            z[q].y = z[q].x * tv.y + z[q].y * tv.x;
                            z[q].x = tr;
                        }
                    }
                    butterfly<RL>(z);
                    if (act[cc / 3])
#pragma unroll
                        for (int q = 0; q < RL; ++q)
                            if (q * HL < NI) dst[(cc << lcomp) + ((bf + q * HL) << lnl2) + t] = z[q];
                }
            }
        } else {
            fft_lds<NI2, CB * TJ, LS2>(s, CB * tj, stw2);
            constexpr int NST = (CB * TJ * NI + kBlock - 1) / kBlock;
            const int nst = CB * tj * NI;
#pragma unroll
            for (int q = 0; q < NST; ++q) {
                const int f = threadIdx.x + q * kBlock;
                if (f < nst) {
                    const int t = f & (tj - 1), r = f >> ltj, i = r & (NI - 1), cc = r >> LNI;
                    if (act[cc / 3]) dst[(cc << lcomp) + (i << lnl2) + t] = s[(cc * tj + t) * LS2 + i];
                }
            }
        }
    }
}

// Grid: x over (J, L) tiles, y over sets of `gpb` consecutive right-hand-side
// groups of RB (RUN; otherwise gpb = 1); each block reads its tile's Â once
// for all of its groups.
template <int NI, int RB, bool RUN, bool COMPACT>
__global__ __launch_bounds__(kBlock) DDA_WAVES_PER_EU(XWaves<NI, RB, RUN, COMPACT>::W) void k_x(Dims d, const SolveState* __restrict__ st,
                                              const float2* __restrict__ b2, float2* __restrict__ b3,
                                              const float2* __restrict__ ahat, int conj,
                                              const float2* __restrict__ wpre, const float2* __restrict__ tw1,
                                              const float2* __restrict__ tw2, int gpb) {
    static_assert(RB == 1 || RB == 2 || RB == 4, "one, two or four right-hand sides per X tile");
    constexpr int TJ = TileX<NI, RB>::T;
    constexpr int LS1 = XStride<NI, RB>::S1, LS2 = XStride<NI, RB>::S2;
    static_assert(LS2 <= 2 * LS1, "the inverse lines fit the forward tile's LDS");
    __shared__ float2 s[3 * RB * TJ * 2 * LS1];
    __shared__ float2 stw1[NI];
    __shared__ float2 stw2[2 * NI];
    const int ngroups = (d.nrhs + RB - 1) / RB;
    const int g0 = blockIdx.y * gpb, g1 = RUN ? min(g0 + gpb, ngroups) : g0 + 1;
    // which right-hand sides of the run are running, read before any other
    // load is issued (so that no later wait on them waits for Â as well)
    const unsigned long long amask = x_active(d, st, g0 * RB, g1 * RB);
    if (!amask) return;
    // Measured on MI355X (2 RHS), against a run-time tile width: n_i = 64
    // 57.8 -> 56.3 us per launch, n_i = 128 612 -> 534 us, n_i = 256
    // 6571 -> 5688 us.
    const int tj = min(TJ, 2 * d.nl);
    if (tj == TJ)
        x_tile<NI, RB, true, RUN, COMPACT>(d, amask, g0, g1, TJ, b2, b3, ahat, conj, wpre, tw1, tw2, s, stw1, stw2);
    else
        x_tile<NI, RB, false, RUN, COMPACT>(d, amask, g0, g1, tj, b2, b3, ahat, conj, wpre, tw1, tw2, s, stw1, stw2);
}

// ---------------------------------------------------------------------------
// YI: lines (c, i, L') along J; 2 n_j -> j < n_j.
// ---------------------------------------------------------------------------

template <int NJ2>
__global__ __launch_bounds__(kBlock) void k_yi(Dims d, const SolveState* __restrict__ st,
                                               const float2* __restrict__ b3, float2* __restrict__ b4,
                                               const float2* __restrict__ tw) {
    constexpr int TL = TileYI<NJ2>::T;
    constexpr int LS = NJ2 + kLpad;
    __shared__ float2 s[TL * LS];
    __shared__ float2 stw[NJ2];
    const int c = blockIdx.y, rhs = c / 3;
    if (st && st[rhs].it >= st[rhs].stop_at) return;
    const int NL2 = 2 * d.nl;
    // the tile body, with the tile width a compile-time constant for full
    // tiles and a run-time one otherwise
    auto tile = [&](auto tile_arg) {
        const int tl = tile_arg, nlt = NL2 / tl;
        const int i = blockIdx.x / nlt, L0 = (blockIdx.x - i * nlt) * tl;
        const int ltl0 = __builtin_ctz(tl), lnl20 = __builtin_ctz(NL2);
        // columns L0 .. L0 + tl - 1 of plane (c, i) of B3 [C][J][i][L']: runs of
        // tl in every row J
        const float2* src = b3 + ((long long)c * NJ2 * d.ni + i) * NL2 + L0;
        const int rowj = d.ni << lnl20;
        constexpr int NSTJ = Plan<NJ2>::NST;
        constexpr bool FULL = !std::is_same_v<decltype(tile_arg), int>;
        if constexpr (FULL && NSTJ >= 2 && YRegStages<NJ2>::ON) {
            // first stage in registers from the loads (item (t, b): the R0 rows
            // of stage-0 butterfly b in column t, the LDS gather's digit-
            // reversed order), the middle stages in LDS, the last stage in
            // registers with its kept outputs (j < n_j) stored to B4: the same
            // butterflies and twiddles as the transform in LDS, with two LDS
            // round trips and a barrier fewer
            constexpr int R0 = plan_radix(Plan<NJ2>::K, 0), NB0 = NJ2 / R0;
            constexpr int NITEM = TL * NB0, IPT = (NITEM + kBlock - 1) / kBlock;
            float2 xv[IPT][R0];
#pragma unroll
            for (int it = 0; it < IPT; ++it) {
                const int f = threadIdx.x + it * kBlock;
                if (NITEM % kBlock == 0 || f < NITEM) {
                    const int t = f % TL, bf = f / TL, j0 = rest_natural<NJ2>(bf);
#pragma unroll
                    for (int q = 0; q < R0; ++q) xv[it][q] = src[(q * NB0 + j0) * rowj + t];
                }
            }
            for (int f = threadIdx.x; f < NJ2; f += kBlock) stw[f] = tw[f];
#pragma unroll
            for (int it = 0; it < IPT; ++it) {
                const int f = threadIdx.x + it * kBlock;
                if (NITEM % kBlock == 0 || f < NITEM) {
                    const int t = f % TL, bf = f / TL;
                    butterfly<R0>(xv[it]);
                    float2* v = s + t * LS + bf * R0;
#pragma unroll
                    for (int q = 0; q < R0; ++q) v[q] = xv[it][q];
                }
            }
            __syncthreads();
            FixedStage<NJ2, TL, LS, 1, kBlock, NSTJ - 1>::run(s, stw);
            constexpr int NJ = NJ2 / 2;
            const int lnl2 = __builtin_ctz(NL2);
            float2* dst = b4 + (((long long)c * d.ni + i) * NJ) * NL2 + L0;
            fft_last_stage_out<NJ2, TL, LS, NJ>(s, stw, [&](int t, int j, float2 v) { dst[(j << lnl2) + t] = v; });
        } else {
            {
                constexpr int NLD = (TL * NJ2 + kBlock - 1) / kBlock;
                float2 v[NLD];
#pragma unroll
                for (int q = 0; q < NLD; ++q) {
                    const int f = threadIdx.x + q * kBlock;
                    if (f < tl * NJ2) v[q] = src[(f >> ltl0) * rowj + (f & (tl - 1))];
                }
#pragma unroll
                for (int q = 0; q < NLD; ++q) {
                    const int f = threadIdx.x + q * kBlock;
                    if (f < tl * NJ2) s[(f & (tl - 1)) * LS + (f >> ltl0)] = v[q];   // line t, point J
                }
            }
            for (int f = threadIdx.x; f < NJ2; f += kBlock) stw[f] = tw[f];
            __syncthreads();
            fft_lds<NJ2, TL, LS>(s, tl, stw);
            constexpr int NJ = NJ2 / 2;
            const int ltl = __builtin_ctz(tl), lnl2 = __builtin_ctz(NL2);
            float2* dst = b4 + (((long long)c * d.ni + i) * NJ) * NL2 + L0;
            constexpr int NST = (TL * NJ + kBlock - 1) / kBlock;
#pragma unroll
            for (int q = 0; q < NST; ++q) {
                const int f = threadIdx.x + q * kBlock;
                if (f < tl * NJ) {
                    const int j = f >> ltl, t = f - (j << ltl);
                    dst[(j << lnl2) + t] = s[t * LS + j];
                }
            }
        }
    };
    if (NL2 >= TL)
        tile(std::integral_constant<int, TL>{});
    else
        tile(NL2);
}

// ---------------------------------------------------------------------------
// ZI: lines (i, j) along L for the 3 components of one right-hand side;
// 2 n_l -> l < n_l, then the epilogue of the calling solver step.
// ---------------------------------------------------------------------------

enum EpiKind : int {
    EK_CONV = 0,      // out = conv                                  (Convolution)
    EK_MV = 1,        // out = x - a*conv on dipoles, else 0
    EK_BICG_R0 = 2,   // R = S = G = -(a*conv); dot1 = (R,R)         (:1114-1135, 1178-1183)
    EK_BICG_H = 3,    // out = x - a*conv; dot1 = (G,out)            (:1192-1216)
    EK_BICG_W = 4,    // out = x - a*conv; dot1 = (out,x); dot2 = (out,out)  (:1231-1258)
    EK_CG_R0 = 5,     // R = S = -(a*conv); err(R)                   (:1114-1138, 1280)
    EK_CG_W = 6,      // out = x - a*conv; dot1 = (x,out); dot2 = (y,x)       (:1325-1361)
    EK_AA_R0 = 7,     // R = -(a*conv); real1 = |R|^2               (:1114-1141)
    EK_AA_S0 = 8,     // conj: out = x - conj(a)*conv; err(x)        (:1143-1176, 1280)
    EK_AA_W = 9,      // out = x - a*conv; real1 = |x|^2             (:1325-1374)
    EK_AA_SUPD = 10,  // conj: W = x - conj(a)*conv; S = W + rB*S on dipoles  (:1386-1416)
    EK_SIM_R0 = 11,   // R = -(a*conv); S = x; P = S - R; err(R)     (:1266-1278)
    EK_SIM = 12,      // W = x - a*conv; R = W - S; P = x - tau*R; err(R)     (:1471-1521)
};


struct EpiArgs {
    int kind;
    int conj;
    const float2* x;     // matvec input
    float2* out;         // primary output
    float2* out2;        // secondary output (S, or P)
    float2* out3;        // tertiary output (G)
    const float2* y;     // extra input (G or R or S)
    const float2* alpha; // [nrhs][nmat]
    int nmat;
    float tau;
    double* partials;    // [nrhs][gridDim.x][kSlots]
};

// ldcScalMul term (dda_func.cpp:734-750): float products, summed wider
__device__ __forceinline__ void dot_acc(float2 a, float2 b, double& re, double& im) {
    re += (double)(a.x * b.x - a.y * b.y);
    im += (double)(a.x * b.y + a.y * b.x);
}
// ldScalMulConj / ldError term (dda_func.cpp:767-775, 796-805)
__device__ __forceinline__ double norm_term(float2 a) { return (double)(a.x * a.x + a.y * a.y); }

template <int NL2>
__global__ __launch_bounds__(kBlock) void k_zi(Dims d, SolveState* __restrict__ st,
                                               const float2* __restrict__ b4, EpiArgs ea,
                                               const int8_t* __restrict__ occ, const int8_t* __restrict__ dipf,
                                               const float2* __restrict__ tw) {
    constexpr int TJ = TileZI<NL2>::T;
    constexpr int LS = ZStride<NL2, TJ>::S;
    constexpr int NLH = NL2 / 2;
    constexpr int LNL2 = ilog2c(NL2), LNLH = LNL2 - 1;
    constexpr int NLD = (3 * TJ * NL2 + kBlock - 1) / kBlock;
    constexpr int NEP = (TJ * NLH + kBlock - 1) / kBlock;
    __shared__ float2 s[3 * TJ * LS];
    __shared__ float2 stw[NL2];
    __shared__ double red[kSlots * 16];
    const int rhs = blockIdx.y;
    float rB = 0.f;
    if (st) {
        if (st[rhs].it >= st[rhs].stop_at) return;
        rB = st[rhs].rB;
    }
    // the tile body, with the row count a compile-time constant for full tiles
    // (nj >= TJ) and a run-time one otherwise
    auto tile = [&](auto tj_arg) {
        const int tj = tj_arg, njt = d.nj / tj;
        const int i = blockIdx.x / njt, j0 = (blockIdx.x - i * njt) * tj;
        const int ltj = __builtin_ctz(tj), lnj = __builtin_ctz(d.nj), lni = __builtin_ctz(d.ni);
        {
            const float2* src = b4 + (((long long)rhs * 3 * d.ni + i) * d.nj + j0) * NL2;
            const int lcs = lni + lnj + LNL2;   // one component plane of B4
            const int nld = 3 * tj * NL2;
            float2 v[NLD];
#pragma unroll
            for (int q = 0; q < NLD; ++q) {
                const int f = threadIdx.x + q * kBlock;
                if (f < nld) {
                    const int LL = f & (NL2 - 1), r = f >> LNL2, t = r & (tj - 1), comp = r >> ltj;
                    v[q] = src[(comp << lcs) + (t << LNL2) + LL];
                }
            }
#pragma unroll
            for (int q = 0; q < NLD; ++q) {
                const int f = threadIdx.x + q * kBlock;
                if (f < nld) {   // line comp * tj + t; L' back to natural order
                    const int Lp = f & (NL2 - 1);
                    s[(f >> LNL2) * LS + (Lp < NLH ? 2 * Lp : 2 * (Lp - NLH) + 1)] = v[q];
                }
            }
        }
        for (int f = threadIdx.x; f < NL2; f += kBlock) stw[f] = tw[f];
        __syncthreads();
        fft_lds<NL2, 3 * TJ, LS>(s, 3 * tj, stw);

        // epilogue operands of every cell of this thread are loaded before the
        // first is used (each cell's outputs overwrite only its own operands)
        const int kind = ea.kind;
        const bool need_x = kind != EK_CONV && kind != EK_BICG_R0 && kind != EK_CG_R0 && kind != EK_AA_R0;
        const bool need_y = kind == EK_BICG_H || kind == EK_CG_W || kind == EK_SIM;
        const bool need_o2 = kind == EK_AA_SUPD;
        const float2* al = ea.alpha + (long long)rhs * ea.nmat;
        const long long rb = (long long)rhs * 3 * d.n;
        const long long k0 = ((long long)i * d.nj + j0) * NLH;   // first cell of the block
        const int nep = tj * NLH;
        // off the dipoles only the Simple solver's first step reads x (every
        // other epilogue writes zeros there or leaves the cell): the operand
        // loads wait on the flags as the polarizability waits on occ
        const bool x_everywhere = kind == EK_SIM_R0;
        int kq[NEP];
        bool dip[NEP];
        float2 a[NEP], xr[NEP][3], yr[NEP][3], o2[NEP][3];
#pragma unroll
        for (int q = 0; q < NEP; ++q) {
            const int f = threadIdx.x + q * kBlock;
            kq[q] = f;   // cell k0 + f: row t = f / nl, l = f % nl
            dip[q] = false;
            a[q] = make_float2(0.f, 0.f);
#pragma unroll
            for (int c3 = 0; c3 < 3; ++c3) xr[q][c3] = yr[q][c3] = o2[q][c3] = make_float2(0.f, 0.f);
            if (f < nep) {
                const int m = occ[k0 + f];
                dip[q] = dipf[k0 + f] != 0;
                a[q] = al[m];
                const bool ldx = need_x && (dip[q] || x_everywhere);
                const bool ldd = dip[q];
#pragma unroll
                for (int c3 = 0; c3 < 3; ++c3) {
                    const long long ck = rb + c3 * d.n + k0 + f;
                    if (ldx) xr[q][c3] = ea.x[ck];
                    if (need_y && ldd) yr[q][c3] = ea.y[ck];
                    if (need_o2 && ldd) o2[q][c3] = ea.out2[ck];
                }
            }
        }
        double acc[kSlots] = {0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int q = 0; q < NEP; ++q) {
            const int f = kq[q];
            if (f >= nep) continue;
            const int t = f >> LNLH, l = f & (NLH - 1);
            const long long k = k0 + f;
            const bool dp = dip[q];
            float2 av = a[q];
            if (ea.conj) av.y = -av.y;
            float2 cv[3], o[3];
#pragma unroll
            for (int comp = 0; comp < 3; ++comp) cv[comp] = cconj(s[(comp * tj + t) * LS + l]);
            double nrm = 0;
            switch (kind) {
                case EK_CONV:
#pragma unroll
                    for (int c3 = 0; c3 < 3; ++c3) ea.out[rb + c3 * d.n + k] = cv[c3];
                    break;
                case EK_MV:
                case EK_BICG_H:
                case EK_BICG_W:
                case EK_CG_W:
                case EK_AA_S0:
                case EK_AA_W:
#pragma unroll
                    for (int c3 = 0; c3 < 3; ++c3) {
                        const long long ck = rb + c3 * d.n + k;
                        const float2 x = xr[q][c3];
                        o[c3] = dp ? csub(x, cmul(av, cv[c3])) : make_float2(0.f, 0.f);
                        ea.out[ck] = o[c3];
                        if (!dp) continue;
                        if (kind == EK_BICG_H) dot_acc(yr[q][c3], o[c3], acc[0], acc[1]);
                        if (kind == EK_BICG_W) {
                            dot_acc(o[c3], x, acc[0], acc[1]);
                            dot_acc(o[c3], o[c3], acc[2], acc[3]);
                        }
                        if (kind == EK_CG_W) {
                            dot_acc(x, o[c3], acc[0], acc[1]);
                            dot_acc(yr[q][c3], x, acc[2], acc[3]);
                        }
                        if (kind == EK_AA_S0) nrm += norm_term(x);
                        if (kind == EK_AA_W) acc[4] += norm_term(x);
                    }
                    if (kind == EK_AA_S0 && dp) acc[5] += sqrt(nrm);
                    break;
                case EK_BICG_R0:
                case EK_CG_R0:
                case EK_AA_R0:
                case EK_SIM_R0:
#pragma unroll
                    for (int c3 = 0; c3 < 3; ++c3) {
                        const long long ck = rb + c3 * d.n + k;
                        o[c3] = dp ? cneg(cmul(av, cv[c3])) : make_float2(0.f, 0.f);
                        ea.out[ck] = o[c3];
                        if (kind == EK_BICG_R0) {
                            ea.out2[ck] = o[c3];
                            ea.out3[ck] = o[c3];
                            if (dp) dot_acc(o[c3], o[c3], acc[0], acc[1]);
                        }
                        if (kind == EK_CG_R0) ea.out2[ck] = o[c3];
                        if (kind == EK_AA_R0 && dp) acc[4] += norm_term(o[c3]);
                        if (kind == EK_SIM_R0) {
                            const float2 x = xr[q][c3];
                            ea.out2[ck] = x;                       // S = P0 (Copying)
                            ea.out3[ck] = csub(x, o[c3]);          // P = S - R (all cells)
                        }
                        if (dp) nrm += norm_term(o[c3]);
                    }
                    if ((kind == EK_CG_R0 || kind == EK_SIM_R0) && dp) acc[5] += sqrt(nrm);
                    break;
                case EK_AA_SUPD:
#pragma unroll
                    for (int c3 = 0; c3 < 3; ++c3) {
                        const long long ck = rb + c3 * d.n + k;
                        if (dp) {
                            const float2 w = csub(xr[q][c3], cmul(av, cv[c3]));
                            ea.out2[ck] = cadd(w, rmul(rB, o2[q][c3]));
                        }
                    }
                    break;
                case EK_SIM:
#pragma unroll
                    for (int c3 = 0; c3 < 3; ++c3) {
                        const long long ck = rb + c3 * d.n + k;
                        if (dp) {
                            const float2 x = xr[q][c3];
                            const float2 w = csub(x, cmul(av, cv[c3]));
                            const float2 rr = csub(w, yr[q][c3]);
                            ea.out[ck] = rr;
                            ea.out2[ck] = csub(x, rmul(ea.tau, rr));
                            nrm += norm_term(rr);
                        } else {
                            ea.out[ck] = make_float2(0.f, 0.f);
                            ea.out2[ck] = make_float2(0.f, 0.f);
                        }
                    }
                    if (dp) acc[5] += sqrt(nrm);
                    break;
                default:
                    break;
            }
        }
        if (ea.partials) {
            block_sum<kSlots>(acc, red);
            if (threadIdx.x == 0) {
                double* p = ea.partials + ((long long)rhs * gridDim.x + blockIdx.x) * kSlots;
#pragma unroll
                for (int q = 0; q < kSlots; ++q) p[q] = acc[q];
            }
        }
    };
    if (d.nj >= TJ)
        tile(std::integral_constant<int, TJ>{});
    else
        tile(d.nj);
}

}  // namespace ddafft
