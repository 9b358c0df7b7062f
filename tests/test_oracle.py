"""Pins the oracle (oracle/ddaoracle.c) before it is trusted as the checker.

Two independent anchors for every function:
  * the reference's own outputs (tests/golden/*, written by the reference
    compiled from its sources): the float build must match bit for bit;
  * mathematics that fixes the result: brute-force DFTs, numpy FFTs of the
    zero-padded field, the closed-form Green tensor, a direct O(N^2) dipole
    sum, a dense linear solve, FP64 direct sums and analytic identities.
Each check is chosen so that a plausible mistake (a dropped term, a wrong
sign, index or conjugation, a transposed operand) fails it.
"""

import numpy as np
import pytest

from conftest import bits, directions, golden


def rnd(rng, shape, dtype=np.complex64):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(dtype)


def relerr(a, b):
    return np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(np.asarray(b))), 1e-300)


# ---------------------------------------------------------------------------
# 1-D transforms
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("n", [4, 8, 16, 32, 64, 128, 256, 512])
def test_fft_lines_golden_bitexact(orc32, n):
    g = golden("fft_lines.npz")
    x = g["x_%d" % n]
    assert np.array_equal(bits(orc32.fft_line(x)), bits(g["fwd_%d" % n]))
    assert np.array_equal(bits(orc32.fft_line(x, inverse=True)), bits(g["inv_%d" % n]))
    assert np.array_equal(bits(orc32.fft_line(x, legacy=True)), bits(g["legacy_%d" % n]))
    assert np.array_equal(bits(orc32.fourier_phase(n)), bits(g["phase_%d" % n]))


@pytest.mark.parametrize("n", [4, 8, 32, 128, 512])
@pytest.mark.parametrize("legacy", [False, True])
def test_fft_line_vs_bruteforce_dft(orc64, n, legacy):
    """DFT_N with forward sign exp(-2 pi i jk/N), natural order; inverse
    conj-DFT-conj with no 1/N (fft_kernel.hpp:8-9, 301-322)."""
    rng = np.random.default_rng(n)
    x = rnd(rng, (n, 3), np.complex128)
    j = np.arange(n)
    W = np.exp(-2j * np.pi * np.outer(j, j) / n)
    assert relerr(orc64.fft_line(x, legacy=legacy), W @ x) < 1e-12
    assert relerr(orc64.fft_line(x, inverse=True, legacy=legacy), W.conj() @ x) < 1e-12
    y = orc64.fft_line(orc64.fft_line(x, legacy=legacy), inverse=True, legacy=legacy)
    assert relerr(y, n * x) < 1e-12


def test_fourier_phase_closed_form(orc64):
    for n in (4, 16, 64, 256):
        w = orc64.fourier_phase(n)
        assert relerr(w, np.exp(-1j * np.pi * np.arange(n) / n)) < 1e-12


# ---------------------------------------------------------------------------
# Pruned 3-D transforms, Green tensor, convolution
# ---------------------------------------------------------------------------

def test_lattice_golden_bitexact(orc32):
    g = golden("lattice.npz")
    for tag in ("a", "b"):
        dims = tuple(int(v) for v in g["dims_" + tag])
        kd = g["kd_" + tag]
        x = g["x_" + tag]
        assert np.array_equal(bits(orc32.amatrix(dims, kd)), bits(g["amat_" + tag]))
        A = orc32.build_ahat(dims, kd)
        assert np.array_equal(bits(A), bits(g["ahat_" + tag]))
        assert np.array_equal(bits(orc32.fft3d(x)), bits(g["spec_" + tag]))
        assert np.array_equal(bits(orc32.ifft3d(g["spec_" + tag])), bits(g["ifft_" + tag]))
        assert np.array_equal(bits(orc32.convolution(A, x)), bits(g["conv_" + tag]))
        assert np.array_equal(bits(orc32.convolution(A, x, conj=True)), bits(g["convc_" + tag]))


@pytest.mark.parametrize("dims", [(4, 4, 4), (8, 4, 16), (16, 8, 4)])
def test_pruned_fft3d_is_zero_padded_dft(orc64, dims):
    rng = np.random.default_rng(1)
    x = rnd(rng, dims + (3,), np.complex128)
    pad = np.zeros(tuple(2 * d for d in dims) + (3,), np.complex128)
    pad[: dims[0], : dims[1], : dims[2]] = x
    ref = np.fft.fftn(pad, axes=(0, 1, 2))
    assert relerr(orc64.fft3d(x), ref) < 1e-12
    # pruned inverse: first octant of the unnormalized inverse transform
    spec = rnd(rng, pad.shape, np.complex128)
    inv = np.fft.ifftn(spec, axes=(0, 1, 2)) * np.prod(pad.shape[:3])
    assert relerr(orc64.ifft3d(spec), inv[: dims[0], : dims[1], : dims[2]]) < 1e-12


def green(r):
    """Closed-form A(r) = e^{ir}/r^3 [(r^2 - 1 + ir) I + (3 - r^2 - 3ir) rr/r^2]."""
    r2 = np.sum(r * r, -1)
    out = np.zeros(r.shape[:-1] + (6,), np.complex128)
    nz = r2 > 0
    rr = np.sqrt(r2[nz])
    pre = np.exp(1j * rr) / rr ** 3
    a = pre * (r2[nz] - 1 + 1j * rr)
    b = pre * (3 - r2[nz] - 3j * rr) / r2[nz]
    rv = r[nz]
    for q, (u, v) in enumerate(((0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2))):
        out[nz, q] = b * rv[:, u] * rv[:, v] + (a if u == v else 0)
    return out


def wrap_coords(dims, kd):
    ax = [np.where(np.arange(2 * n) < n, np.arange(2 * n), np.arange(2 * n) - 2 * n) * kd for n in dims]
    return np.stack(np.meshgrid(*ax, indexing="ij"), -1)


@pytest.mark.parametrize("dims,kd", [((4, 4, 4), 0.5), ((8, 4, 16), 0.7)])
def test_green_tensor_and_ahat(orc64, dims, kd):
    r = wrap_coords(dims, kd)
    A = green(r)
    assert relerr(orc64.amatrix(dims, kd), A) < 1e-12
    ref = np.fft.fftn(A, axes=(0, 1, 2)) / np.prod(r.shape[:3])
    assert relerr(orc64.build_ahat(dims, kd), ref) < 1e-12


def full_matrix(dims, kd, conj=False):
    """Dense interaction matrix A_jk = A(r_j - r_k) of the N-cell lattice."""
    idx = np.stack(np.meshgrid(*[np.arange(n) for n in dims], indexing="ij"), -1).reshape(-1, 3)
    d = (idx[:, None, :] - idx[None, :, :]) * kd
    G = green(d)
    if conj:
        G = G.conj()
    n = idx.shape[0]
    M = np.zeros((n, 3, n, 3), np.complex128)
    for q, (u, v) in enumerate(((0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2))):
        M[:, u, :, v] = G[..., q]
        M[:, v, :, u] = G[..., q]
    return M.reshape(3 * n, 3 * n)


@pytest.mark.parametrize("conj", [False, True])
def test_convolution_is_direct_dipole_sum(orc64, conj):
    dims, kd = (4, 4, 8), 0.6
    rng = np.random.default_rng(2)
    x = rnd(rng, dims + (3,), np.complex128)
    A = orc64.build_ahat(dims, kd)
    y = orc64.convolution(A, x, conj=conj)
    ref = (full_matrix(dims, kd, conj) @ x.reshape(-1)).reshape(x.shape)
    assert relerr(y, ref) < 1e-10


@pytest.mark.parametrize("dims,kd", [((4, 4, 4), 0.5), ((8, 4, 16), 0.7)])
def test_compact_ahat_is_transform_of_zeroed_tensor(orc32, orc64, dims, kd):
    """The compact mode's Â (octant of the transform of A with its Nyquist
    planes zeroed, mirrored): the FP64 DFT of that tensor, so the mirror
    and its per-component signs are right on every octant."""
    r = wrap_coords(dims, kd)
    A = green(r)
    grid = np.meshgrid(*[np.arange(2 * n) for n in dims], indexing="ij")
    nyq = (grid[0] == dims[0]) | (grid[1] == dims[1]) | (grid[2] == dims[2])
    A[nyq] = 0
    ref = np.fft.fftn(A, axes=(0, 1, 2)) / np.prod(r.shape[:3])
    assert relerr(orc64.build_ahat_compact(dims, kd), ref) < 1e-12
    assert relerr(orc32.build_ahat_compact(dims, kd), ref) < 2e-6
    # the reference's Â differs from it: its Nyquist planes are not zero
    assert relerr(orc64.build_ahat(dims, kd), ref) > 1e-3


@pytest.mark.parametrize("conj", [False, True])
def test_compact_convolution_is_direct_dipole_sum(orc32, orc64, conj):
    """The zeroed Nyquist planes never pair two lattice sites: the convolution
    with the compact Â is the direct dipole sum, and in float it agrees with
    the reference's Â to rounding."""
    dims, kd = (4, 4, 8), 0.6
    rng = np.random.default_rng(5)
    x = rnd(rng, dims + (3,), np.complex128)
    y = orc64.convolution(orc64.build_ahat_compact(dims, kd), x, conj=conj)
    ref = (full_matrix(dims, kd, conj) @ x.reshape(-1)).reshape(x.shape)
    assert relerr(y, ref) < 1e-10
    x32 = x.astype(np.complex64)
    y32 = orc32.convolution(orc32.build_ahat_compact(dims, kd), x32, conj=conj)
    r32 = orc32.convolution(orc32.build_ahat(dims, kd), x32, conj=conj)
    assert relerr(y32, r32) < 2e-6
    assert relerr(y32, ref) < 2e-6


def test_convolution_float_vs_double(orc32, orc64):
    dims, kd = (8, 8, 8), 0.4
    rng = np.random.default_rng(3)
    x = rnd(rng, dims + (3,))
    y32 = orc32.convolution(orc32.build_ahat(dims, kd), x)
    y64 = orc64.convolution(orc64.build_ahat(dims, kd), x.astype(np.complex128))
    assert relerr(y32, y64) < 2e-5


# ---------------------------------------------------------------------------
# Physics helpers
# ---------------------------------------------------------------------------

def test_polarizability_golden_and_closed_forms(orc32, orc64):
    g = golden("physics.npz")
    kinc, einc = g["pol_kinc"], g["pol_einc"]
    for a in range(5):
        for b, m in enumerate(g["pol_m"]):
            for c, kd in enumerate(g["pol_kd"]):
                v = orc32.polarizability(a + 1, kd, m, kinc=kinc, einc=einc)
                assert np.array_equal(bits(np.complex64(v)), bits(g["pol"][a, b, c]))
    m, kd = 1.5 + 0.01j, 0.3125
    cm = kd ** 3 * 3 / (4 * np.pi) * (m * m - 1) / (m * m + 2)
    rr = cm / (1 - 2j / 3 * cm)
    a1 = cm / (1 + cm * (-np.sqrt(4 * np.pi / 3) / kd))
    nr = a1 / (1 - 2j / 3 * a1)
    b1, b2, b3 = -1.8915316, 0.1648469, -1.7700004
    ildr = cm / (1 + (cm / kd) * (b1 + b2 * m * m + 0.2 * b3 * m * m - 2j / 3 * kd))
    for model, want in ((1, cm), (2, rr), (3, nr), (4, ildr)):
        got = orc64.polarizability(model, kd, m)
        assert abs(got / want - 1) < 1e-6, model
    # LDR with k = z, E = x: S = 0 (k has no x component)
    ldr0 = cm / (1 + (cm / kd) * (b1 + b2 * m * m - 2j / 3 * kd))
    assert abs(orc64.polarizability(5, kd, m) / ldr0 - 1) < 1e-6


def test_muller_golden_and_identities(orc32):
    g = golden("physics.npz")
    for t in range(g["mu_out"].shape[0]):
        mm = orc32.muller(g["mu_k"][t], g["mu_s"][t], g["mu_n"][t], g["mu_e0"][t], g["mu_e1"][t], g["mu_s0"][t],
                          g["mu_s1"][t])
        assert np.array_equal(bits(mm), bits(g["mu_out"][t]))
        # a Mueller matrix of a pure Jones matrix J satisfies M G M^T = M^T G M = |det J|^2 G with
        # G = diag(1,-1,-1,-1): the first row and the first column give the same non-negative value
        row = mm[0, 0] ** 2 - np.sum(mm[0, 1:] ** 2)
        col = mm[0, 0] ** 2 - np.sum(mm[1:, 0] ** 2)
        assert row >= -1e-4 * mm[0, 0] ** 2
        np.testing.assert_allclose(row, col, atol=1e-4 * mm[0, 0] ** 2)


def test_targets_golden(orc32):
    g = golden("physics.npz")
    for key in ("ell_32_32_32_1_1", "ell_64_64_64_1_0.5", "ell_32_32_64_1_2", "ell_16_8_32_0.5_1"):
        _, ni, nj, nl, b, c = key.split("_")
        occ, cnt = orc32.target_ellipsoid((int(ni), int(nj), int(nl)), float(b), float(c))
        assert cnt == int(g[key + "_count"])
        assert np.array_equal(np.packbits(occ.reshape(-1).astype(np.uint8)), g[key])


def test_ellipsoid_volume(orc32):
    occ, cnt = orc32.target_ellipsoid((64, 64, 64), 1.0, 0.5)
    assert abs(cnt / (np.pi / 6 * 64 * 64 * 32) - 1) < 0.01


# ---------------------------------------------------------------------------
# Observables
# ---------------------------------------------------------------------------

def test_far_field_and_cross_sections_fp64(orc32, orc64):
    dims, kd = (8, 8, 8), 0.5
    occ, _ = orc32.target_ellipsoid(dims, 1.0, 1.0)
    mv = [0, 1]
    rng = np.random.default_rng(4)
    p = rnd(rng, dims + (3,), np.complex128) * (occ[..., None] > 0)
    kinc = np.array([0.0, 0.6, 0.8])
    einc = np.array([1.0, 0, 0], np.complex128)
    alpha = np.array([0, orc64.polarizability(1, kd, 1.5 + 0.1j)])
    idx = np.stack(np.meshgrid(*[np.arange(n) for n in dims], indexing="ij"), -1).reshape(-1, 3)
    r = idx * kd
    P = p.reshape(-1, 3)
    c1 = (np.array(dims) - 1) * kd / 2
    for d in directions(np.arange(0, 181, 45), 30.0).astype(np.float64):
        ph = np.exp(1j * (r - c1) @ d)
        ref = ((P - np.outer(P @ d, d)) * ph[:, None]).sum(0)
        assert relerr(orc64.far_field(occ, mv, kd, p, d), ref) < 1e-12
        assert relerr(orc32.far_field(occ, mv, kd, p.astype(np.complex64), d.astype(np.float32)), ref) < 1e-5
    c2 = np.array(dims) * kd / 2      # the CabsCext centre (dda_func.cpp:1685)
    pre = 4 * np.pi / np.sum(np.abs(einc) ** 2)
    cext = pre * np.sum(np.imag(np.exp(-1j * (r - c2) @ kinc) * (P @ einc.conj())))
    a = alpha[1]
    n2 = np.sum(np.abs(P) ** 2, 1)
    cabs = pre * np.sum(np.imag(n2 / np.conj(a)) - 2 / 3 * n2)
    ca, ce = orc64.cabs_cext(kd, kinc, einc, alpha, occ, p)
    assert abs(ce / cext - 1) < 1e-12 and abs(ca / cabs - 1) < 1e-12


# ---------------------------------------------------------------------------
# Solvers
# ---------------------------------------------------------------------------

def dense_solution(dims, kd, occ, alpha, kinc, einc):
    n = int(np.prod(dims))
    M = full_matrix(dims, kd)
    occ_f = occ.reshape(-1)
    a = np.repeat(alpha[occ_f], 3)
    dip = np.repeat(occ_f > 0, 3)
    idx = np.stack(np.meshgrid(*[np.arange(s) for s in dims], indexing="ij"), -1).reshape(-1, 3)
    c1 = (np.array(dims) - 1) * kd / 2
    b = (alpha[occ_f][:, None] * np.exp(1j * ((idx * kd - c1) @ kinc))[:, None] * einc[None, :]).reshape(-1)
    K = np.eye(3 * n) - a[:, None] * M
    K = K[np.ix_(dip, dip)]
    p = np.zeros(3 * n, np.complex128)
    p[dip] = np.linalg.solve(K, b[dip])
    return p.reshape(dims + (3,))


@pytest.mark.parametrize("solver,steps,tau", [(0, 60, 0.0), (1, 120, 0.0), (2, 40, 0.0), (3, 200, 0.9)])
def test_fp64_solvers_converge_to_dense_solution(orc64, solver, steps, tau):
    dims, kd = (4, 4, 8), 0.5
    occ, _ = orc64.target_ellipsoid(dims, 1.0, 1.0)
    alpha = np.array([0, orc64.polarizability(1, kd, 1.4 + 0.05j)])
    kinc = np.array([0.0, 0.0, 1.0])
    einc = np.array([1.0, 0, 0], np.complex128)
    A = orc64.build_ahat(dims, kd)
    res = orc64.dda_fft(solver, occ, [0, 1], alpha, kd, kinc, einc, A, [[0, 0, 1.0]], max_step=steps,
                        err_max=1e-13, tau=tau, return_p=True)
    ref = dense_solution(dims, kd, occ, alpha, kinc, einc)
    assert relerr(res["p"], ref) < 1e-8, res["err_hist"][-3:]


def test_solves_golden_bitexact(orc32):
    """The float oracle reproduces the reference's DDA_FFT bit for bit."""
    g = golden("solves.npz")
    for tag, dims, solvers in (("s16", (16, 16, 16), (0, 1, 2, 3)), ("e8", (8, 8, 16), (0, 1, 2, 3))):
        occ, cnt = orc32.target_ellipsoid(dims, 1.0, 1.0)
        assert cnt == int(g[tag + "_occ_count"])
        kd, alpha = g[tag + "_kd"], g[tag + "_alpha"]
        A = orc32.build_ahat(dims, kd)
        dirs = directions(np.arange(0, 181, 10))
        for solver in solvers:
            for pol in (0, 1):
                key = "%s_sv%d_p%d" % (tag, solver, pol)
                tol, steps = g[key + "_cfg"]
                einc = np.zeros(3, np.complex64)
                einc[pol] = 1
                r = orc32.dda_fft(solver, occ, [0, 1], alpha, kd, [0, 0, 1], einc, A, dirs, max_step=int(steps),
                                  err_max=float(tol), tau=0.5)
                assert np.array_equal(bits(r["esc"]), bits(g[key + "_esc"])), key
                assert r["cabs"] == g[key + "_cabs"] and r["cext"] == g[key + "_cext"], key
                np.testing.assert_allclose(np.round(r["err_hist"], 7), g[key + "_hist"], atol=1e-12)


def test_residual_recurrence_matches_true_residual(orc32):
    """After k BiCGStab steps the recurrence residual R equals M P - b up to
    float drift: recompute M P - b from the returned P."""
    dims, kd = (8, 8, 8), 0.5
    occ, _ = orc32.target_ellipsoid(dims, 1.0, 1.0)
    alpha = np.array([0, orc32.polarizability(1, kd, 1.5 + 0.01j)], np.complex64)
    A = orc32.build_ahat(dims, kd)
    kinc = np.array([0, 0, 1], np.float32)
    einc = np.array([1, 0, 0], np.complex64)
    res = orc32.dda_fft(2, occ, [0, 1], alpha, kd, kinc, einc, A, [[0, 0, 1]], max_step=3, err_max=1e-9,
                        return_p=True)
    assert not res["restored"]
    b = orc32.first_approx(occ, [0, 1], alpha, kd, kinc, einc)
    R = orc32.matvec(A, occ, [0, 1], alpha, res["p"]) - b
    err = np.mean(np.linalg.norm(R.reshape(-1, 3)[occ.reshape(-1) > 0], axis=1))
    assert abs(err / res["err_hist"][-1] - 1) < 1e-3
