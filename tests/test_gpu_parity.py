"""Parity of the HIP path (libddafft_amd.so, through its C ABI) with the oracle
and with the reference's golden fixtures.

Bit-exact: Â, the pruned transforms inside the convolution, the matvec
epilogue (same float operation order as the reference, no FMA contraction)
and the dipoles of every solve (same iteration count, same printed
discrepancies).  Observables within 1e-5 relative of the oracle (the
reference sums Cext/Cabs/far field sequentially in float; the GPU sums them
in FP64).
"""

import numpy as np
import pytest

from conftest import bits, directions, golden

pytestmark = pytest.mark.gpu

ddafft = pytest.importorskip("ddafft")

DIMS = [(4, 4, 4), (8, 4, 16), (16, 16, 16), (16, 32, 8), (32, 32, 32)]


def rnd(rng, shape):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)


def target(orc32, dims, b_a=1.0, c_a=1.0):
    c_a = min(c_a, dims[2] / dims[0])
    b_a = min(b_a, dims[1] / dims[0])
    occ, cnt = orc32.target_ellipsoid(dims, b_a, c_a)
    return occ


@pytest.mark.parametrize("dims", DIMS)
def test_ahat_bitexact(orc32, dims):
    kd = np.float32(0.3125 if dims[0] >= 16 else 0.7)
    print("ahat", dims, flush=True)
    with ddafft.Context(dims, kd, nrhs_max=1) as ctx:
        A = ctx.ahat()
    R = orc32.build_ahat(dims, kd)
    assert np.array_equal(bits(A), bits(R))


def test_ahat_golden():
    g = golden("lattice.npz")
    for tag in ("a", "b"):
        dims = tuple(int(v) for v in g["dims_" + tag])
        with ddafft.Context(dims, g["kd_" + tag], nrhs_max=1) as ctx:
            A = ctx.ahat()
        assert np.array_equal(bits(A), bits(g["ahat_" + tag])), tag


@pytest.mark.parametrize("dims", DIMS)
@pytest.mark.parametrize("conj", [False, True])
def test_convolution_bitexact(orc32, dims, conj):
    rng = np.random.default_rng(hash(dims) % 2**32)
    kd = np.float32(0.5)
    x = rnd(rng, dims + (3,))
    print("convolution", dims, conj, flush=True)
    with ddafft.Context(dims, kd, nrhs_max=1) as ctx:
        ctx.set_target(np.ones(dims, np.int8), [0, 1])
        y = ctx.convolution(x, conj=conj)
        A = ctx.ahat()
    ref = orc32.convolution(A, x, conj=conj)
    assert np.array_equal(bits(y), bits(ref))


def test_convolution_golden():
    g = golden("lattice.npz")
    for tag in ("a", "b"):
        dims = tuple(int(v) for v in g["dims_" + tag])
        with ddafft.Context(dims, g["kd_" + tag], nrhs_max=1) as ctx:
            ctx.set_target(np.ones(dims, np.int8), [0, 1])
            y = ctx.convolution(g["x_" + tag])
            yc = ctx.convolution(g["x_" + tag], conj=True)
        assert np.array_equal(bits(y), bits(g["conv_" + tag])), tag
        assert np.array_equal(bits(yc), bits(g["convc_" + tag])), tag


@pytest.mark.parametrize("dims", [(8, 8, 8), (16, 16, 32), (32, 32, 32)])
@pytest.mark.parametrize("conj", [False, True])
def test_matvec_bitexact(orc32, dims, conj):
    rng = np.random.default_rng(5)
    kd = np.float32(0.4)
    occ = target(orc32, dims)
    occ = occ * (1 + (rng.random(dims) < 0.3)).astype(np.int8)    # two materials
    mv = np.array([0, 1, 1], np.int32)
    alpha = np.array([0, orc32.polarizability(1, kd, 1.5 + 0.01j), orc32.polarizability(1, kd, 1.7 + 0.1j)],
                     np.complex64)
    x = rnd(rng, dims + (3,))
    print("matvec", dims, conj, flush=True)
    with ddafft.Context(dims, kd, nrhs_max=1) as ctx:
        ctx.set_target(occ, mv)
        y = ctx.matvec(alpha, x, conj=conj)
        A = ctx.ahat()
    ref = orc32.matvec(A, occ, mv, alpha, x, conj=conj)
    assert np.array_equal(bits(y), bits(ref))


SOLVE_CASES = [
    # dims, x, m, tol, max_step
    ((16, 16, 16), 3.0, 1.5 + 0.01j, 1e-5, 8),
    ((8, 8, 16), 2.0, 1.7 + 0.1j, 1e-5, 8),
    ((32, 32, 32), 5.0, 1.5 + 0.01j, 1e-3, 100),
]


def _relerr(a, b):
    return np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(np.asarray(b))), 1e-30)


@pytest.mark.parametrize("solver", [0, 1, 2, 3])
@pytest.mark.parametrize("case", range(len(SOLVE_CASES)))
def test_solve_matches_oracle(orc32, solver, case):
    dims, x, m, tol, steps = SOLVE_CASES[case]
    kd = np.float32(2 * x / dims[0])
    occ = target(orc32, dims)
    mv = np.array([0, 1], np.int32)
    alpha = np.array([0, orc32.polarizability(1, kd, m)], np.complex64)
    kinc = np.array([0, 0, 1], np.float32)
    einc = np.eye(3, dtype=np.complex64)[:2]
    dirs = directions(np.arange(0, 181, 10))
    print("solve", solver, dims, flush=True)
    with ddafft.Context(dims, kd, nrhs_max=2) as ctx:
        ctx.set_target(occ, mv)
        res = ctx.solve(solver, alpha, kinc, einc, dirs=dirs, max_step=steps, err_max=tol, tau=0.5)
        A = ctx.ahat()
        P = [ctx.dipoles(r) for r in range(2)]
    for pol in range(2):
        ref = orc32.dda_fft(solver, occ, mv, alpha, kd, kinc, einc[pol], A, dirs, max_step=steps, err_max=tol,
                            tau=0.5, return_p=True)
        got = res[pol]
        print("pol", pol, "steps", got["steps"], ref["steps"], "P", _relerr(P[pol], ref["p"]),
              "P bit-equal frac", np.mean(bits(P[pol]) == bits(ref["p"])), "esc", _relerr(got["esc"], ref["esc"]),
              "cabs", got["cabs"] / ref["cabs"] - 1, "cext", got["cext"] / ref["cext"] - 1, flush=True)
        assert got["steps"] == ref["steps"]
        assert got["step_min"] == ref["step_min"]
        assert got["restored"] == ref["restored"]
        np.testing.assert_allclose(got["err_hist"], ref["err_hist"], rtol=1e-9)
        assert np.array_equal(bits(P[pol]), bits(ref["p"]))
        assert _relerr(got["esc"], ref["esc"]) < 1e-5
        np.testing.assert_allclose(got["cabs"], ref["cabs"], rtol=1e-4)
        np.testing.assert_allclose(got["cext"], ref["cext"], rtol=1e-5)


@pytest.mark.parametrize("tag,solvers", [("s16", (0, 1, 2, 3)), ("e8", (0, 1, 2, 3)), ("a32", (2,))])
def test_solve_golden(orc32, tag, solvers):
    """Against the reference's own DDA_FFT outputs (tests/golden/solves.npz)."""
    g = golden("solves.npz")
    dims = {"s16": (16, 16, 16), "e8": (8, 8, 16), "a32": (32, 32, 32)}[tag]
    kd = g[tag + "_kd"]
    alpha = g[tag + "_alpha"]
    occ = target(orc32, dims)
    assert int(np.count_nonzero(occ)) == int(g[tag + "_occ_count"])
    dirs = directions(np.arange(0, 181, 10))
    with ddafft.Context(dims, kd, nrhs_max=2) as ctx:
        ctx.set_target(occ, [0, 1])
        for solver in solvers:
            tol, steps = g["%s_sv%d_p0_cfg" % (tag, solver)]
            res = ctx.solve(solver, alpha, [0, 0, 1], np.eye(3, dtype=np.complex64)[:2], dirs=dirs,
                            max_step=int(steps), err_max=float(tol), tau=0.5)
            for pol in range(2):
                key = "%s_sv%d_p%d" % (tag, solver, pol)
                hist = g[key + "_hist"]
                got = res[pol]
                print(key, "steps", len(got["err_hist"]), len(hist), "esc", _relerr(got["esc"], g[key + "_esc"]),
                      "cabs", got["cabs"] / g[key + "_cabs"] - 1, "cext", got["cext"] / g[key + "_cext"] - 1,
                      flush=True)
                assert len(got["err_hist"]) == len(hist), key
                np.testing.assert_allclose(np.round(got["err_hist"], 7), hist, atol=1.5e-7, err_msg=key)
                assert _relerr(got["esc"], g[key + "_esc"]) < 2e-5, key
                np.testing.assert_allclose(got["cabs"], g[key + "_cabs"], rtol=1e-4, err_msg=key)
                np.testing.assert_allclose(got["cext"], g[key + "_cext"], rtol=2e-5, err_msg=key)


@pytest.mark.parametrize("solver", [0, 1, 2, 3])
@pytest.mark.parametrize("case", range(len(SOLVE_CASES)))
def test_reference_sums_bitexact(orc32, solver, case):
    """Sum mode 1 (the reference's sequential float sums): Cext and Cabs are
    the oracle's (the reference's) bits; the far field equals the oracle's
    FarField with correctly rounded phases bit for bit."""
    dims, x, m, tol, steps = SOLVE_CASES[case]
    kd = np.float32(2 * x / dims[0])
    occ = target(orc32, dims)
    mv = np.array([0, 1], np.int32)
    alpha = np.array([0, orc32.polarizability(1, kd, m)], np.complex64)
    kinc = np.array([0, 0, 1], np.float32)
    einc = np.eye(3, dtype=np.complex64)[:2]
    dirs = directions(np.arange(0, 181, 10))
    print("reference sums", solver, dims, flush=True)
    with ddafft.Context(dims, kd, nrhs_max=2) as ctx:
        ctx.set_target(occ, mv)
        ctx.set_sum_mode(1)
        res = ctx.solve(solver, alpha, kinc, einc, dirs=dirs, max_step=steps, err_max=tol, tau=0.5)
        A = ctx.ahat()
    for pol in range(2):
        ref = orc32.dda_fft(solver, occ, mv, alpha, kd, kinc, einc[pol], A, dirs, max_step=steps, err_max=tol,
                            tau=0.5, return_p=True)
        got = res[pol]
        assert got["steps"] == ref["steps"]
        assert bits(np.float32(got["cext"])) == bits(np.float32(ref["cext"]))
        assert bits(np.float32(got["cabs"])) == bits(np.float32(ref["cabs"]))
        orc32.set_phase_rounding(True)
        try:
            esc = np.array([orc32.far_field(occ, mv, kd, ref["p"], d) for d in dirs])
        finally:
            orc32.set_phase_rounding(False)
        assert np.array_equal(bits(got["esc"]), bits(esc.astype(np.complex64)))


@pytest.mark.parametrize("tag,solvers", [("s16", (0, 1, 2, 3)), ("e8", (0, 1, 2, 3)), ("a32", (2,))])
def test_reference_sums_golden(orc32, tag, solvers):
    """Sum mode 1 against the reference's own DDA_FFT outputs: Cext and Cabs
    bit for bit, the far field within 2e-6 of its largest value (correctly
    rounded phases against glibc's)."""
    g = golden("solves.npz")
    dims = {"s16": (16, 16, 16), "e8": (8, 8, 16), "a32": (32, 32, 32)}[tag]
    kd = g[tag + "_kd"]
    alpha = g[tag + "_alpha"]
    occ = target(orc32, dims)
    dirs = directions(np.arange(0, 181, 10))
    with ddafft.Context(dims, kd, nrhs_max=2) as ctx:
        ctx.set_target(occ, [0, 1])
        ctx.set_sum_mode(1)
        for solver in solvers:
            tol, steps = g["%s_sv%d_p0_cfg" % (tag, solver)]
            res = ctx.solve(solver, alpha, [0, 0, 1], np.eye(3, dtype=np.complex64)[:2], dirs=dirs,
                            max_step=int(steps), err_max=float(tol), tau=0.5)
            for pol in range(2):
                key = "%s_sv%d_p%d" % (tag, solver, pol)
                got = res[pol]
                e = _relerr(got["esc"], g[key + "_esc"])
                print(key, "esc", e, "cabs", got["cabs"], g[key + "_cabs"], "cext", got["cext"], g[key + "_cext"],
                      flush=True)
                assert bits(np.float32(got["cext"])) == bits(np.float32(g[key + "_cext"])), key
                assert bits(np.float32(got["cabs"])) == bits(np.float32(g[key + "_cabs"])), key
                assert e < 2e-6, key


@pytest.mark.parametrize("solver", [0, 1, 2, 3])
def test_graph_replays_match_launches(orc32, solver):
    """Every iteration replayed from its captured HIP graph gives the bits of
    the launched iterations: three solves with graphs, each after eager work
    (a launched solve, a matvec) and a synchronization, each checked against
    the launched solve, and against the oracle's dipoles."""
    dims, x, m, tol, steps = SOLVE_CASES[0]
    kd = np.float32(2 * x / dims[0])
    occ = target(orc32, dims)
    mv = np.array([0, 1], np.int32)
    alpha = np.array([0, orc32.polarizability(1, kd, m)], np.complex64)
    kinc = np.array([0, 0, 1], np.float32)
    einc = np.eye(3, dtype=np.complex64)[:2]
    dirs = directions(np.arange(0, 181, 10))
    rng = np.random.default_rng(9)
    print("graph replays, solver", solver, flush=True)
    with ddafft.Context(dims, kd, nrhs_max=2) as ctx:
        ctx.set_target(occ, mv)
        ctx.set_tuning(graph=False)
        eager = ctx.solve(solver, alpha, kinc, einc, dirs=dirs, max_step=steps, err_max=tol, tau=0.5)
        P0 = [ctx.dipoles(r) for r in range(2)]
        A = ctx.ahat()
        for rep in range(3):
            ctx.set_tuning(graph=False)
            ctx.solve(solver, alpha, kinc, einc[::-1].copy(), dirs=dirs, max_step=3, err_max=tol, tau=0.5)
            ctx.matvec(alpha, (rng.standard_normal(dims + (3,)) + 0j).astype(np.complex64))
            ctx.set_tuning(graph=True)
            got = ctx.solve(solver, alpha, kinc, einc, dirs=dirs, max_step=steps, err_max=tol, tau=0.5)
            P = [ctx.dipoles(r) for r in range(2)]
            for r in range(2):
                assert got[r]["steps"] == eager[r]["steps"] and got[r]["step_min"] == eager[r]["step_min"], rep
                assert np.array_equal(got[r]["err_hist"], eager[r]["err_hist"]), rep
                assert np.array_equal(bits(got[r]["esc"]), bits(eager[r]["esc"])), rep
                assert np.array_equal(bits(P[r]), bits(P0[r])), rep
    ref = orc32.dda_fft(solver, occ, mv, alpha, kd, kinc, einc[0], A, dirs, max_step=steps, err_max=tol, tau=0.5,
                        return_p=True)
    assert np.array_equal(bits(P0[0]), bits(ref["p"]))


ZY_PLANES = [(nj, nl) for nj in (8, 16, 32, 64) for nl in (8, 16, 32, 64)]


@pytest.mark.parametrize("plane", ZY_PLANES)
def test_zy_forward_fused_and_split_bitexact(orc32, plane):
    """The z/y-forward transforms as one kernel per i-plane (every plane shape
    it is built for) and as the two passes ZF + YF: convolution, conjugate
    convolution and matvec bit-identical to the oracle either way."""
    dims = (4,) + plane
    rng = np.random.default_rng(plane[0] * 131 + plane[1])
    kd = np.float32(0.4)
    x = rnd(rng, dims + (3,))
    occ = (rng.random(dims) < 0.7).astype(np.int8)
    mv = np.array([0, 1], np.int32)
    alpha = np.array([0, orc32.polarizability(1, kd, 1.5 + 0.02j)], np.complex64)
    print("zy forward", dims, flush=True)
    with ddafft.Context(dims, kd, nrhs_max=1) as ctx:
        ctx.set_target(occ, mv)
        A = ctx.ahat()
        for fuse in (1, 0):
            ctx.set_tuning(fuse_zy=fuse)
            for conj in (False, True):
                y = ctx.convolution(x, conj=conj)
                assert np.array_equal(bits(y), bits(orc32.convolution(A, x, conj=conj))), (fuse, conj)
            y = ctx.matvec(alpha, x, True)
            assert np.array_equal(bits(y), bits(orc32.matvec(A, occ, mv, alpha, x, True))), fuse


@pytest.mark.parametrize("fuse", [1, 0])
def test_zy_forward_solvers_bitexact(orc32, fuse):
    """Every prologue mode of the first pass (the four solvers' AXPYs folded
    into it) through the fused kernel and through ZF: dipoles bit-identical
    to the oracle's."""
    dims, kd = (8, 32, 64), np.float32(0.35)
    occ = target(orc32, dims, 1.0, 1.0)
    mv = np.array([0, 1], np.int32)
    alpha = np.array([0, orc32.polarizability(1, kd, 1.6 + 0.01j)], np.complex64)
    kinc = np.array([0, 0, 1], np.float32)
    einc = np.eye(3, dtype=np.complex64)[:2]
    dirs = directions(np.arange(0, 181, 30))
    with ddafft.Context(dims, kd, nrhs_max=2) as ctx:
        ctx.set_target(occ, mv)
        ctx.set_tuning(fuse_zy=fuse)
        A = ctx.ahat()
        for solver in (0, 1, 2, 3):
            print("zy solve", fuse, solver, flush=True)
            res = ctx.solve(solver, alpha, kinc, einc, dirs=dirs, max_step=6, err_max=1e-6, tau=0.5)
            for pol in range(2):
                ref = orc32.dda_fft(solver, occ, mv, alpha, kd, kinc, einc[pol], A, dirs, max_step=6,
                                    err_max=1e-6, tau=0.5, return_p=True)
                assert res[pol]["steps"] == ref["steps"]
                assert np.array_equal(bits(ctx.dipoles(pol)), bits(ref["p"])), (solver, pol)


@pytest.mark.parametrize("dims", DIMS + [(64, 32, 16)])
def test_compact_ahat_bitexact(orc32, dims):
    """Compact Â (ddafft_set_ahat_mode 1): the octant the GPU builds, mirrored,
    is the oracle's bit for bit; convolution, conjugate convolution and matvec
    through it are the oracle's with that Â bit for bit, and within rounding
    of the reference's (mode 0)."""
    rng = np.random.default_rng(sum(dims))
    kd = np.float32(0.45)
    x = rnd(rng, dims + (3,))
    occ = (rng.random(dims) < 0.6).astype(np.int8)
    mv = np.array([0, 1], np.int32)
    alpha = np.array([0, orc32.polarizability(1, kd, 1.5 + 0.05j)], np.complex64)
    print("compact ahat", dims, flush=True)
    Ac = orc32.build_ahat_compact(dims, kd)
    with ddafft.Context(dims, kd, nrhs_max=1) as ctx:
        ctx.set_target(occ, mv)
        y0 = ctx.convolution(x)
        ctx.set_ahat_mode(1)
        assert np.array_equal(bits(ctx.ahat()), bits(Ac))
        for conj in (False, True):
            y = ctx.convolution(x, conj=conj)
            assert np.array_equal(bits(y), bits(orc32.convolution(Ac, x, conj=conj))), conj
        y = ctx.matvec(alpha, x, False)
        assert np.array_equal(bits(y), bits(orc32.matvec(Ac, occ, mv, alpha, x, False)))
        assert _relerr(ctx.convolution(x), y0) < 2e-6
        ctx.set_ahat_mode(0)   # back to the reference's Â
        assert np.array_equal(bits(ctx.convolution(x)), bits(y0))


@pytest.mark.parametrize("solver", [0, 1, 2, 3])
def test_compact_ahat_solves(orc32, solver):
    """Solves with the compact Â: the iteration count and the dipoles
    bit-identical to the oracle's solve with that Â (eight steps, as the
    reference-Â parity cases)."""
    dims, x, m, tol, steps = SOLVE_CASES[0]
    kd = np.float32(2 * x / dims[0])
    occ = target(orc32, dims)
    mv = np.array([0, 1], np.int32)
    alpha = np.array([0, orc32.polarizability(1, kd, m)], np.complex64)
    kinc = np.array([0, 0, 1], np.float32)
    einc = np.eye(3, dtype=np.complex64)[:2]
    dirs = directions(np.arange(0, 181, 10))
    Ac = orc32.build_ahat_compact(dims, kd)
    with ddafft.Context(dims, kd, nrhs_max=2) as ctx:
        ctx.set_target(occ, mv)
        ctx.set_ahat_mode(1)
        res = ctx.solve(solver, alpha, kinc, einc, dirs=dirs, max_step=steps, err_max=tol, tau=0.5)
        P = [ctx.dipoles(r) for r in range(2)]
    for pol in range(2):
        ref = orc32.dda_fft(solver, occ, mv, alpha, kd, kinc, einc[pol], Ac, dirs, max_step=steps, err_max=tol,
                            tau=0.5, return_p=True)
        assert res[pol]["steps"] == ref["steps"]
        assert np.array_equal(bits(P[pol]), bits(ref["p"]))


def test_compact_ahat_converged_observables(orc64):
    """Converged BiCGStab solves (relative residual 1e-6) with the compact Â
    and with the reference's Â, each against an FP64 solve converged to
    1e-9: the compact mode is as accurate as the reference mode.  (The two
    float solves differ from each other by rounding, which can move the
    stopping step by one: on this case 1e-5 in Cext for one polarization.)"""
    dims, kd = (32, 32, 32), np.float32(0.3125)
    occ, _ = ddafft.target_ellipsoid(dims, 1.0, 0.5)
    mv = np.array([0, 1], np.int32)
    alpha = np.array([0, ddafft.polarizability(1, kd, 1.6 + 0.005j)], np.complex64)
    kinc = np.array([0, 0, 1], np.float32)
    einc = np.eye(3, dtype=np.complex64)[:2]
    dirs = directions(np.arange(0, 181, 5))
    out = []
    with ddafft.Context(dims, kd, nrhs_max=2) as ctx:
        ctx.set_target(occ, mv)
        for mode in (0, 1):
            ctx.set_ahat_mode(mode)
            out.append(ctx.solve("bicgstab", alpha, kinc, einc, dirs=dirs, max_step=300, err_max=1e-6))
    A64 = orc64.build_ahat(dims, float(kd))
    for pol in range(2):
        t = orc64.dda_fft(2, occ, mv, alpha.astype(np.complex128), float(kd), kinc.astype(np.float64),
                          einc[pol].astype(np.complex128), A64, dirs.astype(np.float64), max_step=300, err_max=1e-9)
        e = [dict(cext=abs(o[pol]["cext"] / t["cext"] - 1), cabs=abs(o[pol]["cabs"] / t["cabs"] - 1),
                  esc=_relerr(o[pol]["esc"], t["esc"])) for o in out]
        print("pol", pol, "steps", out[0][pol]["steps"], out[1][pol]["steps"], "vs FP64: reference Â", e[0],
              "compact Â", e[1], flush=True)
        for k in ("cext", "cabs", "esc"):
            assert e[1][k] < max(3 * e[0][k], 3e-5), k


@pytest.mark.parametrize("dims,npol,oracle_too", [((16, 16, 16), 2, True), ((32, 32, 32), 2, True),
                                                  ((32, 32, 64), 1, True), ((64, 64, 64), 2, False),
                                                  ((128, 128, 128), 2, False), ((64, 32, 128), 1, False)])
def test_far_field_grouped(orc32, dims, npol, oracle_too):
    """The far field by groups of directions sharing z (the default where the
    direction set splits so): within 2e-6 of the per-direction kernel
    (ddafft_set_tuning ff_group 0) and, at sizes the oracle finishes quickly,
    within 1e-5 of the oracle's far field.  Three planes of 91 angles leave a
    partial tile of groups; four planes of 181 are the scattering workload's
    set; 128-point axes take the kernel's larger phase-row form."""
    kd = np.float32(0.3)
    occ = target(orc32, dims, 1.0, 0.5)
    mv = np.array([0, 1], np.int32)
    alpha = np.array([0, orc32.polarizability(1, kd, 1.5 + 0.02j)], np.complex64)
    kinc = np.array([0, 0, 1], np.float32)
    einc = np.eye(3, dtype=np.complex64)[:npol]
    if oracle_too:
        dirs = np.concatenate([directions(np.arange(0, 181, 2), p) for p in (0.0, 30.0, 90.0)])   # 273
    else:
        dirs = np.concatenate([directions(np.arange(0, 181, 1), p) for p in (0.0, 45.0, 90.0, 135.0)])   # 724
    out = {}
    with ddafft.Context(dims, kd, nrhs_max=2) as ctx:
        ctx.set_target(occ, mv)
        A = ctx.ahat()
        for mode in (0, 1):
            ctx.set_tuning(ff_group=mode)
            out[mode] = ctx.solve(2, alpha, kinc, einc, dirs=dirs, max_step=6, err_max=1e-6)
    for pol in range(npol):
        e01 = _relerr(out[1][pol]["esc"], out[0][pol]["esc"])
        e = float("nan")
        if oracle_too:
            ref = orc32.dda_fft(2, occ, mv, alpha, kd, kinc, einc[pol], A, dirs, max_step=6, err_max=1e-6)
            e = _relerr(out[1][pol]["esc"], ref["esc"])
            assert e < 1e-5
        print("ff grouped", dims, pol, "vs oracle %.2e, vs per-direction %.2e" % (e, e01), flush=True)
        assert e01 < 2e-6
        assert out[1][pol]["cext"] == out[0][pol]["cext"]
        assert out[1][pol]["steps"] == out[0][pol]["steps"]


@pytest.mark.parametrize("dims,npol", [((16, 16, 16), 2), ((32, 32, 32), 2), ((32, 32, 64), 1)])
def test_far_field_mfma(orc32, dims, npol):
    """The far field summed on the matrix cores (ddafft_set_tuning ff_mfma):
    within 1e-5 of the oracle's far field and within 2e-6 of the direct
    kernel's, with a direction count that leaves a partial 16-direction tile."""
    kd = np.float32(0.3)
    occ = target(orc32, dims, 1.0, 0.5)
    mv = np.array([0, 1], np.int32)
    alpha = np.array([0, orc32.polarizability(1, kd, 1.5 + 0.02j)], np.complex64)
    kinc = np.array([0, 0, 1], np.float32)
    einc = np.eye(3, dtype=np.complex64)[:npol]
    dirs = np.concatenate([directions(np.arange(0, 181, 2), p) for p in (0.0, 30.0, 90.0)])   # 273
    out = {}
    with ddafft.Context(dims, kd, nrhs_max=2) as ctx:
        ctx.set_target(occ, mv)
        ctx.set_tuning(ff_group=0)   # against the per-direction kernel
        A = ctx.ahat()
        for mode in (0, 1):
            ctx.set_tuning(ff_mfma=mode)
            out[mode] = ctx.solve(2, alpha, kinc, einc, dirs=dirs, max_step=6, err_max=1e-6)
    for pol in range(npol):
        ref = orc32.dda_fft(2, occ, mv, alpha, kd, kinc, einc[pol], A, dirs, max_step=6, err_max=1e-6)
        e = _relerr(out[1][pol]["esc"], ref["esc"])
        e01 = _relerr(out[1][pol]["esc"], out[0][pol]["esc"])
        print("ff mfma", dims, pol, "vs oracle %.2e, vs direct %.2e" % (e, e01), flush=True)
        assert e < 1e-5 and e01 < 2e-6
        assert out[1][pol]["cext"] == out[0][pol]["cext"]
