"""The 128-, 256- and 512-point kernels, the 256^3 large-target workload and
the batched Monte-Carlo agglomerate workload on the GPU, against the oracle.

Long axes: every pass instance with 128/256/512-point input lines and
256/512/1024-point inverse lines (X pass with one and two right-hand sides per
tile, Z-inverse, Y-inverse, the legacy radix-2 A-hat transform up to 1024
points) is compared bit for bit with the oracle: A-hat, both convolutions, the
matvec and a 3-step BiCGStab of two right-hand sides (dipoles bit-identical,
same discrepancy log); a 512x64x64 lattice solves bit for bit as well.

Large target (sphere in 256^3, 512^3 padded FFT, x = 30, m = 1.33+1e-4i, LDR,
BiCGStab, 19 directions): A-hat bit for bit, two BiCGStab steps of both
polarizations on the GPU; polarization 0 replayed by the oracle on the same
A-hat: same iteration count and discrepancy log, bit-identical dipoles.  The
observables are held to FP64 sums over those dipoles (the oracle's sequential
float sums over 8.8 M dipoles are themselves off by up to ~1e-3).

Batched Monte-Carlo (one random agglomerate of radius-3 monomers at packing
0.3 on each of 32^3, 32x32x64 and 64^3; x ~ U(6, 12); m = 1.55+0.02i; 8
Monte-Carlo orientations x 2 polarizations = 16 right-hand sides per
instance, through ddafft_solve_batch with the reference's incident phases):
every right-hand side against the oracle's single solve, dipoles bit for bit.
"""

import time

import numpy as np
import pytest

import ddafft
import gradient as G
from conftest import bits, directions

pytestmark = pytest.mark.gpu

LONG_DIMS = [(256, 8, 8), (8, 256, 8), (8, 8, 256), (128, 16, 8), (128, 8, 4), (4, 4, 256), (512, 8, 8),
             (8, 512, 8), (8, 8, 512), (512, 4, 4)]


def rnd(rng, shape):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)


def two_material(orc32, rng, dims, kd):
    occ = (rng.random(dims) < 0.7).astype(np.int8)
    occ = occ * (1 + (rng.random(dims) < 0.3)).astype(np.int8)
    mv = np.array([0, 1, 1], np.int32)
    alpha = np.array([0, orc32.polarizability(1, kd, 1.5 + 0.01j), orc32.polarizability(1, kd, 1.7 + 0.1j)],
                     np.complex64)
    return occ, mv, alpha


@pytest.mark.parametrize("dims", LONG_DIMS)
def test_long_axis_kernels_bitexact(orc32, dims):
    rng = np.random.default_rng(sum(dims))
    kd = np.float32(0.25)
    occ, mv, alpha = two_material(orc32, rng, dims, kd)
    x = rnd(rng, dims + (3,))
    print("long axis", dims, flush=True)
    with ddafft.Context(dims, kd, nrhs_max=2) as ctx:
        A = ctx.ahat()
        assert np.array_equal(bits(A), bits(orc32.build_ahat(dims, kd))), "A-hat"
        ctx.set_target(np.ones(dims, np.int8), [0, 1])
        for conj in (False, True):
            y = ctx.convolution(x, conj=conj)
            assert np.array_equal(bits(y), bits(orc32.convolution(A, x, conj=conj))), ("convolution", conj)
        ctx.set_target(occ, mv)
        for conj in (False, True):
            y = ctx.matvec(alpha, x, conj=conj)
            assert np.array_equal(bits(y), bits(orc32.matvec(A, occ, mv, alpha, x, conj=conj))), ("matvec", conj)
        kinc = np.array([0, 0, 1], np.float32)
        einc = np.eye(3, dtype=np.complex64)[:2]
        dirs = directions(np.arange(0, 181, 30))
        res = ctx.solve("bicgstab", alpha, kinc, einc, dirs=dirs, max_step=3, err_max=1e-12)
        P = [ctx.dipoles(r) for r in range(2)]
    for pol in range(2):
        ref = orc32.dda_fft(2, occ, mv, alpha, kd, kinc, einc[pol], A, dirs, max_step=3, err_max=1e-12,
                            return_p=True)
        got = res[pol]
        print("pol", pol, "steps", got["steps"], ref["steps"], flush=True)
        assert got["steps"] == ref["steps"] == 3
        np.testing.assert_allclose(got["err_hist"], ref["err_hist"], rtol=1e-9)
        assert np.array_equal(bits(P[pol]), bits(ref["p"])), ("dipoles", pol)


def fp64_observables(occ, kd, kinc, einc, alpha, p, dirs):
    """Cext, Cabs and far field of one solution summed in float64, one
    direction at a time (memory stays O(dipoles))."""
    dims = occ.shape
    dip = occ != 0
    idx = np.nonzero(dip)
    rr = np.stack([idx[0], idx[1], idx[2]], 1).astype(np.float64) * float(kd)
    P = p[dip].astype(np.complex128)
    c1 = (np.asarray(dims) - 1) * float(kd) / 2
    c2 = np.asarray(dims) * float(kd) / 2
    E = np.asarray(einc, np.complex128).reshape(3)
    r0 = 4 * np.pi / np.real(np.vdot(E, E))
    a = np.asarray(alpha, np.complex128)[occ[dip]]
    psi = np.exp(-1j * ((rr - c2) @ np.asarray(kinc, np.float64)))
    cext = r0 * np.sum(np.imag(psi * (P @ np.conj(E))))
    cabs = r0 * np.sum(np.sum(abs(P) ** 2, 1) * (np.imag(a) / abs(a) ** 2 - 2.0 / 3.0))
    esc = np.zeros((len(dirs), 3), np.complex128)
    for q, n in enumerate(np.asarray(dirs, np.float64)):
        phi = np.exp(1j * ((rr - c1) @ n))
        s = phi @ P
        esc[q] = s - n * (n @ s)
    return cext, cabs, esc


@pytest.mark.timeout(900)
def test_large_target_256(orc32):
    dims = (256, 256, 256)
    kd = np.float32(2.0 * np.float32(30.0) / np.float32(256))
    t0 = time.time()
    occ, ndip = ddafft.target_ellipsoid(dims, 1.0, 1.0)
    dirs, normals, kinc, einc = ddafft.geometry_fixed(1, 19, 0.0, 10.0)
    dirs = dirs.reshape(-1, 3)
    m = 1.33 + 1e-4j
    alpha = np.array([[0, ddafft.polarizability(5, kd, m, kinc, einc[p])] for p in range(2)], np.complex64)
    print("large 256^3: %d dipoles, LDR alpha %s" % (ndip, alpha[:, 1]), flush=True)
    with ddafft.Context(dims, kd, nrhs_max=2) as ctx:
        ctx.set_target(occ, [0, 1])
        res = ctx.solve("bicgstab", alpha, kinc, einc, dirs=dirs, max_step=2, err_max=1e-4)
        P0 = ctx.dipoles(0)
        A = ctx.ahat()
    print("GPU done %.1f s; steps %s" % (time.time() - t0, [r["steps"] for r in res]), flush=True)
    Ao = orc32.build_ahat(dims, kd)
    same = np.array_equal(bits(A), bits(Ao))
    print("A-hat bit-identical: %s (%.1f s)" % (same, time.time() - t0), flush=True)
    assert same
    del Ao
    ref = orc32.dda_fft(2, occ, [0, 1], alpha[0], kd, kinc, einc[0], A, dirs, max_step=2, err_max=1e-4,
                        return_p=True)
    got = res[0]
    print("oracle done %.1f s: steps %d %d, hist %s %s" % (time.time() - t0, got["steps"], ref["steps"],
                                                           got["err_hist"], ref["err_hist"]), flush=True)
    assert got["steps"] == ref["steps"] == 2 and got["step_min"] == ref["step_min"]
    np.testing.assert_allclose(got["err_hist"], ref["err_hist"], rtol=1e-9)
    assert np.array_equal(bits(P0), bits(ref["p"]))
    del A
    cext, cabs, esc = fp64_observables(occ, kd, kinc, einc[0], alpha[0], P0, dirs)
    e_esc = np.max(np.abs(got["esc"] - esc)) / np.max(np.abs(esc))
    print("vs FP64 sums: cext %.2e cabs %.2e esc %.2e; oracle's float sums: cext %.2e esc %.2e" % (
        got["cext"] / cext - 1, got["cabs"] / cabs - 1, e_esc, ref["cext"] / cext - 1,
        np.max(np.abs(ref["esc"] - esc)) / np.max(np.abs(esc))), flush=True)
    assert abs(got["cext"] / cext - 1) < 1e-5 and abs(got["cabs"] / cabs - 1) < 1e-5
    assert e_esc < 1e-5


MC_LATTICES = ((32, 32, 32), (32, 32, 64), (64, 64, 64))


def agglomerate_instances(seed=20240611, norient=8):
    lib = ddafft.load()
    st = np.array([seed], np.uint64)
    dirs0, _, kinc0, einc0 = ddafft.geometry_fixed(1, 19, 0.0, 10.0)
    dirs0 = dirs0.reshape(-1, 3)
    out = []
    for dims in MC_LATTICES:
        x = 6.0 + 6.0 * lib.ddafft_rng_next(st.ctypes.data) / 2147483647.0
        kd = np.float32(2.0 * np.float32(x) / np.float32(dims[0]))
        occ, ndip, _ = ddafft.target_agglomerate(dims, 3.0, 0.3, st)
        kinc, einc, dirs = [], [], []
        for _ in range(norient):
            R = ddafft.orientation_mc(st)
            k = ddafft.rotate(R, kinc0)
            e = ddafft.rotate(R, np.real(einc0).astype(np.float32))
            d = ddafft.rotate(R, dirs0)
            for p in range(2):
                kinc.append(k)
                einc.append(e[p].astype(np.complex64))
                dirs.append(d)
        out.append(dict(dims=dims, kd=kd, occ=occ, ndip=ndip,
                        alpha=np.array([0, ddafft.polarizability(1, kd, 1.55 + 0.02j)], np.complex64),
                        kinc=np.array(kinc), einc=np.array(einc), dirs=np.array(dirs)))
    return out


@pytest.mark.timeout(600)
def test_batched_agglomerates(orc32):
    insts = agglomerate_instances()
    ctxs = []
    try:
        for it in insts:
            c = ddafft.Context(it["dims"], it["kd"], nrhs_max=16)
            c.set_target(it["occ"], [0, 1])
            ctxs.append(c)
        jobs = [dict(alpha=it["alpha"], kinc=it["kinc"], einc=it["einc"], dirs=it["dirs"], exact_phases=True)
                for it in insts]
        out = ddafft.solve_batch(ctxs, jobs, "bicgstab", max_step=300, err_max=1e-3, width=3)
        P = [[c.dipoles(r) for r in range(16)] for c in ctxs]
        A = [c.ahat() for c in ctxs]
    finally:
        for c in ctxs:
            c.close()
    for k, (it, res) in enumerate(zip(insts, out)):
        print("instance", k, it["dims"], "kd %.4f" % it["kd"], "dipoles", it["ndip"], flush=True)
        for r, got in enumerate(res):
            ref = orc32.dda_fft(2, it["occ"], [0, 1], it["alpha"], it["kd"], it["kinc"][r], it["einc"][r], A[k],
                                it["dirs"][r], max_step=300, err_max=1e-3, return_p=True)
            cext, cabs, esc = fp64_observables(it["occ"], it["kd"], it["kinc"][r], it["einc"][r], it["alpha"],
                                               P[k][r], it["dirs"][r])
            e_esc = np.max(np.abs(got["esc"] - esc)) / np.max(np.abs(esc))
            print("  rhs", r, "steps", got["steps"], ref["steps"], "esc vs FP64 %.1e" % e_esc, flush=True)
            assert got["steps"] == ref["steps"] and got["step_min"] == ref["step_min"]
            assert got["restored"] == ref["restored"]
            assert np.array_equal(bits(P[k][r]), bits(ref["p"])), (k, r)
            assert e_esc < 1e-5
            assert abs(got["cext"] / cext - 1) < 1e-5 and abs(got["cabs"] / cabs - 1) < 1e-5


@pytest.mark.timeout(600)
def test_512_axis_lattice(orc32):
    """A 512x64x64 lattice (1024x128x128 padded FFT): an elongated spheroid,
    four BiCGStab steps of both polarizations, dipoles bit-identical to the
    oracle's, the same discrepancy log."""
    dims = (512, 64, 64)
    kd = np.float32(0.1)
    occ, ndip = ddafft.target_ellipsoid(dims, 0.125, 0.125)
    alpha = np.array([0, orc32.polarizability(1, kd, 1.4 + 0.01j)], np.complex64)
    kinc = np.array([1, 0, 0], np.float32)   # along the long axis
    einc = np.array([[0, 1, 0], [0, 0, 1]], np.complex64)
    dirs = directions(np.arange(0, 181, 20))
    print("512x64x64: %d dipoles" % ndip, flush=True)
    with ddafft.Context(dims, kd, nrhs_max=2) as ctx:
        ctx.set_target(occ, [0, 1])
        res = ctx.solve("bicgstab", alpha, kinc, einc, dirs=dirs, max_step=4, err_max=1e-12)
        P = [ctx.dipoles(r) for r in range(2)]
        A = ctx.ahat()
    assert np.array_equal(bits(A), bits(orc32.build_ahat(dims, kd)))
    for pol in range(2):
        ref = orc32.dda_fft(2, occ, [0, 1], alpha, kd, kinc, einc[pol], A, dirs, max_step=4, err_max=1e-12,
                            return_p=True)
        print("pol", pol, "steps", res[pol]["steps"], ref["steps"], res[pol]["err_hist"], flush=True)
        assert res[pol]["steps"] == ref["steps"] == 4
        np.testing.assert_allclose(res[pol]["err_hist"], ref["err_hist"], rtol=1e-9)
        assert np.array_equal(bits(P[pol]), bits(ref["p"]))


def test_compact_ahat_agglomerates_converged():
    """The batched workload's agglomerates (32^3, 32x32x64, 64^3; 16
    right-hand sides each) solved to a relative residual of 1e-7 with the
    compact Â and with the reference's Â: far fields and cross sections
    within 1e-4 (measured: far fields 2e-6 to 1.6e-5, Cext and Cabs up to
    2.6e-5 at 64^3).  (At the workload's own 1e-3 the two float solves
    can part by a few percent on these ill-conditioned targets, where the
    iterate at a residual of 1e-3 is itself that far from the solution.)"""
    for it in agglomerate_instances():
        out = []
        with ddafft.Context(it["dims"], it["kd"], nrhs_max=16) as ctx:
            ctx.set_target(it["occ"], [0, 1])
            for mode in (0, 1):
                ctx.set_ahat_mode(mode)
                out.append(ctx.solve_ex("bicgstab", it["alpha"], it["kinc"], it["einc"], dirs=it["dirs"],
                                        max_step=1000, err_max=1e-7))
        e = max(np.max(np.abs(a["esc"] - b["esc"])) / np.max(np.abs(a["esc"])) for a, b in zip(*out))
        c = max(max(abs(a["cext"] / b["cext"] - 1), abs(a["cabs"] / b["cabs"] - 1)) for a, b in zip(*out))
        steps = [(a["steps"], b["steps"]) for a, b in zip(*out)]
        print(it["dims"], "steps", steps[:4], "max esc diff %.2e, cross sections %.2e" % (e, c), flush=True)
        assert e < 1e-4 and c < 1e-4
