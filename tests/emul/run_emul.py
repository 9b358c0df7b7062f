"""Drives the CPU-emulated libddafft (tests/emul/libddafft_emul.so) through its
C ABI and checks it against the oracle.  Run under AddressSanitizer, from
tests/test_emul.py:

    LD_PRELOAD=<libasan.so> python tests/emul/run_emul.py [quick]

Exit status 0 when every check passes; ASan aborts the process on any
out-of-bounds or null access made by a kernel or by the host sequencing.
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "dda-fft_amd"), os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")]

import ddafft  # noqa: E402
import oracle  # noqa: E402
from conftest import bits, directions  # noqa: E402


def rnd(rng, shape):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)


def check(cond, what):
    print(("ok   " if cond else "FAIL ") + what, flush=True)
    return bool(cond)


def main():
    quick = "quick" in sys.argv[1:]
    ddafft.load(os.path.join(HERE, "libddafft_emul.so"))
    orc = oracle.Oracle("f32")
    ok = True
    rng = np.random.default_rng(3)
    cases = [(8, 8, 8)] if quick else [(8, 8, 8), (4, 8, 16)]
    for dims in cases:
        kd = np.float32(0.6)
        occ, _ = orc.target_ellipsoid(dims, min(1.0, dims[1] / dims[0]), min(1.0, dims[2] / dims[0]))
        occ = (occ * (1 + (rng.random(dims) < 0.3))).astype(np.int8)
        mv = np.array([0, 1, 1], np.int32)
        alpha = np.array([0, orc.polarizability(1, kd, 1.5 + 0.01j), orc.polarizability(1, kd, 1.7 + 0.1j)],
                         np.complex64)
        with ddafft.Context(dims, kd, nrhs_max=2) as ctx:
            A = ctx.ahat()
            ok &= check(np.array_equal(bits(A), bits(orc.build_ahat(dims, kd))), "ahat %s" % (dims,))
            ctx.set_target(occ, mv)
            x = rnd(rng, dims + (3,))
            for conj in (False, True):
                y = ctx.convolution(x, conj)
                ok &= check(np.array_equal(bits(y), bits(orc.convolution(A, x, conj))), "conv %s %d" % (dims, conj))
                y = ctx.matvec(alpha, x, conj)
                ok &= check(np.array_equal(bits(y), bits(orc.matvec(A, occ, mv, alpha, x, conj))),
                            "matvec %s %d" % (dims, conj))
            dirs = np.concatenate([directions(np.arange(0, 181, 30), p) for p in (0.0, 90.0)])
            kinc = np.array([0, 0, 1], np.float32)
            einc = np.eye(3, dtype=np.complex64)[:2]
            for solver in (0, 1, 2, 3):
                res = ctx.solve(solver, alpha, kinc, einc, dirs=dirs, max_step=4, err_max=1e-6, tau=0.5)
                for pol in range(2):
                    ref = orc.dda_fft(solver, occ, mv, alpha, kd, kinc, einc[pol], A, dirs, max_step=4,
                                      err_max=1e-6, tau=0.5, return_p=True)
                    got = res[pol]
                    p = ctx.dipoles(pol)
                    rel = np.max(np.abs(p - ref["p"])) / np.max(np.abs(ref["p"]))
                    erel = np.max(np.abs(got["esc"] - ref["esc"])) / np.max(np.abs(ref["esc"]))
                    good = (got["steps"] == ref["steps"] and got["step_min"] == ref["step_min"]
                            and got["restored"] == ref["restored"]
                            and np.allclose(got["err_hist"], ref["err_hist"], rtol=1e-9) and rel < 1e-5
                            and erel < 1e-5 and abs(got["cext"] / ref["cext"] - 1) < 1e-5
                            and abs(got["cabs"] / ref["cabs"] - 1) < 1e-4)
                    ok &= check(good, "solve %s solver %d pol %d steps %d/%d P %.2e esc %.2e" % (
                        dims, solver, pol, got["steps"], ref["steps"], rel, erel))
            # resident directions, one right-hand side
            ctx.set_directions(dirs)
            res1 = ctx.solve(2, alpha, kinc, einc[:1], max_step=3, err_max=1e-6)
            ok &= check(len(res1) == 1 and res1[0]["esc"].shape == (dirs.shape[0], 3), "solve resident dirs")
            p = ctx.dipoles(0)
            esc = ctx.far_field(p, dirs)
            ok &= check(np.max(np.abs(esc - res1[0]["esc"])) <= 1e-6 * np.max(np.abs(esc)), "far_field")
            ca, ce = ctx.cabs_cext(alpha, kinc, einc[0], p)
            ok &= check(abs(ce / res1[0]["cext"] - 1) < 1e-6 and abs(ca / res1[0]["cabs"] - 1) < 1e-6,
                        "cabs_cext")
            k, t = ctx.time_matvec(2, 1)
            ok &= check(t >= 0, "time_matvec")
    ok &= check_batch(orc)
    ok &= check_rhs_groups(orc)
    ok &= check_reference_sums(orc)
    ok &= check_noalloc(orc)
    ok &= check_graph(orc)
    ok &= check_fuse(orc)
    ok &= check_compact_ahat(orc)
    ok &= check_x_lines(orc)
    ok &= check_y_lines(orc)
    ok &= check_z_lines(orc)
    ok &= check_ff_mfma(orc)
    ok &= check_ff_group(orc)
    if not quick:
        ok &= check_gradient()
        ok &= check_long_axes(orc)
        ok &= check_exact_phases(orc)
    print("ALL OK" if ok else "SOME CHECKS FAILED", flush=True)
    return 0 if ok else 1




def check_exact_phases(orc):
    """A rotated incidence with exact_phases: the same dipoles as the oracle
    with the reference's (C library) phases, bit for bit."""
    ok = True
    dims, kd = (8, 8, 16), np.float32(0.5)
    occ, _ = orc.target_ellipsoid(dims, 1.0, 1.0)
    alpha = np.array([0, orc.polarizability(1, kd, 1.33 + 0.05j)], np.complex64)
    st = np.array([99], np.uint64)
    R = ddafft.orientation_mc(st)
    k = ddafft.rotate(R, np.array([0, 0, 1], np.float32))
    e = ddafft.rotate(R, np.eye(3, dtype=np.float32)[:2]).astype(np.complex64)
    A = orc.build_ahat(dims, kd)
    with ddafft.Context(dims, kd, nrhs_max=2) as ctx:
        ctx.set_target(occ, [0, 1])
        res = ctx.solve_ex("bicgstab", alpha, np.stack([k, k]), e, max_step=8, err_max=1e-9, exact_phases=True)
        for p in range(2):
            o = orc.dda_fft(2, occ, [0, 1], alpha, kd, k, e[p], A, np.zeros((0, 3), np.float32), 8, 1e-9,
                            return_p=True)
            ok &= check(res[p]["steps"] == o["steps"] and np.array_equal(bits(ctx.dipoles(p)), bits(o["p"])),
                        "exact phases, rotated incidence, pol %d" % p)
    return ok

def check_fuse(orc):
    """The fused z/y-forward kernel and the two-pass ZF + YF path, each forced
    on, against the oracle bit for bit: convolution, matvec and the four
    solvers (every prologue mode of the first pass)."""
    rng = np.random.default_rng(17)
    ok = True
    dims, kd = (4, 8, 16), np.float32(0.55)
    occ, _ = orc.target_ellipsoid(dims, 1.0, 1.0)
    occ = (occ * (1 + (rng.random(dims) < 0.3))).astype(np.int8)
    mv = np.array([0, 1, 1], np.int32)
    alpha = np.array([0, orc.polarizability(1, kd, 1.5 + 0.01j), orc.polarizability(1, kd, 1.7 + 0.1j)],
                     np.complex64)
    A = orc.build_ahat(dims, kd)
    x = rnd(rng, dims + (3,))
    kinc = np.array([0, 0, 1], np.float32)
    einc = np.eye(3, dtype=np.complex64)[:2]
    dirs = directions(np.arange(0, 181, 45), 0.0)
    refs = {}
    for fuse in (1, 0):
        with ddafft.Context(dims, kd, nrhs_max=2) as ctx:
            ctx.set_target(occ, mv)
            ctx.set_tuning(fuse_zy=fuse)
            for conj in (False, True):
                y = ctx.convolution(x, conj)
                ok &= check(np.array_equal(bits(y), bits(orc.convolution(A, x, conj))),
                            "fuse %d conv %d" % (fuse, conj))
                y = ctx.matvec(alpha, x, conj)
                ok &= check(np.array_equal(bits(y), bits(orc.matvec(A, occ, mv, alpha, x, conj))),
                            "fuse %d matvec %d" % (fuse, conj))
            for solver in (0, 1, 2, 3):
                res = ctx.solve(solver, alpha, kinc, einc, dirs=dirs, max_step=5, err_max=1e-6, tau=0.5)
                for pol in range(2):
                    if (solver, pol) not in refs:
                        refs[solver, pol] = orc.dda_fft(solver, occ, mv, alpha, kd, kinc, einc[pol], A, dirs,
                                                        max_step=5, err_max=1e-6, tau=0.5, return_p=True)
                    ref = refs[solver, pol]
                    good = (res[pol]["steps"] == ref["steps"]
                            and np.array_equal(bits(ctx.dipoles(pol)), bits(ref["p"])))
                    ok &= check(good, "fuse %d solve solver %d pol %d" % (fuse, solver, pol))
    return ok


def check_compact_ahat(orc):
    """The compact Â mode (octant of the transform of A with its Nyquist
    planes zeroed): Â, convolutions, matvec and a BiCGStab solve bit for bit
    against the oracle with that Â, then back to the reference's Â."""
    rng = np.random.default_rng(23)
    ok = True
    for dims in ((4, 8, 16), (8, 4, 4)):
        kd = np.float32(0.45)
        occ = (rng.random(dims) < 0.6).astype(np.int8)
        mv = np.array([0, 1], np.int32)
        alpha = np.array([0, orc.polarizability(1, kd, 1.5 + 0.05j)], np.complex64)
        x = rnd(rng, dims + (3,))
        Ac = orc.build_ahat_compact(dims, kd)
        with ddafft.Context(dims, kd, nrhs_max=2) as ctx:
            ctx.set_target(occ, mv)
            ctx.set_ahat_mode(1)
            ok &= check(np.array_equal(bits(ctx.ahat()), bits(Ac)), "compact ahat %s" % (dims,))
            for conj in (False, True):
                y = ctx.convolution(x, conj)
                ok &= check(np.array_equal(bits(y), bits(orc.convolution(Ac, x, conj))),
                            "compact conv %s %d" % (dims, conj))
            y = ctx.matvec(alpha, x, False)
            ok &= check(np.array_equal(bits(y), bits(orc.matvec(Ac, occ, mv, alpha, x, False))),
                        "compact matvec %s" % (dims,))
            kinc = np.array([0, 0, 1], np.float32)
            einc = np.eye(3, dtype=np.complex64)[:2]
            dirs = directions(np.arange(0, 181, 45), 0.0)
            res = ctx.solve(2, alpha, kinc, einc, dirs=dirs, max_step=5, err_max=1e-6)
            ref = orc.dda_fft(2, occ, mv, alpha, kd, kinc, einc[0], Ac, dirs, max_step=5, err_max=1e-6,
                              return_p=True)
            ok &= check(res[0]["steps"] == ref["steps"] and np.array_equal(bits(ctx.dipoles(0)), bits(ref["p"])),
                        "compact solve %s" % (dims,))
            ctx.set_ahat_mode(0)
            A = orc.build_ahat(dims, kd)
            ok &= check(np.array_equal(bits(ctx.convolution(x, False)), bits(orc.convolution(A, x, False))),
                        "reference ahat again %s" % (dims,))
    return ok


def check_x_lines(orc):
    """The X pass on its full-tile path at 32-, 64- and 128-point lines, one
    and two right-hand sides: convolutions (both conjugations) and a matvec
    bit for bit against the oracle."""
    rng = np.random.default_rng(29)
    ok = True
    for dims in ((32, 8, 4), (64, 4, 4), (128, 4, 4)):
        kd = np.float32(0.35)
        A = orc.build_ahat(dims, kd)
        occ = (rng.random(dims) < 0.7).astype(np.int8)
        mv = np.array([0, 1], np.int32)
        alpha = np.array([0, orc.polarizability(1, kd, 1.5 + 0.05j)], np.complex64)
        for nrhs in (1, 2):
            with ddafft.Context(dims, kd, nrhs_max=nrhs) as ctx:
                ctx.set_target(occ, mv)
                ctx.set_tuning(x_rb=nrhs)
                x = rnd(rng, dims + (3,))
                for conj in (False, True):
                    y = ctx.convolution(x, conj)
                    ok &= check(np.array_equal(bits(y), bits(orc.convolution(A, x, conj))),
                                "x lines conv %s R%d %d" % (dims, nrhs, conj))
                if nrhs == 2:
                    kinc = np.array([0, 0, 1], np.float32)
                    einc = np.eye(3, dtype=np.complex64)[:2]
                    dirs = directions(np.arange(0, 181, 90), 0.0)
                    res = ctx.solve(2, np.stack([alpha, alpha]), kinc, einc, dirs=dirs, max_step=3, err_max=1e-9)
                    for pol in range(2):
                        ref = orc.dda_fft(2, occ, mv, alpha, kd, kinc, einc[pol], A, dirs, max_step=3,
                                          err_max=1e-9, return_p=True)
                        ok &= check(res[pol]["steps"] == ref["steps"] and
                                    np.array_equal(bits(ctx.dipoles(pol)), bits(ref["p"])),
                                    "x lines solve %s pol %d" % (dims, pol))
    return ok


def check_y_lines(orc):
    """The Y-inverse pass on its full-tile path at 16-, 64- and 128-point
    lines: convolutions (both conjugations) and a BiCGStab solve bit for bit
    against the oracle."""
    rng = np.random.default_rng(31)
    ok = True
    for dims in ((8, 8, 8), (4, 32, 8), (4, 64, 8)):
        kd = np.float32(0.35)
        A = orc.build_ahat(dims, kd)
        occ = (rng.random(dims) < 0.7).astype(np.int8)
        mv = np.array([0, 1], np.int32)
        alpha = np.array([0, orc.polarizability(1, kd, 1.5 + 0.05j)], np.complex64)
        with ddafft.Context(dims, kd, nrhs_max=2) as ctx:
            ctx.set_target(occ, mv)
            x = rnd(rng, dims + (3,))
            for conj in (False, True):
                y = ctx.convolution(x, conj)
                ok &= check(np.array_equal(bits(y), bits(orc.convolution(A, x, conj))),
                            "y lines conv %s %d" % (dims, conj))
            kinc = np.array([0, 0, 1], np.float32)
            einc = np.eye(3, dtype=np.complex64)[:2]
            dirs = directions(np.arange(0, 181, 90), 0.0)
            res = ctx.solve(2, np.stack([alpha, alpha]), kinc, einc, dirs=dirs, max_step=3, err_max=1e-9)
            for pol in range(2):
                ref = orc.dda_fft(2, occ, mv, alpha, kd, kinc, einc[pol], A, dirs, max_step=3, err_max=1e-9,
                                  return_p=True)
                ok &= check(res[pol]["steps"] == ref["steps"] and
                            np.array_equal(bits(ctx.dipoles(pol)), bits(ref["p"])),
                            "y lines solve %s pol %d" % (dims, pol))
    return ok


def check_z_lines(orc):
    """The Z- and Y-forward passes (split, not fused) on their full-tile paths
    at 16-, 32- and 64-point lines: convolutions (both conjugations) and a
    BiCGStab solve (every prologue of its matvecs) bit for bit against the
    oracle."""
    rng = np.random.default_rng(37)
    ok = True
    for dims in ((4, 16, 16), (4, 16, 64), (4, 32, 32)):
        kd = np.float32(0.35)
        A = orc.build_ahat(dims, kd)
        occ = (rng.random(dims) < 0.7).astype(np.int8)
        mv = np.array([0, 1], np.int32)
        alpha = np.array([0, orc.polarizability(1, kd, 1.5 + 0.05j)], np.complex64)
        with ddafft.Context(dims, kd, nrhs_max=2) as ctx:
            ctx.set_target(occ, mv)
            ctx.set_tuning(fuse_zy=0)
            x = rnd(rng, dims + (3,))
            for conj in (False, True):
                y = ctx.convolution(x, conj)
                ok &= check(np.array_equal(bits(y), bits(orc.convolution(A, x, conj))),
                            "z lines conv %s %d" % (dims, conj))
            kinc = np.array([0, 0, 1], np.float32)
            einc = np.eye(3, dtype=np.complex64)[:2]
            dirs = directions(np.arange(0, 181, 90), 0.0)
            res = ctx.solve(2, np.stack([alpha, alpha]), kinc, einc, dirs=dirs, max_step=3, err_max=1e-9)
            for pol in range(2):
                ref = orc.dda_fft(2, occ, mv, alpha, kd, kinc, einc[pol], A, dirs, max_step=3, err_max=1e-9,
                                  return_p=True)
                ok &= check(res[pol]["steps"] == ref["steps"] and
                            np.array_equal(bits(ctx.dipoles(pol)), bits(ref["p"])),
                            "z lines solve %s pol %d" % (dims, pol))
    return ok


def check_ff_mfma(orc):
    """The far field on the matrix cores (ddafft_set_tuning ff_mfma) against
    the oracle's far field and against the direct kernel's."""
    rng = np.random.default_rng(41)
    ok = True
    for dims in ((8, 8, 8), (4, 8, 16)):
        kd = np.float32(0.5)
        A = orc.build_ahat(dims, kd)
        occ = (rng.random(dims) < 0.7).astype(np.int8)
        mv = np.array([0, 1], np.int32)
        alpha = np.array([0, orc.polarizability(1, kd, 1.5 + 0.05j)], np.complex64)
        kinc = np.array([0, 0, 1], np.float32)
        einc = np.eye(3, dtype=np.complex64)[:2]
        dirs = np.concatenate([directions(np.arange(0, 181, 10), p) for p in (0.0, 45.0)])   # 38: a partial tile
        out = {}
        with ddafft.Context(dims, kd, nrhs_max=2) as ctx:
            ctx.set_target(occ, mv)
            for mode in (0, 1):
                ctx.set_tuning(ff_mfma=mode)
                out[mode] = ctx.solve(2, np.stack([alpha, alpha]), kinc, einc, dirs=dirs, max_step=4, err_max=1e-9)
        for pol in range(2):
            ref = orc.dda_fft(2, occ, mv, alpha, kd, kinc, einc[pol], A, dirs, max_step=4, err_max=1e-9)
            e1 = np.max(np.abs(out[1][pol]["esc"] - ref["esc"])) / np.max(np.abs(ref["esc"]))
            e01 = np.max(np.abs(out[1][pol]["esc"] - out[0][pol]["esc"])) / np.max(np.abs(out[0][pol]["esc"]))
            ok &= check(e1 < 1e-5 and e01 < 2e-6, "ff mfma %s pol %d vs oracle %.1e vs direct %.1e" % (
                dims, pol, e1, e01))
    return ok


def check_ff_group(orc):
    """The far field by groups of directions sharing z (ddafft_set_tuning
    ff_group) against the per-direction kernel and the oracle: 2 and 3
    planes (a partial last tile of groups), one and two right-hand sides,
    columns cut into several segments (n_l = 32), and a set that does not
    split (every direction its own z), which takes the per-direction kernel."""
    rng = np.random.default_rng(43)
    ok = True
    cases = (((8, 8, 8), (0.0, 45.0), 2), ((4, 4, 32), (0.0, 60.0, 120.0), 2), ((4, 8, 16), (0.0, 60.0, 120.0), 1),
             ((8, 4, 8), None, 2))
    for dims, planes, nrhs in cases:
        kd = np.float32(0.5)
        A = orc.build_ahat(dims, kd)
        occ = (rng.random(dims) < 0.8).astype(np.int8)
        mv = np.array([0, 1], np.int32)
        alpha = np.array([0, orc.polarizability(1, kd, 1.5 + 0.05j)], np.complex64)
        kinc = np.array([0, 0, 1], np.float32)
        einc = np.eye(3, dtype=np.complex64)[:nrhs]
        if planes is None:
            v = rng.standard_normal((21, 3))
            dirs = (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32)
        else:
            dirs = np.concatenate([directions(np.arange(0, 181, 10), p) for p in planes])
        out = {}
        with ddafft.Context(dims, kd, nrhs_max=2) as ctx:
            ctx.set_target(occ, mv)
            for mode in (0, 1):
                ctx.set_tuning(ff_group=mode)
                out[mode] = ctx.solve(2, np.stack([alpha] * nrhs), kinc, einc, dirs=dirs, max_step=4, err_max=1e-9)
        for pol in range(nrhs):
            ref = orc.dda_fft(2, occ, mv, alpha, kd, kinc, einc[pol], A, dirs, max_step=4, err_max=1e-9)
            e1 = np.max(np.abs(out[1][pol]["esc"] - ref["esc"])) / np.max(np.abs(ref["esc"]))
            e01 = np.max(np.abs(out[1][pol]["esc"] - out[0][pol]["esc"])) / np.max(np.abs(out[0][pol]["esc"]))
            ok &= check(e1 < 1e-5 and e01 < 2e-6, "ff group %s planes %s rhs %d/%d vs oracle %.1e vs direct %.1e" % (
                dims, planes, pol, nrhs, e1, e01))
    return ok


def check_long_axes(orc):
    """256- and 128-point axes (the transform sizes of the 256^3 workload, 512
    points padded): A-hat, both convolutions and a short solve, bit for bit."""
    ok = True
    rng = np.random.default_rng(5)
    for dims in [(256, 4, 4), (4, 256, 4), (4, 4, 256), (128, 8, 4)]:
        kd = np.float32(0.3)
        occ, _ = orc.target_ellipsoid(dims, min(1.0, dims[1] / dims[0]), min(1.0, dims[2] / dims[0]))
        mv = np.array([0, 1], np.int32)
        alpha = np.array([0, orc.polarizability(1, kd, 1.5 + 0.01j)], np.complex64)
        with ddafft.Context(dims, kd, nrhs_max=2) as ctx:
            A = ctx.ahat()
            ok &= check(np.array_equal(bits(A), bits(orc.build_ahat(dims, kd))), "ahat %s" % (dims,))
            ctx.set_target(occ, mv)
            x = rnd(rng, dims + (3,))
            for conj in (False, True):
                ok &= check(np.array_equal(bits(ctx.convolution(x, conj)), bits(orc.convolution(A, x, conj))),
                            "conv %s %d" % (dims, conj))
            kinc = np.array([0, 0, 1], np.float32)
            einc = np.eye(3, dtype=np.complex64)[:2]
            res = ctx.solve("bicgstab", np.stack([alpha, alpha]), kinc, einc, max_step=3, err_max=1e-9)
            for p in range(2):
                o = orc.dda_fft(2, occ, mv, alpha, kd, kinc, einc[p], A, np.zeros((0, 3), np.float32), 3, 1e-9,
                                return_p=True)
                ok &= check(res[p]["steps"] == o["steps"] and np.array_equal(bits(ctx.dipoles(p)), bits(o["p"])),
                            "solve %s pol %d" % (dims, p))
    return ok

def check_rhs_groups(orc):
    """Six right-hand sides with the X pass taking all its groups in one
    block (x_blocks = 1: one Â read for all of them) or one group per block,
    in groups of two or of four (the second group half empty): dipoles
    bit-identical to the oracle's."""
    ok = True
    dims, kd = (8, 4, 8), np.float32(0.5)
    occ, _ = orc.target_ellipsoid(dims, 0.5, 1.0)
    mv = np.array([0, 1], np.int32)
    ms = [1.5 + 0.01j, 1.6 + 0.02j, 1.33 + 0.001j]
    alpha = np.array([[0, orc.polarizability(1, kd, ms[r // 2])] for r in range(6)], np.complex64)
    kinc = np.array([0, 0, 1], np.float32)
    einc = np.array([np.eye(3)[r % 2] for r in range(6)], np.complex64)
    A = orc.build_ahat(dims, kd)
    for xb, rb in ((1, 0), (4096, 0), (1 << 30, 4), (1, 4)):
        with ddafft.Context(dims, kd, nrhs_max=6) as ctx:
            ctx.set_target(occ, mv)
            ctx.set_tuning(x_blocks=xb, x_rb=rb)
            res = ctx.solve("bicgstab", alpha, kinc, einc, max_step=3, err_max=1e-9)
            for r in range(6):
                o = orc.dda_fft(2, occ, mv, alpha[r], kd, kinc, einc[r], A, np.zeros((0, 3), np.float32), 3, 1e-9,
                                return_p=True)
                ok &= check(res[r]["steps"] == o["steps"] and np.array_equal(bits(ctx.dipoles(r)), bits(o["p"])),
                            "6 right-hand sides, x_blocks %d, x_rb %d, rhs %d" % (xb, rb, r))
    return ok


def check_graph(orc):
    """Iterations replayed as captured graphs against launched ones: the same
    bits for every solver."""
    ok = True
    dims, kd = (8, 4, 8), np.float32(0.5)
    occ, _ = orc.target_ellipsoid(dims, 0.5, 1.0)
    alpha = np.array([0, orc.polarizability(1, kd, 1.5 + 0.01j)], np.complex64)
    kinc = np.array([0, 0, 1], np.float32)
    einc = np.eye(3, dtype=np.complex64)[:2]
    dirs = directions(np.arange(0, 181, 45))
    with ddafft.Context(dims, kd, nrhs_max=2) as ctx:
        ctx.set_target(occ, [0, 1])
        for solver in (0, 1, 2, 3):
            out = []
            for graph in (False, True, True):
                ctx.set_tuning(graph=graph)
                res = ctx.solve(solver, alpha, kinc, einc, dirs=dirs, max_step=6, err_max=1e-9, tau=0.5)
                out.append((res, [ctx.dipoles(r) for r in range(2)]))
            good = all(a["steps"] == b["steps"] and np.array_equal(bits(a["esc"]), bits(b["esc"]))
                       and np.array_equal(a["err_hist"], b["err_hist"]) and np.array_equal(bits(pa), bits(pb))
                       for k in (1, 2) for a, b, pa, pb in zip(out[0][0], out[k][0], out[0][1], out[k][1]))
            ok &= check(good, "graph replays, solver %d" % solver)
    return ok


def check_noalloc(orc):
    """ddafft_solve_noalloc: the results of ddafft_solve bit for bit, in both
    sum modes, twice in a row (the second smaller); two scattering planes, so
    the FP64 mode takes the grouped far field with its moments in the scratch."""
    ok = True
    for dims, nrhs in (((8, 8, 8), 2), ((4, 4, 8), 1)):
        kd = np.float32(0.6)
        occ, _ = orc.target_ellipsoid(dims, 1.0, 1.0)
        alpha = np.array([0, orc.polarizability(1, kd, 1.5 + 0.01j)], np.complex64)
        kinc = np.array([0, 0, 1], np.float32)
        einc = np.eye(3, dtype=np.complex64)[:nrhs]
        dirs = np.concatenate([directions(np.arange(0, 181, 30), p) for p in (0.0, 90.0)])
        for mode in (0, 1):
            with ddafft.Context(dims, kd, nrhs_max=2) as ctx:
                ctx.set_target(occ, [0, 1])
                ctx.set_sum_mode(mode)
                a = ctx.solve("bicgstab", alpha, kinc, einc, dirs=dirs, max_step=5, err_max=1e-6, noalloc=True)
                b = ctx.solve("bicgstab", alpha, kinc, einc, dirs=dirs, max_step=5, err_max=1e-6)
            good = all(x["steps"] == y["steps"] and np.array_equal(bits(x["esc"]), bits(y["esc"]))
                       and np.float32(x["cext"]).tobytes() == np.float32(y["cext"]).tobytes()
                       and np.array_equal(x["err_hist"], y["err_hist"]) for x, y in zip(a, b))
            ok &= check(good, "no-alloc solve %s mode %d" % (dims, mode))
    return ok


def check_reference_sums(orc):
    """Sum mode 1: Cext and Cabs bit-identical to the oracle's sequential
    float sums, the far field to its FarField with correctly rounded phases,
    for shared and for per-right-hand-side directions."""
    ok = True
    dims, kd = (8, 8, 4), np.float32(0.6)
    occ, _ = orc.target_ellipsoid(dims, 1.0, 0.5)
    mv = np.array([0, 1], np.int32)
    alpha = np.array([0, orc.polarizability(1, kd, 1.5 + 0.01j)], np.complex64)
    kinc = np.array([0, 0, 1], np.float32)
    einc = np.eye(3, dtype=np.complex64)[:2]
    dirs = np.concatenate([directions(np.arange(0, 181, 20), p) for p in (0.0, 45.0)])
    A = orc.build_ahat(dims, kd)
    with ddafft.Context(dims, kd, nrhs_max=2) as ctx:
        ctx.set_target(occ, mv)
        ctx.set_sum_mode(1)
        for per_rhs in (False, True):
            d2 = np.stack([dirs, dirs[::-1].copy()]) if per_rhs else dirs
            res = ctx.solve_ex("bicgstab", alpha, np.stack([kinc, kinc]), einc, dirs=d2, max_step=4, err_max=1e-6)
            for p in range(2):
                o = orc.dda_fft(2, occ, mv, alpha, kd, kinc, einc[p], A, np.zeros((0, 3), np.float32), 4, 1e-6,
                                return_p=True)
                dd = d2[p] if per_rhs else dirs
                orc.set_phase_rounding(True)
                esc = np.array([orc.far_field(occ, mv, kd, o["p"], d) for d in dd]).astype(np.complex64)
                orc.set_phase_rounding(False)
                good = (np.float32(res[p]["cext"]).tobytes() == np.float32(o["cext"]).tobytes()
                        and np.float32(res[p]["cabs"]).tobytes() == np.float32(o["cabs"]).tobytes()
                        and np.array_equal(bits(res[p]["esc"]), bits(esc)))
                ok &= check(good, "reference sums, per-rhs directions %d, pol %d" % (per_rhs, p))
    return ok


def check_batch(orc):
    """ddafft_solve_batch over two uneven lattices (width 1: the refill path),
    per-right-hand-side incidence and directions, against single oracle solves."""
    from test_gpu_batch import check_instance, mc_instance
    insts = [mc_instance(orc, (4, 4, 8), 0.6, 1.5 + 0.01j, 1, 3), mc_instance(orc, (8, 8, 8), 0.5, 1.33 + 0.05j, 1, 4)]
    ctxs = []
    for inst in insts:
        ctx = ddafft.Context(inst["occ"].shape, inst["kd"], nrhs_max=2)
        ctx.set_target(inst["occ"], inst["mv"])
        ctxs.append(ctx)
    jobs = [dict(alpha=i["alpha"], kinc=i["kinc"], einc=i["einc"], dirs=i["dirs"]) for i in insts]
    out = ddafft.solve_batch(ctxs, jobs, "bicgstab", max_step=6, err_max=1e-6, width=1)
    good = True
    try:
        for k, (inst, res) in enumerate(zip(insts, out)):
            check_instance(orc, inst, res, 2, 6, 1e-6, "emul batch %d" % k)
    except AssertionError as e:
        print(e)
        good = False
    for c in ctxs:
        c.close()
    return check(good, "solve_batch")


def check_gradient():
    """ddafft_gradient (forward, adjoint sources, adjoint solve, assembly) on a
    small two-material target against the FP64 CPU reference gradient."""
    import gradient as G
    o64 = oracle.Oracle("f64")
    dims, kd = (4, 4, 8), 0.5
    occ, _ = o64.target_ellipsoid(dims, 1.0, 2.0)
    occ = occ.copy()
    occ[0, 1, 3] = occ[2, 2, 4] = 2
    mv, m = [0, 1, 1], [1, 1.5 + 0.001j, 1.7 + 0.1j]
    th = np.arange(0, 181, 45.0) * np.pi / 180
    dirs = np.stack([np.sin(th), 0 * th, np.cos(th)], 1).astype(np.float32)
    normals = np.tile([0, 1.0, 0], (len(th), 1)).astype(np.float32)
    W = np.zeros((2, 2 + len(th)))
    W[0, 0] = 1.0
    W[1] = np.linspace(-1, 1, 2 + len(th))
    kinc = np.array([0, 0, 1.0], np.float32)
    einc = np.array([[1, 0, 0], [0, 1, 0]], np.complex64)
    alpha = np.array([[0] + [ddafft.polarizability(1, kd, mm) for mm in m[1:]]] * 2)
    dadm = np.array([[0] + [ddafft.polarizability_deriv(1, kd, mm) for mm in m[1:]]] * 2)
    with ddafft.Context(dims, kd, nrhs_max=4) as ctx:
        ctx.set_target(occ, mv)
        res = ctx.gradient("bicgstab", alpha, dadm, m, kinc, einc, W, dirs=dirs, normals=normals, max_step=100,
                           err_max=1e-8)
    ref = G.gradient(o64, 2, occ, mv, m, 1, kd, kinc.astype(np.float64), einc.astype(complex),
                     o64.build_ahat(dims, kd), dirs.astype(np.float64), normals, W, max_step=300, err_max=1e-14)
    em = np.abs(res["grad_m"] - ref["grad_m"]).max() / np.abs(ref["grad_m"]).max()
    ec = np.abs(res["grad_chi"] - ref["grad_chi"]).max() / np.abs(ref["grad_chi"]).max()
    return check(em < 1e-4 and ec < 1e-4, "gradient grad_m %.1e grad_chi %.1e" % (em, ec))


if __name__ == "__main__":
    sys.exit(main())
