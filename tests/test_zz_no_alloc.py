"""ddafft_solve_noalloc (the solve that allocates nothing: its scratch is a
fixed per-GPU device array) against ddafft_solve, through the C ABI: two
calls in a row, the second with a smaller input (smaller lattice, one
right-hand side, fewer directions), both bit for bit the results of the
allocating solve; the reference-order sum mode likewise; a request larger
than the scratch is refused with a clear error."""

import numpy as np
import pytest

import ddafft
from conftest import bits, directions

pytestmark = pytest.mark.gpu


def _case(dims, x, m):
    kd = np.float32(2 * x / dims[0])
    occ, _ = ddafft.target_ellipsoid(dims, 1.0, 0.5)
    alpha = np.array([0, ddafft.polarizability(1, kd, m)], np.complex64)
    return kd, occ, alpha


def _same(a, b, what):
    assert len(a) == len(b)
    for r, (x, y) in enumerate(zip(a, b)):
        assert x["steps"] == y["steps"] and x["step_min"] == y["step_min"], (what, r)
        assert x["restored"] == y["restored"] and x["matvecs"] == y["matvecs"], (what, r)
        assert np.array_equal(bits(x["esc"]), bits(y["esc"])), (what, r)
        assert bits(np.float32(x["cext"])) == bits(np.float32(y["cext"])), (what, r)
        assert bits(np.float32(x["cabs"])) == bits(np.float32(y["cabs"])), (what, r)
        assert np.array_equal(x["err_hist"], y["err_hist"]), (what, r)


@pytest.mark.parametrize("sum_mode", [0, 1])
def test_noalloc_matches_solve(sum_mode):
    kinc = np.array([0, 0, 1], np.float32)
    einc = np.eye(3, dtype=np.complex64)[:2]
    # first call: 32^3, two right-hand sides, 2 planes x 91 directions
    dims = (32, 32, 32)
    kd, occ, alpha = _case(dims, 4.0, 1.5 + 0.01j)
    dirs = np.concatenate([directions(np.arange(0, 181, 2), p) for p in (0.0, 90.0)])
    print("no-alloc call 1:", dims, "sum mode", sum_mode, flush=True)
    with ddafft.Context(dims, kd, nrhs_max=2) as ctx:
        ctx.set_target(occ, [0, 1])
        ctx.set_sum_mode(sum_mode)
        ref = ctx.solve("bicgstab", alpha, kinc, einc, dirs=dirs, max_step=50, err_max=1e-4)
        Pr = [ctx.dipoles(r) for r in range(2)]
        got = ctx.solve("bicgstab", alpha, kinc, einc, dirs=dirs, max_step=50, err_max=1e-4, noalloc=True)
        Pg = [ctx.dipoles(r) for r in range(2)]
    _same(got, ref, "call 1")
    for a, b in zip(Pg, Pr):
        assert np.array_equal(bits(a), bits(b))
    # second call, smaller: 16^3, one right-hand side, 19 directions
    dims = (16, 16, 16)
    kd, occ, alpha = _case(dims, 2.0, 1.33 + 0.05j)
    dirs = directions(np.arange(0, 181, 10))
    print("no-alloc call 2:", dims, flush=True)
    with ddafft.Context(dims, kd, nrhs_max=2) as ctx:
        ctx.set_target(occ, [0, 1])
        ctx.set_sum_mode(sum_mode)
        got = ctx.solve("cg", alpha, kinc, einc[:1], dirs=dirs, max_step=40, err_max=1e-5, noalloc=True)
        Pg = ctx.dipoles(0)
        ref = ctx.solve("cg", alpha, kinc, einc[:1], dirs=dirs, max_step=40, err_max=1e-5)
        Pr = ctx.dipoles(0)
    _same(got, ref, "call 2")
    assert np.array_equal(bits(Pg), bits(Pr))


def test_noalloc_refuses_oversized_request():
    dims = (16, 16, 16)
    kd, occ, alpha = _case(dims, 2.0, 1.5 + 0.01j)
    th = np.linspace(0, 180, 100000)
    dirs = directions(th)   # 100000 directions: phase tables and partials beyond the scratch
    with ddafft.Context(dims, kd, nrhs_max=1) as ctx:
        ctx.set_target(occ, [0, 1])
        with pytest.raises(ddafft.DdafftError, match="allocation-free solve has"):
            ctx.solve("bicgstab", alpha, [0, 0, 1], np.eye(3, dtype=np.complex64)[:1], dirs=dirs, max_step=5,
                      err_max=1e-3, noalloc=True)
        # the context still works afterwards, with and without allocation
        a = ctx.solve("bicgstab", alpha, [0, 0, 1], np.eye(3, dtype=np.complex64)[:1], dirs=dirs[:7], max_step=5,
                      err_max=1e-3, noalloc=True)
        b = ctx.solve("bicgstab", alpha, [0, 0, 1], np.eye(3, dtype=np.complex64)[:1], dirs=dirs[:7], max_step=5,
                      err_max=1e-3)
    _same(a, b, "after refusal")
