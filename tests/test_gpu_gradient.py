"""GPU gradients (ddafft_gradient) against the CPU reference gradient and FP64
central finite differences.

The reference has no gradients; the forward model they differentiate is the
reference's (the oracle), and the gradients are pinned by finite differences
(tests/test_gradient_oracle.py on the CPU; tests/golden/grad_fd_c.npz at the
full 32x32x64 workload).  The GPU computes in FP32 with FP64 reductions, so the
comparisons carry FP32-level tolerances written in each test.
"""

import os

import numpy as np
import pytest

import ddafft
import gradient as G

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def _weights(nd, rng):
    W = np.zeros((3 + nd, 2 + nd))
    W[0, 0] = 1
    W[1, 1] = 1
    for a in range(nd):
        W[2 + a, 2 + a] = 1
    W[2 + nd] = rng.normal(size=2 + nd)
    return W


def _gpu_gradient(occ, mv, m, kd, W, dirs, normals, model=1, err_max=1e-7, max_step=400, solver="bicgstab"):
    kinc = np.array([0, 0, 1.0], np.float32)
    einc = np.array([[1, 0, 0], [0, 1, 0]], np.complex64)
    alpha = np.array([[0] + [ddafft.polarizability(model, kd, mm, kinc, einc[p]) for mm in m[1:]] for p in range(2)])
    dadm = np.array([[0] + [ddafft.polarizability_deriv(model, kd, mm, kinc, einc[p]) for mm in m[1:]]
                     for p in range(2)])
    with ddafft.Context(occ.shape, kd, nrhs_max=2 * W.shape[0]) as ctx:
        ctx.set_target(occ, mv)
        return ctx.gradient(solver, alpha, dadm, m, kinc, einc, W, dirs=dirs, normals=normals, max_step=max_step,
                            err_max=err_max)


@pytest.mark.parametrize("model", [1, 5])
def test_gradient_matches_cpu_reference(orc64, model):
    dims, kd = (8, 8, 16), 0.5
    occ, cnt = orc64.target_ellipsoid(dims, 1.0, 2.0)
    rng = np.random.default_rng(7)
    occ = occ.copy()
    occ[(occ == 1) & (rng.random(dims) < 0.2)] = 2
    mv, m = [0, 1, 1], [1, 1.5 + 0.001j, 1.7 + 0.1j]
    th = np.arange(0, 181, 30.0) * np.pi / 180
    dirs = np.stack([np.sin(th), 0 * th, np.cos(th)], 1).astype(np.float32)
    normals = np.tile([0, 1.0, 0], (len(th), 1)).astype(np.float32)
    W = _weights(len(th), rng)
    print("case dims", dims, "model", model, flush=True)
    res = _gpu_gradient(occ, mv, m, kd, W, dirs, normals, model=model)
    ref = G.gradient(orc64, 2, occ, mv, m, model, kd, np.array([0, 0, 1.0]),
                     np.array([[1, 0, 0], [0, 1, 0]], complex), orc64.build_ahat(dims, kd), dirs.astype(np.float64),
                     normals, W, max_step=500, err_max=1e-14)
    # FP32 solve to 1e-7 vs FP64 to 1e-14: objectives within 1e-5, gradients within 1e-4 (normwise per objective)
    np.testing.assert_allclose(res["f"], ref["f"], rtol=1e-5, atol=1e-5 * abs(ref["f"]).max())
    np.testing.assert_allclose(res["s11"], ref["m00"], rtol=1e-5, atol=1e-6 * abs(ref["m00"]).max())
    em = np.abs(res["grad_m"] - ref["grad_m"]).max(axis=(1, 2)) / np.abs(ref["grad_m"]).max(axis=(1, 2))
    ec = np.abs(res["grad_chi"] - ref["grad_chi"]).max(axis=(1, 2, 3, 4)) / np.abs(ref["grad_chi"]).max(
        axis=(1, 2, 3, 4))
    print("grad_m rel err", em.max(), "grad_chi rel err", ec.max(), flush=True)
    assert em.max() < 1e-4
    assert ec.max() < 1e-4


def test_gradient_single_polarization_and_empty_cells(orc64):
    """npol = 1 (no S11 weights): Cext/Cabs gradients; empty cells get 0."""
    dims, kd = (8, 8, 8), 0.6
    occ, cnt = orc64.target_ellipsoid(dims, 1.0, 1.0)
    mv, m = [0, 1], [1, 1.33 + 0.05j]
    W = np.array([[1.0, 0.0], [0.0, 1.0], [0.3, -2.0]])
    kinc = np.array([0, 0, 1.0], np.float32)
    einc = np.array([[1, 0, 0]], np.complex64)
    alpha = np.array([[0, ddafft.polarizability(1, kd, m[1])]])
    dadm = np.array([[0, ddafft.polarizability_deriv(1, kd, m[1])]])
    with ddafft.Context(dims, kd, nrhs_max=3) as ctx:
        ctx.set_target(occ, mv)
        res = ctx.gradient("bicgstab", alpha, dadm, m, kinc, einc, W, dirs=np.zeros((0, 3), np.float32),
                           max_step=300, err_max=1e-8)
    ref = G.gradient(orc64, 2, occ, mv, m, 1, kd, np.array([0, 0, 1.0]), np.array([[1, 0, 0]], complex),
                     orc64.build_ahat(dims, kd), np.zeros((0, 3)), np.zeros((0, 3)), W, max_step=500,
                     err_max=1e-14)
    assert np.all(res["grad_chi"][:, occ == 0] == 0)
    em = np.abs(res["grad_m"] - ref["grad_m"]).max() / np.abs(ref["grad_m"]).max()
    ec = np.abs(res["grad_chi"] - ref["grad_chi"]).max() / np.abs(ref["grad_chi"]).max()
    print("npol=1 grad_m rel err", em, "grad_chi rel err", ec, flush=True)
    assert em < 1e-4 and ec < 1e-4


def test_gradient_workload_against_fp64_differences():
    """32x32x64 two-material spheroid, 19 angles: GPU adjoint gradient vs FP64
    central differences (h = 1e-4) on both materials and 16 seeded dipoles."""
    path = os.path.join(HERE, "golden", "grad_fd_c.npz")
    if not os.path.exists(path):
        pytest.skip("fixture not generated (tests/golden/make_grad_fd.py)")
    fx = np.load(path)
    occ = fx["occ"]
    kd = float(fx["kd"])
    m = list(fx["m"])
    th = fx["theta"] * np.pi / 180
    dirs = np.stack([np.sin(th), 0 * th, np.cos(th)], 1).astype(np.float32)
    normals = np.tile([0, 1.0, 0], (len(th), 1)).astype(np.float32)
    nd = len(th)
    W = np.eye(2 + nd)
    print("case 32x32x64 gradient vs FP64 differences", flush=True)
    res = _gpu_gradient(occ, [0, 1, 1], m, kd, W, dirs, normals, err_max=1e-6, max_step=1000)
    fd_m, fd_d = fx["fd_m"], fx["fd_d"]
    gd = np.stack([res["grad_chi"][:, i, j, l] for i, j, l in fx["pick"]], 1)
    em = np.abs(res["grad_m"][:, 1:] - fd_m[:, 1:]).max(axis=(1, 2)) / np.abs(fd_m[:, 1:]).max(axis=(1, 2))
    ed = np.abs(gd - fd_d).max(axis=(1, 2)) / np.abs(fd_d).max(axis=(1, 2))
    print("objective errors vs FP64:", np.abs(res["f"] - fx["f0"]).max() / np.abs(fx["f0"]).max(), flush=True)
    print("material grad rel err (max over objectives)", em.max(), "dipole grad rel err", ed.max(), flush=True)
    assert em.max() < 2e-3
    assert ed.max() < 2e-3


def test_gradient_into_pinned_buffer():
    """grad_chi written into a page-locked buffer (ddafft_host_alloc) that is
    reused across calls equals the freshly allocated result, bit for bit."""
    dims, kd = (8, 8, 16), 0.5
    occ, _ = ddafft.target_ellipsoid(dims, 1.0, 2.0)
    kinc = np.array([0, 0, 1.0], np.float32)
    einc = np.array([[1, 0, 0], [0, 1, 0]], np.complex64)
    m = [1, 1.5 + 0.01j]
    alpha = np.array([[0, ddafft.polarizability(1, kd, m[1])]] * 2)
    dadm = np.array([[0, ddafft.polarizability_deriv(1, kd, m[1])]] * 2)
    W = np.array([[1.0, 0.5]], np.float32)
    out = ddafft.pinned_empty((1,) + dims + (2,), np.float32)
    print("case pinned grad_chi", dims, flush=True)
    with ddafft.Context(dims, kd, nrhs_max=2) as ctx:
        ctx.set_target(occ, [0, 1])
        ref = ctx.gradient("bicgstab", alpha, dadm, m, kinc, einc, W, ndirs=0, max_step=50, err_max=1e-5)
        for _ in range(2):
            out[:] = np.nan
            res = ctx.gradient("bicgstab", alpha, dadm, m, kinc, einc, W, ndirs=0, max_step=50, err_max=1e-5,
                               grad_chi_out=out)
            assert res["grad_chi"] is out
            assert np.array_equal(out.view(np.uint32), ref["grad_chi"].view(np.uint32))
            assert np.array_equal(res["grad_m"], ref["grad_m"])
