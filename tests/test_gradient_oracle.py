"""Pins the CPU reference gradient (oracle/gradient.py) without a GPU.

The reference has no gradients, so the adjoint restatement is pinned by the
mathematics: FP64 central finite differences of the oracle's own forward solve
(the reference's algorithm in double), plus the pieces it is assembled from
checked against the oracle's restatement of the reference's functions.
"""

import numpy as np
import pytest

import gradient as G


def _case(orc, dims=(8, 8, 16), kd=0.5, seed=7):
    occ, cnt = orc.target_ellipsoid(dims, 1.0, 2.0)
    rng = np.random.default_rng(seed)
    occ = occ.copy()
    occ[(occ == 1) & (rng.random(dims) < 0.2)] = 2
    th = np.arange(0, 181, 30.0) * np.pi / 180
    dirs = np.stack([np.sin(th), 0 * th, np.cos(th)], 1)
    normals = np.tile([0, 1.0, 0], (len(th), 1))
    nd = len(th)
    W = np.zeros((3 + nd, 2 + nd))
    W[0, 0] = 1
    W[1, 1] = 1
    for a in range(nd):
        W[2 + a, 2 + a] = 1
    W[2 + nd] = rng.normal(size=2 + nd)
    return dict(occ=occ, cnt=cnt, mv=[0, 1, 1], m=[1, 1.5 + 0.001j, 1.7 + 0.1j], kd=kd, kinc=np.array([0, 0, 1.0]),
                einc=np.array([[1, 0, 0], [0, 1, 0]], complex), ahat=orc.build_ahat(dims, kd), dirs=dirs,
                normals=normals, W=W, rng=rng)


@pytest.mark.parametrize("model", [1, 2, 3, 4, 5])
def test_polarizability_derivative(orc64, model):
    kinc = (0.6, 0.0, 0.8)
    einc_re = (0.8, 0.0, -0.6)
    e = np.stack([np.asarray(einc_re), np.zeros(3)], 1)
    for m in (1.5 + 0.001j, 1.7 + 0.1j, 1.33 + 1e-4j):
        a, da = G.polarizability(model, 0.5, m, kinc, einc_re)
        # value: the oracle's restatement of dda_func.cpp:1631-1677 in double
        assert abs(a - orc64.polarizability(model, 0.5, m, kinc, e)) < 1e-13 * abs(a)
        # holomorphic derivative: central differences along both axes
        h = 1e-6
        for dz in (h, 1j * h):
            fd = (G.polarizability(model, 0.5, m + dz, kinc, einc_re)[0]
                  - G.polarizability(model, 0.5, m - dz, kinc, einc_re)[0]) / (2 * dz)
            assert abs(fd - da) < 1e-7 * abs(da)


def test_m00_against_muller_and_fd(orc64):
    rng = np.random.default_rng(3)
    th = np.arange(5, 180, 25.0) * np.pi / 180
    dirs = np.stack([np.sin(th), 0 * th, np.cos(th)], 1)
    normals = np.tile([0, 1.0, 0], (len(th), 1))
    kinc = np.array([0, 0, 1.0])
    e0, e1 = np.array([1, 0, 0], complex), np.array([0, 1, 0], complex)
    s0 = rng.normal(size=(len(th), 3)) + 1j * rng.normal(size=(len(th), 3))
    s1 = rng.normal(size=(len(th), 3)) + 1j * rng.normal(size=(len(th), 3))
    m00, g0, g1 = G.m00_and_grad(kinc, dirs, normals, e0, e1, s0, s1)
    for a in range(len(th)):
        mm = orc64.muller(kinc, dirs[a], normals[a], e0, e1, s0[a], s1[a])
        assert abs(mm[0, 0] - m00[a]) < 1e-12 * abs(m00[a])
    h = 1e-6
    for which, g in ((0, g0), (1, g1)):
        for c in range(3):
            for dz, part in ((h, lambda w: 2 * w.real), (1j * h, lambda w: -2 * w.imag)):
                sp = [s0.copy(), s1.copy()]
                sm = [s0.copy(), s1.copy()]
                sp[which][:, c] += dz
                sm[which][:, c] -= dz
                fd = (G.m00_and_grad(kinc, dirs, normals, e0, e1, *sp)[0]
                      - G.m00_and_grad(kinc, dirs, normals, e0, e1, *sm)[0]) / (2 * h)
                np.testing.assert_allclose(fd, part(g[:, c]), rtol=1e-6, atol=1e-9 * abs(m00).max())


def test_observables_restatement(orc64):
    c = _case(orc64)
    al = np.array([0, orc64.polarizability(1, c["kd"], c["m"][1]), orc64.polarizability(1, c["kd"], c["m"][2])])
    out = orc64.dda_fft(2, c["occ"], c["mv"], al, c["kd"], c["kinc"], c["einc"][0], c["ahat"], c["dirs"], 100,
                        1e-10, return_p=True)
    ce, ca, esc = G.observables(c["occ"], c["mv"], al, c["kd"], c["kinc"], c["einc"][0], out["p"], c["dirs"])
    assert abs(ce - out["cext"]) < 1e-12 * abs(ce)
    assert abs(ca - out["cabs"]) < 1e-12 * abs(ca)
    assert abs(esc - out["esc"]).max() < 1e-12 * abs(esc).max()


def _fd_objective(orc, c, model, m=None, occ=None, mv=None, alpha_override=None):
    return G.objective(orc, 2, c["occ"] if occ is None else occ, c["mv"] if mv is None else mv,
                       c["m"] if m is None else m, model, c["kd"], c["kinc"], c["einc"], c["ahat"], c["dirs"],
                       c["normals"], c["W"], 500, 1e-14, alpha_override=alpha_override)


@pytest.mark.parametrize("model", [1, 5])
def test_gradient_vs_central_differences(orc64, model):
    c = _case(orc64)
    res = G.gradient(orc64, 2, c["occ"], c["mv"], c["m"], model, c["kd"], c["kinc"], c["einc"], c["ahat"],
                     c["dirs"], c["normals"], c["W"], max_step=500, err_max=1e-14)
    h = 1e-4
    # per material (Re m, Im m)
    for q in (1, 2):
        for comp, dz in ((0, h), (1, 1j * h)):
            mp, mm = list(c["m"]), list(c["m"])
            mp[q] += dz
            mm[q] -= dz
            fd = (_fd_objective(orc64, c, model, m=mp) - _fd_objective(orc64, c, model, m=mm)) / (2 * h)
            np.testing.assert_allclose(res["grad_m"][:, q, comp], fd, rtol=2e-6, atol=2e-6 * abs(fd).max())
    if model != 1:
        return
    # per dipole chi = m^2 - 1: the dipole gets a material index of its own
    idx = np.argwhere(c["occ"] > 0)[c["rng"].choice(c["cnt"], 3, replace=False)]
    for i, j, l in idx:
        q = c["occ"][i, j, l]
        occ2 = c["occ"].copy()
        occ2[i, j, l] = 3
        chi = c["m"][q] ** 2 - 1
        for comp, dz in ((0, h), (1, 1j * h)):
            fs = []
            for sgn in (1, -1):
                mq = np.sqrt(1 + chi + sgn * dz)
                ao = np.zeros((2, 4), complex)
                ao[:, :3] = res["alpha"]
                ao[:, 3] = [orc64.polarizability(model, c["kd"], mq)] * 2
                fs.append(_fd_objective(orc64, c, model, occ=occ2, mv=c["mv"] + [1], alpha_override=ao))
            fd = (fs[0] - fs[1]) / (2 * h)
            np.testing.assert_allclose(res["grad_chi"][:, i, j, l, comp], fd, rtol=2e-6,
                                       atol=2e-6 * abs(fd).max())


def test_gradient_catches_a_dropped_term(orc64):
    """A planted mistake (the explicit Cabs term left out) must fail the FD check."""
    c = _case(orc64)
    W = np.zeros((1, 2 + len(c["dirs"])))
    W[0, 1] = 1.0
    res = G.gradient(orc64, 2, c["occ"], c["mv"], c["m"], 1, c["kd"], c["kinc"], c["einc"], c["ahat"], c["dirs"],
                     c["normals"], W, max_step=500, err_max=1e-14)
    c["W"] = W
    h = 1e-4
    mp, mm = list(c["m"]), list(c["m"])
    mp[2] += 1j * h
    mm[2] -= 1j * h
    fd = (_fd_objective(orc64, c, 1, m=mp) - _fd_objective(orc64, c, 1, m=mm)) / (2 * h)
    good = res["grad_m"][0, 2, 1]
    assert abs(good - fd[0]) < 1e-5 * abs(fd[0])
    # the explicit term enters as (w_abs/npol) r0 |P|^2/(2i alpha^2); dropping it
    # changes the Im m derivative of Cabs by a visible amount
    al = res["alpha"]
    P = [np.asarray(f["p"], np.complex128) for f in res["fw"]]
    dip = c["occ"] == 2
    a, da = G.polarizability(1, c["kd"], c["m"][2])
    expl = sum(0.5 * (4 * np.pi) * np.sum(abs(P[p][dip]) ** 2) / 2j / al[p, 2] ** 2 for p in range(2)) * da
    wrong = good - (-2 * expl.imag)
    assert abs(wrong - fd[0]) > 1e-3 * abs(fd[0])
