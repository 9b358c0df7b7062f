"""CPU reference of the adjoint gradients (TEST INFRASTRUCTURE ONLY).

Only tests/ use this module, as the checker of the library's gradient entry
point.  The reference has no gradients; what it fixes is the forward model
that is differentiated:

  system        K P = e,  K = diag(alpha)^-1 - A  on occupied cells, with
                e_j = E exp(i k.(r_j - c1))  (FirstApproximation,
                dda_func.cpp:654-689, and the matvec x - alpha (A x),
                dda_func.cpp:1114-1135).  A(r) is a symmetric tensor, even in
                r, so K is complex symmetric: K^T = K.
  Cext, Cabs    CabsCext, dda_func.cpp:1679-1727 (phase centre c2 = n d/2)
  far field     FarField, dda_func.cpp:1576-1625 (phase centre c1 = (n-1) d/2)
  S11 = M00     BuildMatrixJones + JonesToMuller, mmuller.cpp:175-314
  alpha(m)      Polarizability_*, dda_func.cpp:1631-1677

The objective of one gradient is the real number
  f = w_ext * mean_p Cext_p + w_abs * mean_p Cabs_p + sum_a w_a * M00(a)
over the incident polarizations p (the reference reports 0.5 (pol0 + pol1),
main.cpp:939) and the scattering directions a.  With Wirtinger derivatives
g_p = df/dP_p, one adjoint solve per polarization, K lambda_p = g_p (the
forward iteration with the source diag(alpha) g_p), gives
  df/dalpha_j = sum_p [lambda_pj . P_pj + (w_abs/npol) r0 |P_pj|^2 / (2i)] / alpha_pj^2,
and the chain rule through alpha(m) gives the per-material (Re m, Im m) and the
per-dipole susceptibility chi = m^2 - 1 gradients:
  df/dRe z = 2 Re(df/dz),  df/dIm z = -2 Im(df/dz).

Pinned by central finite differences of the FP64 oracle forward solve
(tests/test_gradient_oracle.py).
"""

import numpy as np

B1, B2, B3 = -1.8915316, 0.1648469, -1.7700004


def _cm(kd, m):
    f = kd ** 3 * (0.75 / np.pi)
    e = m * m
    return f * (e - 1) / (e + 2), f * 6 * m / (e + 2) ** 2


def _ldr_s(model, kinc, einc_re):
    if model == 4:
        return 0.2
    k = np.asarray(kinc, np.float64)
    e = np.asarray(einc_re, np.float64)
    return float(np.sum((k * e) ** 2))


def polarizability(model, kd, m, kinc=(0, 0, 1), einc_re=(1, 0, 0)):
    """alpha(m) in complex128 and its holomorphic derivative d alpha / d m."""
    m = complex(m)
    a, da = _cm(kd, m)
    if model == 1:
        return a, da
    if model in (2, 3):
        if model == 3:
            b = -np.sqrt(4 * np.pi / 3) / kd
            a, da = a / (1 + b * a), da / (1 + b * a) ** 2
        u = 1 - (2j / 3) * a
        return a / u, da / u ** 2
    if model in (4, 5):
        S = _ldr_s(model, kinc, einc_re)
        t = B1 + B2 * m * m + S * B3 * m * m - (2.0 / 3.0) * kd * 1j
        dt = (B2 + S * B3) * 2 * m
        u = 1 + (a / kd) * t
        du = (da / kd) * t + (a / kd) * dt
        return a / u, (da * u - a * du) / u ** 2
    raise ValueError("model must be 1..5")


def _cross(a, b):
    return np.cross(a, b)


def m00_and_grad(kinc, dirs, normals, einc0, einc1, esc0, esc1):
    """M00 per direction (mmuller.cpp:175-314 in float64) and its Wirtinger
    derivatives with respect to the far fields of both polarizations."""
    k = np.asarray(kinc, np.float64)
    s = np.asarray(dirs, np.float64).reshape(-1, 3)
    nv = np.asarray(normals, np.float64).reshape(-1, 3)
    E0 = np.asarray(einc0, np.complex128).reshape(3)
    E1 = np.asarray(einc1, np.complex128).reshape(3)
    S0 = np.asarray(esc0, np.complex128).reshape(-1, 3)
    S1 = np.asarray(esc1, np.complex128).reshape(-1, 3)
    v0 = _cross(np.broadcast_to(k, s.shape), nv)
    pin = v0 / np.linalg.norm(v0, axis=1)[:, None]
    v1 = _cross(s, nv)
    psc = -v1 / np.linalg.norm(v1, axis=1)[:, None]
    cc0, cc1 = (nv * E0).sum(1), (nv * E1).sum(1)
    cc2, cc3 = (pin * E0).sum(1), (pin * E1).sum(1)
    den = 1j * (cc0 * cc3 - cc1 * cc2)
    cc4, cc5 = (nv * S0).sum(1), (nv * S1).sum(1)
    cc6, cc7 = (psc * S0).sum(1), (psc * S1).sum(1)
    J00 = (cc7 * cc0 - cc6 * cc1) / den
    J01 = (cc6 * cc3 - cc7 * cc2) / den
    J10 = (cc5 * cc0 - cc4 * cc1) / den
    J11 = (cc4 * cc3 - cc5 * cc2) / den
    m00 = 0.5 * (abs(J00) ** 2 + abs(J01) ** 2 + abs(J10) ** 2 + abs(J11) ** 2)
    c00, c01, c10, c11 = np.conj(J00), np.conj(J01), np.conj(J10), np.conj(J11)
    g0 = 0.5 * (((c00 * -cc1 + c01 * cc3) / den)[:, None] * psc + ((c10 * -cc1 + c11 * cc3) / den)[:, None] * nv)
    g1 = 0.5 * (((c00 * cc0 - c01 * cc2) / den)[:, None] * psc + ((c10 * cc0 - c11 * cc2) / den)[:, None] * nv)
    return m00, g0, g1


def _positions(dims, kd):
    ni, nj, nl = dims
    i, j, l = np.meshgrid(np.arange(ni), np.arange(nj), np.arange(nl), indexing="ij")
    return np.stack([i, j, l], -1).astype(np.float64) * kd


def observables(occ, mv, alpha, kd, kinc, einc, p, dirs):
    """Cext, Cabs and the far field of one solution in float64."""
    dims = occ.shape
    r = _positions(dims, kd)
    dip = np.asarray(mv)[occ] != 0
    c1 = (np.asarray(dims) - 1) * kd / 2
    c2 = np.asarray(dims) * kd / 2
    k = np.asarray(kinc, np.float64)
    E = np.asarray(einc, np.complex128).reshape(3)
    r0 = 4 * np.pi / np.real(np.vdot(E, E))
    P = np.asarray(p, np.complex128)[dip]
    rr = r[dip]
    a = np.asarray(alpha, np.complex128)[occ[dip]]
    psi = np.exp(-1j * ((rr - c2) @ k))
    cext = r0 * np.sum(np.imag(psi * (P @ np.conj(E))))
    pp = np.sum(abs(P) ** 2, 1)
    cabs = r0 * np.sum(pp * (np.imag(a) / abs(a) ** 2 - 2.0 / 3.0))
    n = np.asarray(dirs, np.float64).reshape(-1, 3)
    phi = np.exp(1j * ((rr - c1) @ n.T))                     # [ndip, ndirs]
    tp = P[:, None, :] - n[None] * (P @ n.T)[:, :, None]
    esc = np.einsum("ja,jac->ac", phi, tp)
    return cext, cabs, esc


def adjoint_source(occ, mv, alpha, kd, kinc, einc, p, dirs, u, w_ext, w_abs, npol):
    """g = df/dP (Wirtinger) of one polarization on the occupied cells.
    u[a] = w_a (I - n n^T) dM00(a)/dEsc_p."""
    dims = occ.shape
    r = _positions(dims, kd)
    dip = np.asarray(mv)[occ] != 0
    c1 = (np.asarray(dims) - 1) * kd / 2
    c2 = np.asarray(dims) * kd / 2
    k = np.asarray(kinc, np.float64)
    E = np.asarray(einc, np.complex128).reshape(3)
    r0 = 4 * np.pi / np.real(np.vdot(E, E))
    P = np.asarray(p, np.complex128)[dip]
    rr = r[dip]
    a = np.asarray(alpha, np.complex128)[occ[dip]]
    psi = np.exp(-1j * ((rr - c2) @ k))
    g = (w_ext / npol) * r0 * psi[:, None] * np.conj(E)[None] / 2j
    g = g + (w_abs / npol) * r0 * (np.imag(a) / abs(a) ** 2 - 2.0 / 3.0)[:, None] * np.conj(P)
    if u is not None and len(u):
        n = np.asarray(dirs, np.float64).reshape(-1, 3)
        phi = np.exp(1j * ((rr - c1) @ n.T))
        g = g + phi @ np.asarray(u, np.complex128)
    out = np.zeros(dims + (3,), np.complex128)
    out[dip] = g
    return out, r0


def gradient(orc, solver, occ, mv, m, model, kd, kinc, einc, ahat, dirs, normals, weights, max_step=200,
             err_max=1e-8, tau=0.0):
    """Forward + adjoint gradients with the oracle's solver.

    m        [nmat] complex refractive index per material (m[0] unused)
    einc     [npol, 3] complex incident polarizations
    weights  [nobj, 2 + ndirs]: (w_ext, w_abs, w_s11[ndirs]) per objective
    Returns dict with f [nobj], grad_m [nobj, nmat, 2], grad_chi [nobj, ni, nj, nl, 2],
    cext/cabs [npol], esc [npol, ndirs, 3], m00 [ndirs], info (forward, adjoint).
    """
    occ = np.ascontiguousarray(occ, np.int8)
    mv = np.asarray(mv, np.int32)
    nmat = len(mv)
    einc = np.asarray(einc, np.complex128).reshape(-1, 3)
    npol = einc.shape[0]
    dirs = np.asarray(dirs, np.float64).reshape(-1, 3)
    nd = dirs.shape[0]
    weights = np.asarray(weights, np.float64).reshape(-1, 2 + nd)
    alpha = np.zeros((npol, nmat), np.complex128)
    dadm = np.zeros((npol, nmat), np.complex128)
    for p in range(npol):
        for q in range(1, nmat):
            if mv[q]:
                alpha[p, q] = orc.polarizability(model, kd, m[q], kinc, np.stack([einc[p].real, einc[p].imag], 1))
                dadm[p, q] = polarizability(model, kd, m[q], kinc, einc[p].real)[1]
    fw = []
    for p in range(npol):
        out = orc.dda_fft(solver, occ, mv, alpha[p], kd, kinc, einc[p], ahat, dirs.astype(orc.real), max_step,
                          err_max, tau, return_p=True)
        fw.append(out)
    P = [np.asarray(o["p"], np.complex128) for o in fw]
    esc = np.stack([np.asarray(o["esc"], np.complex128) for o in fw])
    cext = np.array([float(o["cext"]) for o in fw])
    cabs = np.array([float(o["cabs"]) for o in fw])
    if npol == 2:
        m00, v0, v1 = m00_and_grad(kinc, dirs, normals, einc[0], einc[1], esc[0], esc[1])
        v = [v0, v1]
    else:
        m00, v = np.zeros(nd), [np.zeros((nd, 3))]
    dip = mv[occ] != 0
    res = dict(cext=cext, cabs=cabs, esc=esc, m00=m00, f=[], grad_m=[], grad_chi=[], info=[])
    for w in weights:
        w_ext, w_abs, ws = w[0], w[1], w[2:]
        if npol != 2 and np.any(ws != 0):
            raise ValueError("S11 weights need both polarizations")
        res["f"].append(w_ext * cext.mean() + w_abs * cabs.mean() + float(np.dot(ws, m00)))
        H = np.zeros(occ.shape, np.complex128)
        Hm = np.zeros((npol,) + occ.shape, np.complex128)
        for p in range(npol):
            tv = v[p] - dirs * np.sum(dirs * v[p], 1)[:, None]
            u = ws[:, None] * tv
            g, r0 = adjoint_source(occ, mv, alpha[p], kd, kinc, einc[p], P[p], dirs, u, w_ext, w_abs, npol)
            a = alpha[p][occ]
            b = np.where(dip[..., None], a[..., None] * g, 0)
            nb = np.mean(np.linalg.norm(b[dip], axis=1))
            na = np.mean(abs(a[dip])) * np.linalg.norm(einc[p])
            s = nb / na if nb > 0 else 1.0
            lam, info, _ = orc.solve_b(solver, occ, mv, alpha[p], ahat, (b / s).astype(orc.cplx), max_step,
                                       err_max, tau)
            lam = np.asarray(lam, np.complex128) * s
            res["info"].append(info)
            lp = np.sum(lam * P[p], -1)
            pp = np.sum(abs(P[p]) ** 2, -1)
            with np.errstate(divide="ignore", invalid="ignore"):
                h = np.where(dip, (lp + (w_abs / npol) * r0 * pp / 2j) / a ** 2, 0)
            Hm[p] = h
        # per-material and per-dipole chain rule
        gm = np.zeros((nmat, 2))
        gchi = np.zeros(occ.shape + (2,))
        for q in range(1, nmat):
            if not mv[q]:
                continue
            sel = occ == q
            dz_m = sum(Hm[p][sel].sum() * dadm[p, q] for p in range(npol))
            gm[q] = [2 * dz_m.real, -2 * dz_m.imag]
            dchi = sum(Hm[p][sel] * dadm[p, q] / (2 * complex(m[q])) for p in range(npol))
            gchi[sel, 0] = 2 * dchi.real
            gchi[sel, 1] = -2 * dchi.imag
        res["grad_m"].append(gm)
        res["grad_chi"].append(gchi)
    res["f"] = np.array(res["f"])
    res["grad_m"] = np.array(res["grad_m"])
    res["grad_chi"] = np.array(res["grad_chi"])
    res["alpha"] = alpha
    res["fw"] = fw
    return res


def objective(orc, solver, occ, mv, m, model, kd, kinc, einc, ahat, dirs, normals, weights, max_step=200,
              err_max=1e-8, tau=0.0, alpha_override=None):
    """f for each weight vector from one forward solve (FD checks).
    alpha_override: optional [npol, nmat] polarizabilities used as they are."""
    occ = np.ascontiguousarray(occ, np.int8)
    mv = np.asarray(mv, np.int32)
    nmat = len(mv)
    einc = np.asarray(einc, np.complex128).reshape(-1, 3)
    npol = einc.shape[0]
    dirs = np.asarray(dirs, np.float64).reshape(-1, 3)
    nd = dirs.shape[0]
    weights = np.asarray(weights, np.float64).reshape(-1, 2 + nd)
    esc, cext, cabs = [], [], []
    for p in range(npol):
        if alpha_override is not None:
            al = np.asarray(alpha_override[p], np.complex128)
        else:
            al = np.zeros(nmat, np.complex128)
            for q in range(1, nmat):
                if mv[q]:
                    al[q] = orc.polarizability(model, kd, m[q], kinc, np.stack([einc[p].real, einc[p].imag], 1))
        out = orc.dda_fft(solver, occ, mv, al, kd, kinc, einc[p], ahat, dirs.astype(orc.real), max_step, err_max,
                          tau)
        esc.append(np.asarray(out["esc"], np.complex128))
        cext.append(float(out["cext"]))
        cabs.append(float(out["cabs"]))
    if npol == 2:
        m00 = m00_and_grad(kinc, dirs, normals, einc[0], einc[1], esc[0], esc[1])[0]
    else:
        m00 = np.zeros(nd)
    return np.array([w[0] * np.mean(cext) + w[1] * np.mean(cabs) + float(np.dot(w[2:], m00)) for w in weights])
