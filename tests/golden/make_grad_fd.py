"""Writes tests/golden/grad_fd_c.npz: FP64 central finite differences of the
adjoint-gradient workload (32x32x64 prolate spheroid 1:1:2, x = 8, two-material
random mixture m1 = 1.5+0.001i, m2 = 1.7+0.1i with P(material 2) = 0.2 from a
seeded uniform draw, CM polarizability, BiCGStab, 19 angles 0-180 step 10 in
the phi = 0 plane, both polarizations).

Objectives: Cext, Cabs (polarization means) and S11 = M00 at each of the 19
angles.  Parameters: (Re m, Im m) of both materials and (Re chi, Im chi) of 16
seeded dipoles, chi = m^2 - 1, step h = 1e-4, central differences of the
oracle's FP64 forward solve converged to 1e-16.  Run from the repository root
(about 40 minutes on 8 cores):

    python tests/golden/make_grad_fd.py
"""

import os
import sys
import time
from multiprocessing import Pool

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [os.path.join(ROOT, "oracle")]

import gradient as G  # noqa: E402
import oracle  # noqa: E402

DIMS = (32, 32, 64)
X = 8.0
KD = 2 * X / DIMS[0]
M = [1.0, 1.5 + 0.001j, 1.7 + 0.1j]
MV = [0, 1, 1]
SEED = 2024
H = 1e-4
ERR = 1e-16
MAX_STEP = 4000
THETA = np.arange(0, 181, 10.0)


def setup():
    o = oracle.Oracle("f64")
    occ, cnt = o.target_ellipsoid(DIMS, 1.0, 2.0)
    rng = np.random.default_rng(SEED)
    occ = occ.copy()
    occ[(occ == 1) & (rng.random(DIMS) < 0.2)] = 2
    cells = np.argwhere(occ > 0)
    pick = cells[rng.choice(len(cells), 16, replace=False)]
    th = THETA * np.pi / 180
    dirs = np.stack([np.sin(th), 0 * th, np.cos(th)], 1)
    normals = np.tile([0.0, 1.0, 0.0], (len(th), 1))
    nd = len(th)
    W = np.zeros((2 + nd, 2 + nd))
    for q in range(2 + nd):
        W[q, q] = 1.0
    return o, occ, pick, dirs, normals, W


def _alpha(o, m):
    return np.array([0] + [o.polarizability(1, KD, mm) for mm in m[1:]], np.complex128)


def job(args):
    kind, idx, comp, sgn = args
    o, occ, pick, dirs, normals, W = setup()
    A = o.build_ahat(DIMS, KD)
    dz = (H if comp == 0 else 1j * H) * sgn
    kinc = np.array([0, 0, 1.0])
    einc = np.array([[1, 0, 0], [0, 1, 0]], complex)
    if kind == "m":
        m = list(M)
        m[idx] = m[idx] + dz
        return args, G.objective(o, 2, occ, MV, m, 1, KD, kinc, einc, A, dirs, normals, W, MAX_STEP, ERR)
    i, j, l = pick[idx]
    q = occ[i, j, l]
    occ2 = occ.copy()
    occ2[i, j, l] = 3
    mq = np.sqrt(1 + (M[q] ** 2 - 1) + dz)
    al = np.concatenate([_alpha(o, M), [o.polarizability(1, KD, mq)]])
    ao = np.stack([al, al])
    return args, G.objective(o, 2, occ2, MV + [1], M + [mq], 1, KD, kinc, einc, A, dirs, normals, W, MAX_STEP, ERR,
                             alpha_override=ao)


def main():
    o, occ, pick, dirs, normals, W = setup()
    jobs = [("m", q, c, s) for q in (1, 2) for c in (0, 1) for s in (1, -1)]
    jobs += [("d", k, c, s) for k in range(16) for c in (0, 1) for s in (1, -1)]
    t0 = time.time()
    res = {}
    with Pool(int(os.environ.get("NPROC", "8"))) as pool:
        for args, f in pool.imap_unordered(job, jobs):
            res[args] = f
            print("%4d/%d %s %.0fs" % (len(res), len(jobs), args, time.time() - t0), flush=True)
    nobj = W.shape[0]
    fd_m = np.zeros((nobj, 3, 2))
    for q in (1, 2):
        for c in (0, 1):
            fd_m[:, q, c] = (res[("m", q, c, 1)] - res[("m", q, c, -1)]) / (2 * H)
    fd_d = np.zeros((nobj, 16, 2))
    for k in range(16):
        for c in (0, 1):
            fd_d[:, k, c] = (res[("d", k, c, 1)] - res[("d", k, c, -1)]) / (2 * H)
    A = o.build_ahat(DIMS, KD)
    kinc = np.array([0, 0, 1.0])
    einc = np.array([[1, 0, 0], [0, 1, 0]], complex)
    ref = G.gradient(o, 2, occ, MV, M, 1, KD, kinc, einc, A, dirs, normals, W, MAX_STEP, ERR)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "grad_fd_c.npz")
    np.savez_compressed(out, occ=occ, pick=pick, theta=THETA, kd=KD, m=np.array(M), h=H, fd_m=fd_m, fd_d=fd_d,
                        f0=ref["f"], adj_m=ref["grad_m"],
                        adj_d=np.stack([ref["grad_chi"][:, i, j, l] for i, j, l in pick], 1),
                        cext=ref["cext"], cabs=ref["cabs"], m00=ref["m00"])
    print("wrote", out, "adjoint vs FD (material):",
          np.abs(ref["grad_m"] - fd_m).max() / np.abs(fd_m).max(), flush=True)


if __name__ == "__main__":
    main()
