"""Host-side pieces of the library (no GPU): targets, polarizabilities,
scattering geometry and Mueller matrices, bit-exact against the reference's
golden fixtures and the oracle."""

import ctypes as C
import re

import numpy as np
import pytest

from conftest import ROOT, bits, directions, golden

ddafft = pytest.importorskip("ddafft")


@pytest.fixture(scope="module")
def lib():
    return ddafft.load()


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def test_exports_match_header(lib):
    """Every function declared in include/ddafft.h is exported by the library."""
    hdr = open(ROOT + "/include/ddafft.h").read()
    names = sorted(set(re.findall(r"\b(ddafft_[a-z_0-9]+)\s*\(", hdr)))
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n


def test_create_validates_lattice(lib):
    """Axes are powers of two in [4, 512] (the reference accepts any power of
    two, options.cpp:506-513; 512 is this build's longest line) and a
    lattice has at most 2^24 cells; both are checked before any GPU work."""
    h = C.c_void_p()
    for dims in ((1024, 4, 4), (4, 2, 4), (12, 4, 4), (512, 512, 128)):
        assert lib.ddafft_create(*dims, C.c_float(0.5), 2, C.byref(h)) == -1, dims
        assert not h.value
    assert "2^24" in lib.ddafft_last_error().decode()


@pytest.mark.parametrize("key", ["ell_32_32_32_1_1", "ell_64_64_64_1_0.5", "ell_32_32_64_1_2", "ell_16_8_32_0.5_1",
                                 "par_16_16_16_0.5_0.75"])
def test_targets_bitexact(lib, key):
    g = golden("physics.npz")
    kind, ni, nj, nl, b, c = key.split("_")
    dims = (int(ni), int(nj), int(nl))
    occ = np.zeros(dims, np.int8)
    fn = lib.ddafft_target_ellipsoid if kind == "ell" else lib.ddafft_target_parallelepiped
    fn.argtypes = [C.c_int] * 3 + [C.c_double, C.c_double, C.c_void_p]
    cnt = fn(*dims, float(b), float(c), _p(occ))
    assert cnt == int(g[key + "_count"])
    assert np.array_equal(np.packbits(occ.reshape(-1).astype(np.uint8)), g[key])


def test_bisphere_matches_oracle_restatement(lib, orc32):
    lib.ddafft_target_bisphere.argtypes = [C.c_int] * 3 + [C.c_double, C.c_void_p]
    for ratio in (0.5, 1.0, 2.0):
        occ = np.zeros((32, 16, 16), np.int8)
        cnt = lib.ddafft_target_bisphere(32, 16, 16, ratio, _p(occ))
        assert cnt == int(np.count_nonzero(occ))
        assert set(np.unique(occ)) <= {0, 1, 2}


def test_polarizability_bitexact(lib):
    g = golden("physics.npz")
    lib.ddafft_polarizability.argtypes = [C.c_int, C.c_float, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    kinc = np.ascontiguousarray(g["pol_kinc"], np.float32)
    einc = np.ascontiguousarray(g["pol_einc"], np.float32)
    for a in range(5):
        for b, m in enumerate(g["pol_m"]):
            for c, kd in enumerate(g["pol_kd"]):
                mm = np.array([m.real, m.imag], np.float32)
                out = np.zeros(2, np.float32)
                assert lib.ddafft_polarizability(a + 1, C.c_float(kd), _p(mm), _p(kinc), _p(einc), _p(out)) == 0
                want = np.array([g["pol"][a, b, c].real, g["pol"][a, b, c].imag], np.float32)
                assert np.array_equal(bits(out), bits(want)), (a, m, kd)


def test_muller_bitexact(lib):
    g = golden("physics.npz")
    for t in range(g["mu_out"].shape[0]):
        args = [np.ascontiguousarray(g[k][t]) for k in ("mu_k", "mu_s", "mu_n")]
        args = [a.astype(np.float32) for a in args]
        cargs = [np.ascontiguousarray(g[k][t], np.complex64) for k in ("mu_e0", "mu_e1", "mu_s0", "mu_s1")]
        out = np.zeros((4, 4), np.float32)
        assert lib.ddafft_muller(*[_p(a) for a in args], *[_p(a) for a in cargs], _p(out)) == 0
        assert np.array_equal(bits(out), bits(g["mu_out"][t])), t


def geometry(lib, nplanes, theta_min, theta_max, step, psi=0.0, iota=0.0, kappa=0.0):
    lib.ddafft_geometry_fixed.argtypes = [C.c_int, C.c_int] + [C.c_double] * 5 + [C.c_void_p] * 4
    tmin, tmax, tst = np.float32(theta_min), np.float32(theta_max), np.float32(step)
    v = np.float32(np.float32(tmax - tmin) / tst) + np.float32(1.0)
    frac, whole = np.modf(np.float64(v))
    nang = int(whole) + (1 if frac > 0.5 else 0)
    dirs = np.zeros((nplanes, nang, 3), np.float32)
    normals = np.zeros((nplanes, 3), np.float32)
    kinc = np.zeros(3, np.float32)
    einc = np.zeros((2, 3), np.complex64)
    assert lib.ddafft_geometry_fixed(nplanes, nang, theta_min, step, psi, iota, kappa, _p(dirs), _p(normals),
                                     _p(kinc), _p(einc)) == 0
    return dirs, normals, kinc, einc


def test_geometry_first_plane(lib):
    dirs, normals, kinc, einc = geometry(lib, 4, 0, 180, 10)
    assert dirs.shape == (4, 19, 3)
    np.testing.assert_array_equal(dirs[0], directions(np.arange(0, 181, 10)))
    np.testing.assert_array_equal(kinc, [0, 0, 1])
    np.testing.assert_array_equal(einc, np.eye(3)[:2])
    np.testing.assert_allclose(normals[2], [1, 0, 0], atol=1e-6)   # Rotation_z (dda_func.cpp:27-34) of (0,1,0) by 90 deg
    # every direction of plane m is perpendicular to that plane's normal
    for m in range(4):
        np.testing.assert_allclose(dirs[m] @ normals[m], 0, atol=1e-6)


def test_geometry_rotated_orthonormal(lib):
    dirs, normals, kinc, einc = geometry(lib, 3, 0, 180, 15, psi=30, iota=40, kappa=50)
    np.testing.assert_allclose(np.linalg.norm(kinc), 1, atol=1e-6)
    for p in range(2):
        np.testing.assert_allclose(np.abs(einc[p]) @ np.abs(einc[p]), 1, atol=1e-6)
        np.testing.assert_allclose(einc[p].real @ kinc, 0, atol=1e-6)
    np.testing.assert_allclose(einc[0].real @ einc[1].real, 0, atol=1e-6)
