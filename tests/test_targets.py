"""Host target generators, random stream and orientation (no GPU).

Bit-exact against the reference's own generators: the committed fixtures
(tests/golden/targets.npz, made by tests/golden/make_targets.py from the
reference's functions) always, and the compiled reference itself over more
seeds and lattices when oracle/_ref is built in this container.
"""

import ctypes as C
import os

import numpy as np
import pytest

import ddafft

HERE = os.path.dirname(os.path.abspath(__file__))
P = C.c_void_p


def _p(a):
    return a.ctypes.data_as(P)


@pytest.fixture(scope="module")
def lib():
    lib = ddafft.load()
    lib.ddafft_rng_next.argtypes = [P]
    lib.ddafft_rng_worker_seed.argtypes = [C.c_ulonglong, C.c_int]
    lib.ddafft_rng_worker_seed.restype = C.c_ulonglong
    for f in ("ddafft_target_random_ellipsoid", "ddafft_target_random_parallelepiped"):
        getattr(lib, f).argtypes = [C.c_int] * 3 + [C.c_double] * 2 + [C.c_int, P, P, P]
    lib.ddafft_target_nibble_ellipsoid.argtypes = [C.c_int] * 3 + [C.c_double] * 2 + [C.c_int, P, C.c_int,
                                                                                       C.c_double, P, P, P]
    lib.ddafft_target_hierarchical_ellipsoid.argtypes = [C.c_int] * 4 + [C.c_double] * 2 + [C.c_int, P, P, P]
    lib.ddafft_target_sphere_spec.argtypes = [C.c_int] * 3 + [C.c_double] * 2 + [P]
    lib.ddafft_target_worms.argtypes = [C.c_int] * 4 + [C.c_double] * 2 + [C.c_int, C.c_longlong, P, P]
    lib.ddafft_worm_count.argtypes = [C.c_int] * 3 + [C.c_double] * 3
    lib.ddafft_orientation_mc.argtypes = [P, P]
    lib.ddafft_rotate.argtypes = [P, C.c_int, P, P]
    lib.ddafft_target_file_read.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, P]
    lib.ddafft_target_file_write.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, P]
    return lib


NSEEDS = (5, 3, 2)
MV = (0, 1, 0, 1, 1)
BITES, DEPTH, LEVELS = 4, 0.3, 2
RW, LR, RHO = 1.5, 2.0, 0.02


def generate(lib, kind, dims, seed, nw=None):
    """The library's generator for `kind` with the fixture parameters;
    returns (occupancy, return value, next draw of the stream)."""
    ni, nj, nl = dims
    occ = np.full(dims, 99, np.int8)
    st = np.array([seed], np.uint64)
    ns = np.array(NSEEDS, np.int32)
    mv = np.array(MV, np.int32)
    b_a, c_a = float(np.float32(nj / ni * 0.9)), float(np.float32(nl / ni * 0.9))
    if kind == 5:
        cnt = lib.ddafft_target_random_ellipsoid(ni, nj, nl, b_a, c_a, 3, _p(ns), _p(st), _p(occ))
    elif kind == 9:
        cnt = lib.ddafft_target_random_parallelepiped(ni, nj, nl, b_a, c_a, 3, _p(ns), _p(st), _p(occ))
    elif kind == 10:
        cnt = lib.ddafft_target_nibble_ellipsoid(ni, nj, nl, b_a, c_a, 3, _p(ns), BITES, float(np.float32(DEPTH)),
                                                 _p(mv), _p(st), _p(occ))
    elif kind == 13:
        cnt = lib.ddafft_target_hierarchical_ellipsoid(LEVELS, ni, nj, nl, b_a, c_a, 3, _p(ns), _p(st), _p(occ))
    elif kind == 14:
        cnt = lib.ddafft_target_sphere_spec(ni, nj, nl, ni / 2.0, float(np.float32(RHO)), _p(occ))
    else:
        nw = lib.ddafft_worm_count(ni, nj, nl, RHO, RW, LR) if nw is None else nw
        lw = float(np.float32(RW) * np.float32(LR))
        cnt = lib.ddafft_target_worms(kind, ni, nj, nl, RW, lw, nw, 0, _p(st), _p(occ))
    nxt = lib.ddafft_rng_next(_p(st))
    return occ, cnt, nxt


def test_fixtures_bit_exact(lib):
    fx = np.load(os.path.join(HERE, "golden", "targets.npz"))
    n = len([k for k in fx.files if k.startswith("meta_")])
    assert n == 32
    for c in range(n):
        seed, ni, nj, nl, kind, cnt, nxt, nw = [int(v) for v in fx["meta_%d" % c]]
        occ, got, gnext = generate(lib, kind, (ni, nj, nl), seed)
        if kind in (6, 7, 8):
            assert lib.ddafft_worm_count(ni, nj, nl, RHO, RW, LR) == nw
        assert got == cnt, (seed, kind)
        assert np.array_equal(occ, fx["occ_%d" % c]), (seed, (ni, nj, nl), kind)
        if kind != 14:      # the partial sphere draws nothing
            assert gnext == nxt, (seed, kind)


def test_against_compiled_reference(lib):
    path = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "libddaref.so")
    if not os.path.exists(path):
        pytest.skip("reference not built in this container")
    ref = C.CDLL(path)
    ref.ref_target_rand.argtypes = ([C.c_int] * 4 + [C.c_float] * 2 + [C.c_int, P, C.c_int, C.c_float, P, C.c_int,
                                    C.c_float, C.c_float, C.c_float, C.c_int, C.c_ulonglong, P])
    for seed in (1, 99, 31415, 2 ** 40 + 17):
        for dims in ((16, 16, 32), (32, 24, 24)):
            ni, nj, nl = dims
            for kind in (5, 9, 10, 13, 14, 6, 7, 8):
                occ_r = np.full(dims, 99, np.int8)
                ns = np.array(NSEEDS, np.int32)
                mv = np.array(MV, np.int32)
                nw = lib.ddafft_worm_count(ni, nj, nl, RHO, RW, LR)
                lw = float(np.float32(RW) * np.float32(LR))
                cnt = ref.ref_target_rand(kind, ni, nj, nl, nj / ni * 0.9, nl / ni * 0.9, 3, _p(ns), BITES, DEPTH,
                                          _p(mv), LEVELS, RHO, RW, lw, nw, seed, _p(occ_r))
                nxt = ref.ref_rand_next()
                occ, got, gnext = generate(lib, kind, dims, seed)
                assert got == cnt and np.array_equal(occ, occ_r), (seed, dims, kind)
                if kind != 14:
                    assert gnext == nxt


def test_rng_stream(lib):
    # xorshift64* known answers, recomputed from the recurrence in Python
    def py_stream(s, n):
        out = []
        m = (1 << 64) - 1
        for _ in range(n):
            s ^= s >> 12
            s ^= (s << 25) & m
            s ^= s >> 27
            out.append(((s * 2685821657736338717) & m) >> 33)
        return out

    for seed in (1, 12345, 2 ** 63 + 5):
        st = np.array([seed], np.uint64)
        assert [lib.ddafft_rng_next(_p(st)) for _ in range(50)] == py_stream(seed, 50)
    st = np.array([0], np.uint64)   # dda_srand(0) uses the golden-ratio constant
    assert lib.ddafft_rng_next(_p(st)) == py_stream(0x9E3779B97F4A7C15, 1)[0]
    assert lib.ddafft_rng_worker_seed(10, 0) == (10 + 0x9E3779B97F4A7C15) % 2 ** 64
    assert lib.ddafft_rng_worker_seed(10, 2) == (10 + 3 * 0x9E3779B97F4A7C15) % 2 ** 64


def test_orientation_mc(lib):
    st = np.array([777], np.uint64)
    for _ in range(20):
        s0 = st.copy()
        rot = np.zeros(9, np.float32)
        assert lib.ddafft_orientation_mc(_p(st), _p(rot)) == 0
        R = rot.reshape(3, 3).astype(np.float64)
        np.testing.assert_allclose(R @ R.T, np.eye(3), atol=2e-6)
        assert abs(np.linalg.det(R) - 1) < 2e-6
        # three draws: cos(psi) = 2 u - 1 is the (3,3) element
        draws = [lib.ddafft_rng_next(_p(s0)) for _ in range(3)]
        cps = np.float32(np.float64(np.float32(draws[0])) * 2.0 / np.float64(np.float32(2147483647)) - 1.0)
        assert R[2, 2] == np.float64(cps)
        assert s0[0] == st[0]
        v = np.array([[0, 0, 1], [1, 0, 0]], np.float32)
        out = np.zeros_like(v)
        assert lib.ddafft_rotate(_p(rot), 2, _p(v), _p(out)) == 0
        np.testing.assert_allclose(out, (R @ v.T.astype(np.float64)).T, atol=1e-6)


def test_shape_file_roundtrip(lib, tmp_path):
    occ = np.zeros((6, 5, 4), np.int8)
    occ[1, 2, 3] = 1
    occ[5, 4, 0] = 2
    occ[0, 0, 0] = 1
    path = str(tmp_path / "shape.txt").encode()
    assert lib.ddafft_target_file_write(path, 6, 5, 4, _p(occ)) == 3
    back = np.full((6, 5, 4), 9, np.int8)
    assert lib.ddafft_target_file_read(path, 6, 5, 4, _p(back)) == 3
    assert np.array_equal(back, occ)
    # the reference's exit-3 case: a cell outside the lattice
    assert lib.ddafft_target_file_read(path, 5, 5, 4, _p(np.zeros((5, 5, 4), np.int8))) == -3
    # a short file
    (tmp_path / "short.txt").write_text("4\n1 1 1 1\n")
    assert lib.ddafft_target_file_read(str(tmp_path / "short.txt").encode(), 6, 5, 4, _p(back)) == -4


def test_worm_attempt_cap(lib):
    """Worms that cannot all fit: the reference loops forever (REPORT.md:1139-1211);
    with a cap the generator returns -3."""
    occ = np.zeros((8, 8, 8), np.int8)
    st = np.array([5], np.uint64)
    assert lib.ddafft_target_worms(6, 8, 8, 8, 1.5, 3.0, 500, 20000, _p(st), _p(occ)) == -3


# ---- the two benchmark generators without a reference target (mixture of the
# adjoint workload, agglomerates of the batched Monte-Carlo workload)

def _gen():
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(HERE), "oracle"))
    import generators
    return generators


@pytest.mark.parametrize("seed", [7, 2024])
def test_mixture_bit_exact(seed):
    gen = _gen()
    base, _ = ddafft.target_ellipsoid((32, 32, 64), 1.0, 2.0)
    st = np.array([seed], np.uint64)
    occ, n = ddafft.target_mixture(base, 2, 0.2, st)
    ref, s_end = gen.mixture(base, 2, 0.2, seed)
    assert np.array_equal(occ, ref) and int(st[0]) == s_end
    n1 = int((base == 1).sum())
    assert n == int((occ == 2).sum())
    assert abs(n - 0.2 * n1) < 5 * np.sqrt(n1 * 0.2 * 0.8)
    assert np.array_equal(occ != 0, base != 0)


@pytest.mark.parametrize("dims,seed", [((32, 32, 32), 1), ((32, 32, 64), 99), ((64, 64, 64), 5),
                                       ((32, 32, 64), 0xEA52A8A33F56EC11)])   # this one jams and starts over
def test_agglomerate(dims, seed):
    gen = _gen()
    st = np.array([seed], np.uint64)
    occ, cnt, cen = ddafft.target_agglomerate(dims, 3.0, 0.3, st)
    ref, rcen, s_end = gen.agglomerate(dims, 3.0, 0.3, seed)
    assert np.array_equal(occ, ref) and np.array_equal(cen, rcen) and int(st[0]) == s_end
    # invariants: count, inside the ball, disjoint, union of digital spheres
    K = gen.agglomerate_count(dims, 3.0, 0.3)
    assert len(cen) == K == int(np.floor(0.3 * (min(dims) / 6.0) ** 3 + 0.5))
    c0 = (np.array(dims) - 1) / 2.0
    assert np.all(np.linalg.norm(cen - c0, axis=1) <= min(dims) / 2.0 - 3.0)
    d = np.linalg.norm(cen[:, None] - cen[None], axis=2) + np.eye(K) * 100
    assert d.min() >= 6.0
    assert cnt == int(occ.sum())
    # disjoint monomers: each covers the same number of cells up to its lattice offset
    per = cnt / K
    assert 100 <= per <= 130   # digital ball of radius 3: 4/3 pi 27 = 113 +- boundary


def test_agglomerate_attempt_cap():
    st = np.array([3], np.uint64)
    with pytest.raises(ddafft.DdafftError):
        ddafft.target_agglomerate((16, 16, 16), 3.0, 0.9, st, max_attempts=10000)
