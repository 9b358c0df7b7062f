"""CPU restatement of the two workload generators that have no reference
target (test infrastructure only: the tests use it to check the library's
ddafft_target_mixture and ddafft_target_agglomerate).

They feed the reference through its target 11 shape file (main.cpp:645-677)
and draw from the reference's xorshift64* stream (randomize.cpp:44-58).
With no reference counterpart, parity rests on this restatement (bit-exact)
and on the invariants that tests/test_targets.py checks: monomers inside
the envelope and disjoint, occupancy equal to the union of the digital
spheres, mixture fraction within its binomial spread.
"""

import numpy as np

MASK = (1 << 64) - 1
STALL = 100000   # consecutive rejected centre draws before a packing starts over


class Stream:
    """xorshift64* with 31-bit outputs (randomize.cpp:51-58)."""

    def __init__(self, state):
        self.s = int(state) or 0x9E3779B97F4A7C15

    def next(self):
        t = self.s
        t ^= t >> 12
        t ^= (t << 25) & MASK
        t ^= t >> 27
        self.s = t
        return ((t * 2685821657736338717) & MASK) >> 33


def mixture(base, material, fraction, state):
    """Each cell of material 1 becomes `material` when its draw is below
    `fraction`, one draw per such cell in storage order."""
    rng = Stream(state)
    occ = np.array(base, np.int8, copy=True).reshape(-1)
    for k in np.flatnonzero(occ == 1):
        if rng.next() / 2147483647.0 < fraction:
            occ[k] = material
    return occ.reshape(np.shape(base)), rng.s


def agglomerate_count(dims, r_mono, packing):
    R = 0.5 * min(dims)
    return int(np.floor(packing * (R / r_mono) ** 3 + 0.5))


def agglomerate(dims, r_mono, packing, state):
    """Rejection-sampled disjoint spheres in the inscribed ball; returns
    (occ, centres, final state)."""
    ni, nj, nl = dims
    K = agglomerate_count(dims, r_mono, packing)
    R = 0.5 * min(dims)
    Rc = R - r_mono
    c0 = np.array([0.5 * (ni - 1), 0.5 * (nj - 1), 0.5 * (nl - 1)])
    rng = Stream(state)
    cen = []
    stall = 0
    while len(cen) < K:
        stall += 1
        if stall > STALL:   # jammed short of K: start over
            cen, stall = [], 0
        v = [Rc * (2.0 * (rng.next() / 2147483647.0) - 1.0) for _ in range(3)]
        if v[0] * v[0] + v[1] * v[1] + v[2] * v[2] > Rc * Rc:
            continue
        p = (c0[0] + v[0], c0[1] + v[1], c0[2] + v[2])
        if all((p[0] - q[0]) ** 2 + (p[1] - q[1]) ** 2 + (p[2] - q[2]) ** 2 >= 4.0 * r_mono * r_mono for q in cen):
            cen.append(p)
            stall = 0
    cen = np.array(cen, np.float64).reshape(-1, 3)
    i, j, l = np.meshgrid(np.arange(ni), np.arange(nj), np.arange(nl), indexing="ij")
    occ = np.zeros(dims, np.int8)
    for x, y, z in cen:
        occ[(i - x) ** 2 + (j - y) ** 2 + (l - z) ** 2 <= r_mono * r_mono] = 1
    return occ, cen, rng.s
