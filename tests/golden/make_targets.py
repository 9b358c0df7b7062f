"""Writes tests/golden/targets.npz: occupancy maps of the reference's random
and parametric target generators (particle.cpp:21-995), produced by the
reference's own functions (oracle/_ref/libddaref.so, built from its sources in
this container) after dda_srand(seed), with the next draw of its stream.

    python tests/golden/make_targets.py
"""

import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(os.path.dirname(os.path.dirname(HERE)), "oracle")]

import oracle  # noqa: E402

CASES = [(seed, dims, kind) for seed in (7, 2024) for dims in ((16, 16, 16), (20, 16, 12))
         for kind in (5, 9, 10, 13, 14, 6, 7, 8)]
NSEEDS = (5, 3, 2)
MV = (0, 1, 0, 1, 1)
BITES, DEPTH, LEVELS = 4, 0.3, 2
RW, LR, RHO = 1.5, 2.0, 0.02


def ref_lib():
    lib = oracle.Reference().lib
    P = C.c_void_p
    lib.ref_target_rand.argtypes = ([C.c_int] * 4 + [C.c_float] * 2 + [C.c_int, P, C.c_int, C.c_float, P, C.c_int,
                                    C.c_float, C.c_float, C.c_float, C.c_int, C.c_ulonglong, P])
    return lib


def worm_count(ni, nj, nl):
    r, rw, lr = np.float32(RHO), np.float32(RW), np.float32(LR)
    return int(float(r * np.float32(ni * nj * nl)) / (8.0 * float(rw * rw * rw) * (0.75 * float(lr) + 1))) + 1


def main():
    lib = ref_lib()
    out = {}
    for n, (seed, dims, kind) in enumerate(CASES):
        ni, nj, nl = dims
        occ = np.full(dims, 99, np.int8)
        ns = np.array(NSEEDS, np.int32)
        mv = np.array(MV, np.int32)
        b_a, c_a = nj / ni * 0.9, nl / ni * 0.9
        nw = worm_count(ni, nj, nl)
        lw = float(np.float32(RW) * np.float32(LR))
        cnt = lib.ref_target_rand(kind, ni, nj, nl, b_a, c_a, 3, ns.ctypes.data_as(C.c_void_p), BITES, DEPTH,
                                  mv.ctypes.data_as(C.c_void_p), LEVELS, RHO, RW, lw, nw, seed,
                                  occ.ctypes.data_as(C.c_void_p))
        nxt = lib.ref_rand_next()
        out["occ_%d" % n] = occ
        out["meta_%d" % n] = np.array([seed, ni, nj, nl, kind, cnt, nxt, nw], np.int64)
    np.savez_compressed(os.path.join(HERE, "targets.npz"), **out)
    print("wrote", len(CASES), "cases")


if __name__ == "__main__":
    main()
