"""Writes tests/golden/<case>.fp64.npz for the seven reference CLI runs: the
Mueller matrices, Cabs and Cext of the same sample with the far-field and
cross-section sums taken in float64 over the same float dipoles.

The dipoles come from the oracle's float solve, which reproduces the
reference's bit for bit (and the HIP library's on axis-aligned incidence).
The reference then sums 10^4..10^5 float terms sequentially; on strongly
cancelling directions (backscatter) that float sum carries relative errors up
to ~1e-2 of small Mueller elements, so the reference's own out.dat is not the
accurate value there.  These fixtures are: geometry, polarizabilities and
solve as the reference (main.cpp, dda_func.cpp), far field and cross sections
in float64, Jones->Mueller from the far fields rounded to complex64 as the
reference rounds them (mmuller.cpp:175-314 in float).

The second set, <case>.f64solve.npz, is the whole sample in FP64: the
oracle's double build (kd, polarizabilities, A-hat, FFTs, solver, far field,
Cabs/Cext, Jones->Mueller all in double) on the same occupancy and geometry.
It gives the error of the float path against an FP64 build of the same solve.

    python tests/golden/make_cli_fp64.py
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "oracle"), os.path.join(ROOT, "dda-fft_amd")]

import ddafft  # noqa: E402  (host functions only: geometry, polarizability, targets)
import gradient as G  # noqa: E402
import oracle  # noqa: E402

SOLVERS = {"cg": 0, "aa": 1, "bicgstab": 2, "simple": 3}
MODELS = {"cm": 1, "cmrr": 2, "cmnr": 3, "ildr": 4, "ldr": 5}


def parse(args):
    a = args.split()
    o = dict(t="sphere", g=None, aspect=(1.0, 1.0), x=None, m=None, P="cm", M="cg", e=1e-3, S=20, p=1, a=0.0,
             b=180.0, s=10.0, tau=0.0, psi=0.0, iota=0.0, kappa=0.0)
    i = 0
    while i < len(a):
        k = a[i]
        if k == "-g":
            v = []
            while i + 1 < len(a) and a[i + 1].lstrip("-").replace(".", "").isdigit():
                v.append(int(a[i + 1]))
                i += 1
            o["g"] = tuple(v * 3 if len(v) == 1 else v)
        elif k == "--aspect":
            o["aspect"] = (float(a[i + 1]), float(a[i + 2]))
            i += 2
        elif k == "-m":
            o["m"] = complex(float(a[i + 1]), float(a[i + 2]))
            i += 2
        else:
            o[k.lstrip("-")] = a[i + 1]
            i += 1
        i += 1
    return o


def case(name, args):
    o = parse(args)
    dims = o["g"]
    x = np.float32(float(o["x"]))
    kd = np.float32(2.0 * x / np.float32(dims[0]))
    t = o["t"]
    ba, ca = [float(np.float32(v)) for v in o["aspect"]]
    if t == "sphere":
        occ, _ = ddafft.target_ellipsoid(dims, 1.0, 1.0)
    elif t == "ellipsoid":
        occ, _ = ddafft.target_ellipsoid(dims, ba, ca)
    else:
        import ctypes as C
        lib = ddafft.load()
        occ = np.zeros(dims, np.int8)
        lib.ddafft_target_parallelepiped(*dims, ba, ca, occ.ctypes.data_as(C.c_void_p))
    npl, tmin, tstep = int(o["p"]), float(o["a"]), float(o["s"])
    nang = int(round((float(o["b"]) - tmin) / tstep + 1))
    dirs, normals, kinc, einc = ddafft.geometry_fixed(npl, nang, tmin, tstep, float(o["psi"]), float(o["iota"]),
                                                      float(o["kappa"]))
    dirs = dirs.reshape(-1, 3)
    model, solver = MODELS[o["P"]], SOLVERS[o["M"]]
    orc = oracle.Oracle("f32")
    A = orc.build_ahat(dims, kd)
    esc, cabs, cext = [], [], []
    for p in range(2):
        alpha = np.array([0, ddafft.polarizability(model, kd, o["m"], kinc, einc[p])], np.complex64)
        out = orc.dda_fft(solver, occ, [0, 1], alpha, kd, kinc, einc[p], A, dirs, int(o["S"]), float(o["e"]),
                          float(o["tau"]), return_p=True)
        ce, cb, e = G.observables(occ, [0, 1], alpha, kd, kinc.astype(np.float64), einc[p], out["p"], dirs)
        esc.append(e.astype(np.complex64))
        cabs.append(cb)
        cext.append(ce)
    mm = np.zeros((nang, 4, 4))
    for a in range(nang):
        acc = np.zeros((4, 4), np.float32)
        for pl in range(npl):
            d = pl * nang + a
            acc = acc + orc.muller(kinc, dirs[d], normals[pl], einc[0], einc[1], esc[0][d], esc[1][d])
        mm[a] = acc / np.float32(npl)
    np.savez_compressed(os.path.join(HERE, name + ".fp64.npz"), mm=mm, cabs=0.5 * (cabs[0] + cabs[1]),
                        cext=0.5 * (cext[0] + cext[1]))
    print(name, "done", flush=True)


def case_f64(name, args):
    o = parse(args)
    dims = o["g"]
    kd = 2.0 * float(o["x"]) / float(dims[0])
    t = o["t"]
    ba, ca = [float(np.float32(v)) for v in o["aspect"]]
    if t == "sphere":
        occ, _ = ddafft.target_ellipsoid(dims, 1.0, 1.0)
    elif t == "ellipsoid":
        occ, _ = ddafft.target_ellipsoid(dims, ba, ca)
    else:
        import ctypes as C
        lib = ddafft.load()
        occ = np.zeros(dims, np.int8)
        lib.ddafft_target_parallelepiped(*dims, ba, ca, occ.ctypes.data_as(C.c_void_p))
    npl, tmin, tstep = int(o["p"]), float(o["a"]), float(o["s"])
    nang = int(round((float(o["b"]) - tmin) / tstep + 1))
    dirs, normals, kinc, einc = ddafft.geometry_fixed(npl, nang, tmin, tstep, float(o["psi"]), float(o["iota"]),
                                                      float(o["kappa"]))
    dirs = dirs.reshape(-1, 3).astype(np.float64)
    normals = normals.astype(np.float64)
    kinc = kinc.astype(np.float64)
    einc = einc.astype(np.complex128)
    model, solver = MODELS[o["P"]], SOLVERS[o["M"]]
    orc = oracle.Oracle("f64")
    A = orc.build_ahat(dims, kd)
    esc, cabs, cext, steps = [], [], [], []
    for p in range(2):
        a = orc.polarizability(model, kd, o["m"], kinc, np.stack([einc[p].real, einc[p].imag], 1))
        alpha = np.array([0, a], np.complex128)
        out = orc.dda_fft(solver, occ, [0, 1], alpha, kd, kinc, einc[p], A, dirs, int(o["S"]), float(o["e"]),
                          float(o["tau"]))
        esc.append(out["esc"])
        cabs.append(out["cabs"])
        cext.append(out["cext"])
        steps.append(out["steps"])
    mm = np.zeros((nang, 4, 4))
    for a in range(nang):
        acc = np.zeros((4, 4))
        for pl in range(npl):
            d = pl * nang + a
            acc = acc + orc.muller(kinc, dirs[d], normals[pl], einc[0], einc[1], esc[0][d], esc[1][d])
        mm[a] = acc / npl
    np.savez_compressed(os.path.join(HERE, name + ".f64solve.npz"), mm=mm, cabs=0.5 * (cabs[0] + cabs[1]),
                        cext=0.5 * (cext[0] + cext[1]), steps=np.array(steps))
    print(name, "f64 solve done, steps", steps, flush=True)


TARGETS = {"sphere": 1, "ellipsoid": 2, "parallelepiped": 3, "random-sphere": 4, "random-ellipsoid": 5,
           "anthrax": 6, "hedgehog": 7, "dead-anthrax": 8, "random-parallelepiped": 9, "nibble-ellipsoid": 10,
           "bisphere": 12, "hierarchical": 13, "sphere-spec": 14}


def parse_samples(args):
    """The options of a sample-loop run (repeated -m, --seeds lists)."""
    a = args.split()
    o = dict(t="sphere", aspect=(1.0, 1.0), m=[], seeds=[], P="cm", M="cg", e=1e-3, S=20, p=1, a=0.0, b=180.0,
             s=10.0, tau=0.0, n=1, O="fixed")
    i = 0
    while i < len(a):
        k = a[i]
        if k in ("-g", "--seeds"):
            v = []
            while i + 1 < len(a) and a[i + 1].replace(".", "").isdigit():
                v.append(int(a[i + 1]))
                i += 1
            if k == "-g":
                o["g"] = tuple(v * 3 if len(v) == 1 else v)
            else:
                o["seeds"] = v
        elif k == "--aspect":
            o["aspect"] = (float(a[i + 1]), float(a[i + 2]))
            i += 2
        elif k == "-m":
            o["m"].append(complex(float(a[i + 1]), float(a[i + 2])))
            i += 2
        else:
            o[k.lstrip("-")] = a[i + 1]
            i += 1
        i += 1
    return o


def case_samples(name, args):
    """A run of the sample loop (main.cpp:774-996) with seed 1 and one worker:
    targets regenerated from the stream per sample, Monte-Carlo orientations
    drawn after them, float solves as the reference's (the oracle), far fields
    and cross sections in float64, samples averaged in float64."""
    import ctypes as C
    o = parse_samples(args)
    dims = o["g"]
    ni, nj, nl = dims
    t = TARGETS[o["t"]]
    x = np.float32(float(o["x"]))
    kd = np.float32(2.0 * x / np.float32(ni))
    lib = ddafft.load()
    P_, I_, D_ = C.c_void_p, C.c_int, C.c_double
    mats = o["m"]
    nsorts = len(mats)
    mv = np.array([0] + [0 if m == 1 + 0j else 1 for m in mats], np.int32)
    seeds = np.array(o["seeds"] or [0], np.int32)
    st = np.array([lib.ddafft_rng_worker_seed(C.c_ulonglong(1), 0)], np.uint64)
    if int(st[0]) == 0:
        st[0] = 0x9E3779B97F4A7C15
    ba, ca = [float(np.float32(v)) for v in o["aspect"]]
    occ = np.zeros(dims, np.int8)
    pp = lambda arr: arr.ctypes.data_as(P_)   # noqa: E731

    def generate():
        if t == 4:
            return lib.ddafft_target_random_ellipsoid(I_(ni), I_(nj), I_(nl), D_(1.0), D_(1.0), I_(nsorts), pp(seeds),
                                                      pp(st), pp(occ))
        if t == 5:
            return lib.ddafft_target_random_ellipsoid(I_(ni), I_(nj), I_(nl), D_(ba), D_(ca), I_(nsorts), pp(seeds),
                                                      pp(st), pp(occ))
        if t == 9:
            return lib.ddafft_target_random_parallelepiped(I_(ni), I_(nj), I_(nl), D_(ba), D_(ca), I_(nsorts),
                                                           pp(seeds), pp(st), pp(occ))
        if t == 10:
            return lib.ddafft_target_nibble_ellipsoid(I_(ni), I_(nj), I_(nl), D_(ba), D_(ca), I_(nsorts), pp(seeds),
                                                      I_(int(o["bites"])), D_(float(np.float32(float(o["depth"])))),
                                                      pp(mv), pp(st), pp(occ))
        if t == 13:
            return lib.ddafft_target_hierarchical_ellipsoid(I_(int(o["levels"])), I_(ni), I_(nj), I_(nl), D_(ba),
                                                            D_(ca), I_(nsorts), pp(seeds), pp(st), pp(occ))
        if t in (6, 7, 8):
            rw = float(o["monomer-radius"])
            lr = float(o["monomer-length-ratio"])
            lib.ddafft_worm_count.argtypes = [I_] * 3 + [D_] * 3
            nw = lib.ddafft_worm_count(ni, nj, nl, float(o["packing"]), rw, lr)
            lw = float(np.float32(rw) * np.float32(lr))
            lib.ddafft_target_worms.argtypes = [I_] * 4 + [D_] * 2 + [I_, C.c_longlong, P_, P_]
            return lib.ddafft_target_worms(t, ni, nj, nl, rw, lw, nw, 0, pp(st), pp(occ))
        if t == 12:
            return lib.ddafft_target_bisphere(I_(ni), I_(nj), I_(nl), D_(float(np.float32(float(o["radius-ratio"])))),
                                              pp(occ))
        if t == 14:
            return lib.ddafft_target_sphere_spec(I_(ni), I_(nj), I_(nl), D_(ni / 2.0),
                                                 D_(float(np.float32(float(o["packing"])))), pp(occ))
        occ[...], _ = ddafft.target_ellipsoid(dims, 1.0, 1.0) if t == 1 else ddafft.target_ellipsoid(dims, ba, ca)
        return int((occ != 0).sum())

    fixed_target = t not in (4, 5, 6, 7, 8, 9, 10, 13)
    if fixed_target:
        assert generate() >= 0
    npl, tmin, tstep = int(o["p"]), float(o["a"]), float(o["s"])
    nang = int(round((float(o["b"]) - tmin) / tstep + 1))
    dirs0, normals0, kinc0, einc0 = ddafft.geometry_fixed(npl, nang, tmin, tstep)
    dirs0 = dirs0.reshape(-1, 3)
    model, solver = MODELS[o["P"]], SOLVERS[o["M"]]
    orc = oracle.Oracle("f32")
    A = orc.build_ahat(dims, kd)
    nsamp = int(o["n"])
    mm = np.zeros((nang, 4, 4))
    cabs_s = cext_s = 0.0
    for _ in range(nsamp):
        if not fixed_target:
            assert generate() >= 0
        if o["O"] == "mc":
            R = ddafft.orientation_mc(st)
            k = ddafft.rotate(R, kinc0)
            e = ddafft.rotate(R, np.real(einc0).astype(np.float32)).astype(np.complex64)
            d = ddafft.rotate(R, dirs0)
            nrm = ddafft.rotate(R, normals0)
        else:
            k, e, d, nrm = kinc0, einc0, dirs0, normals0
        esc, cab, cex = [], [], []
        for p in range(2):
            alpha = np.array([0] + [ddafft.polarizability(model, kd, m, k, e[p]) for m in mats], np.complex64)
            out = orc.dda_fft(solver, occ, mv, alpha, kd, k, e[p], A, d, int(o["S"]), float(o["e"]), float(o["tau"]),
                              return_p=True)
            ce, cb, es = G.observables(occ, mv, alpha, kd, k.astype(np.float64), e[p], out["p"], d)
            esc.append(es.astype(np.complex64))
            cab.append(cb)
            cex.append(ce)
        for a in range(nang):
            acc = np.zeros((4, 4))
            for pl in range(npl):
                q = pl * nang + a
                acc += orc.muller(k, d[q], nrm[pl], e[0], e[1], esc[0][q], esc[1][q])
            mm[a] += acc / npl
        cabs_s += 0.5 * (cab[0] + cab[1])
        cext_s += 0.5 * (cex[0] + cex[1])
    np.savez_compressed(os.path.join(HERE, name + ".fp64.npz"), mm=mm / nsamp, cabs=cabs_s / nsamp,
                        cext=cext_s / nsamp)
    print(name, "done (%d samples)" % nsamp, flush=True)


SAMPLE_RUNS = ("rand16", "worms16", "mc16", "rsphere16", "rpar16", "nibble16", "hier16", "bisph16", "hedgehog16",
               "dead16", "spec16")


def main():
    if "--samples" in sys.argv:
        for name in SAMPLE_RUNS:
            with open(os.path.join(HERE, name + ".steps.txt")) as f:
                case_samples(name, f.readline().split(":", 1)[1].strip())
        return
    for name in ("sphere32", "spheroid64", "cg16", "aa16", "simple16", "ldr16", "ildr16"):
        with open(os.path.join(HERE, name + ".steps.txt")) as f:
            args = f.readline().split(":", 1)[1].strip()
        if "--f64-solve-only" not in sys.argv:
            case(name, args)
        case_f64(name, args)


if __name__ == "__main__":
    main()
