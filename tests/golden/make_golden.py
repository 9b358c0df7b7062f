"""Writes the golden fixtures of tests/golden/ from the reference build.

Run in the container that holds the reference tree (it is never run on the
GPU host):  python tests/golden/make_golden.py

It builds oracle/_ref (the reference compiled from its own sources by
oracle/Makefile), calls the reference's functions through
oracle/ref_harness.cpp and runs the reference command-line program, and
stores only their inputs and outputs (numpy .npz without pickles, and the
reference's own out.dat text files).
"""

import os
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle as orc  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))
CLI = os.path.join(ROOT, "oracle", "_ref", "dda_fft.out")


def directions(theta_deg):
    """Scattering directions of the first plane (main.cpp:354-357)."""
    th = np.asarray(theta_deg, np.float32).astype(np.float64) * np.pi / 180.0
    return np.stack([np.sin(th), 0 * th, np.cos(th)], 1).astype(np.float32)


def rnd(rng, shape):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)


def fft_fixtures(ref):
    out = {}
    rng = np.random.default_rng(12345)
    for n in (4, 8, 16, 32, 64, 128, 256, 512):
        x = rnd(rng, (n, 3))
        out["x_%d" % n] = x
        out["fwd_%d" % n] = ref.fft_line(x)
        out["inv_%d" % n] = ref.fft_line(x, inverse=True)
        out["legacy_%d" % n] = ref.fft_line(x, legacy=True)
        out["phase_%d" % n] = ref.fourier_phase(n)
    np.savez_compressed(os.path.join(OUT, "fft_lines.npz"), **out)


def lattice_fixtures(ref):
    out = {}
    rng = np.random.default_rng(2024)
    for tag, dims, kd in (("a", (4, 4, 4), 0.5), ("b", (8, 4, 16), 0.7)):
        out["dims_" + tag] = np.array(dims)
        out["kd_" + tag] = np.float32(kd)
        out["amat_" + tag] = ref.amatrix(dims, kd)
        A = ref.build_ahat(dims, kd)
        out["ahat_" + tag] = A
        x = rnd(rng, dims + (3,))
        out["x_" + tag] = x
        out["spec_" + tag] = ref.fft3d(x)
        out["ifft_" + tag] = ref.ifft3d(out["spec_" + tag])
        out["conv_" + tag] = ref.convolution(A, x)
        out["convc_" + tag] = ref.convolution(A, x, conj=True)
    np.savez_compressed(os.path.join(OUT, "lattice.npz"), **out)


def solve_fixtures(ref):
    """DDA_FFT of all four solvers on small ellipsoids, both polarizations."""
    out = {}
    cases = [
        ("s16", (16, 16, 16), 3.0, (1.5 + 0.01j,), 1.0, 1.0, 1e-5, 8),
        ("e8", (8, 8, 16), 2.0, (1.7 + 0.1j,), 1.0, 1.0, 1e-5, 8),
        ("a32", (32, 32, 32), 5.0, (1.5 + 0.01j,), 1.0, 1.0, 1e-3, 100),
    ]
    for tag, dims, x, mats, b_a, c_a, tol, steps in cases:
        kd = np.float32(2 * x / dims[0])
        occ, cnt = ref.target(2, dims, b_a, c_a)
        alpha = np.array([0] + [ref.polarizability(1, kd, m) for m in mats], np.complex64)
        mv = np.array([0] + [1] * len(mats), np.int32)
        A = ref.build_ahat(dims, kd)
        dirs = directions(np.arange(0, 181, 10))
        kinc = np.array([0, 0, 1], np.float32)
        out["%s_occ_count" % tag] = np.int32(cnt)
        out["%s_kd" % tag] = kd
        out["%s_alpha" % tag] = alpha
        solvers = (0, 1, 2, 3) if tag != "a32" else (2,)
        for solver in solvers:
            for pol in (0, 1):
                einc = np.zeros(3, np.complex64)
                einc[pol] = 1
                with tempfile.NamedTemporaryFile("r", suffix=".log") as log:
                    r = ref.dda_fft(solver, occ, mv, alpha, kd, kinc, einc, A, dirs, max_step=steps, err_max=tol,
                                    tau=0.5, log_path=log.name)
                    steps_log = orc.parse_steps(open(log.name).read())
                key = "%s_sv%d_p%d" % (tag, solver, pol)
                out[key + "_esc"] = r["esc"]
                out[key + "_cabs"] = np.float32(r["cabs"])
                out[key + "_cext"] = np.float32(r["cext"])
                out[key + "_hist"] = np.array([v for _, v in steps_log])
                out[key + "_cfg"] = np.array([tol, steps])
    np.savez_compressed(os.path.join(OUT, "solves.npz"), **out)


def physics_fixtures(ref):
    out = {}
    ms = np.array([1.5 + 0.01j, 1.33 + 1e-4j, 1.7 + 0.1j, 1.6 + 0.005j, 1.55 + 0.02j], np.complex64)
    kds = np.array([0.3125, 0.5, 0.234375, 0.15625], np.float32)
    kinc = np.array([0.3, 0.4, np.sqrt(0.75)], np.float32)
    einc = np.array([[0.8, 0], [0.6, 0], [0, 0]], np.float32)
    pol = np.zeros((5, len(ms), len(kds)), np.complex64)
    for a in range(5):
        for b, m in enumerate(ms):
            for c, kd in enumerate(kds):
                pol[a, b, c] = ref.polarizability(a + 1, kd, m, kinc=kinc, einc=einc)
    out.update(pol_m=ms, pol_kd=kds, pol_kinc=kinc, pol_einc=einc, pol=pol)
    rng = np.random.default_rng(7)
    n = 64
    ks, ss, ns, e0, e1, s0, s1, mm = [], [], [], [], [], [], [], []
    for _ in range(n):
        k = rng.standard_normal(3); k /= np.linalg.norm(k)
        s = rng.standard_normal(3); s /= np.linalg.norm(s)
        nv = rng.standard_normal(3); nv -= k * np.dot(nv, k); nv /= np.linalg.norm(nv)
        a0 = rng.standard_normal(3).astype(np.complex64)
        a1 = rng.standard_normal(3).astype(np.complex64)
        b0 = rnd(rng, 3)
        b1 = rnd(rng, 3)
        ks.append(k); ss.append(s); ns.append(nv); e0.append(a0); e1.append(a1); s0.append(b0); s1.append(b1)
        mm.append(ref.muller(k, s, nv, a0, a1, b0, b1))
    out.update(mu_k=np.array(ks, np.float32), mu_s=np.array(ss, np.float32), mu_n=np.array(ns, np.float32),
               mu_e0=np.array(e0), mu_e1=np.array(e1), mu_s0=np.array(s0), mu_s1=np.array(s1),
               mu_out=np.array(mm, np.float32))
    tg = {}
    for dims, b_a, c_a in (((32, 32, 32), 1.0, 1.0), ((64, 64, 64), 1.0, 0.5), ((32, 32, 64), 1.0, 2.0),
                           ((16, 8, 32), 0.5, 1.0)):
        occ, cnt = ref.target(2, dims, b_a, c_a)
        key = "ell_%d_%d_%d_%g_%g" % (dims + (b_a, c_a))
        tg[key] = np.packbits(occ.reshape(-1).astype(np.uint8))
        tg[key + "_count"] = np.int32(cnt)
    for dims, b_a, c_a in (((16, 16, 16), 0.5, 0.75),):
        occ, cnt = ref.target(3, dims, b_a, c_a)
        key = "par_%d_%d_%d_%g_%g" % (dims + (b_a, c_a))
        tg[key] = np.packbits(occ.reshape(-1).astype(np.uint8))
        tg[key + "_count"] = np.int32(cnt)
    out.update(tg)
    np.savez_compressed(os.path.join(OUT, "physics.npz"), **out)


CLI_CASES = {
    # the plain CPU reference configuration (32^3 sphere, x = 5)
    "sphere32": "-t sphere -g 32 -x 5 -m 1.5 0.01 -P cm -M bicgstab -e 1e-3 -S 100 -O fixed --psi 0 --iota 0 "
                "--kappa 0 -p 1 -a 0 -b 180 -s 10 -n 1",
    # the single-instance GPU configuration (64^3 oblate spheroid, 4 planes, 1 degree)
    "spheroid64": "-t ellipsoid -g 64 --aspect 1 0.5 -x 10 -m 1.6 0.005 -P cm -M bicgstab -e 1e-3 -S 300 -p 4 "
                  "-a 0 -b 180 -s 1 -n 1",
    # every solver and polarizability model on a small lattice
    "cg16": "-t ellipsoid -g 16 --aspect 1 0.5 -x 3 -m 1.5 0.01 -P cm -M cg -e 1e-4 -S 30 -p 2 -s 15",
    "aa16": "-t sphere -g 16 -x 3 -m 1.5 0.01 -P cmrr -M aa -e 1e-4 -S 30 -p 2 -s 15",
    "simple16": "-t sphere -g 16 -x 2 -m 1.2 0.01 -P cmnr -M simple --tau 0.8 -e 1e-4 -S 15 -p 2 -s 15",
    "ldr16": "-t ellipsoid -g 16 16 32 --aspect 1 2 -x 3 -m 1.33 0.001 -P ldr -M bicgstab -e 1e-4 -S 40 -p 3 -s 20",
    "ildr16": "-t parallelepiped -g 16 --aspect 0.5 0.75 -x 2 -m 1.7 0.1 -P ildr -M bicgstab -e 1e-4 -S 40 "
              "-p 2 -s 20 --psi 30 --iota 40 --kappa 50",
    # the sample loop: random targets regenerated per sample from the stream,
    # Monte-Carlo orientations, worm bodies
    "rand16": "-t random-ellipsoid -g 16 --aspect 1 0.75 -x 3 -m 1.5 0.01 -m 1.7 0.05 --seeds 3 2 -P cm "
              "-M bicgstab -e 1e-4 -S 60 -p 2 -s 20 -n 3",
    "mc16": "-t ellipsoid -g 16 --aspect 1 0.5 -x 3 -m 1.5 0.01 -P cm -M bicgstab -e 1e-4 -S 60 -O mc -p 2 -s 20 "
            "-n 4",
    "worms16": "-t anthrax -g 16 -x 2 -m 1.5 0.01 --monomer-radius 2 --monomer-length-ratio 1 --packing 0.05 "
               "-M bicgstab -e 1e-4 -S 60 -p 1 -s 30 -n 2",
    # the remaining target kinds, one short run each
    "rsphere16": "-t random-sphere -g 16 -x 2 -m 1.5 0.01 -m 1.7 0.05 --seeds 3 2 -M bicgstab -e 1e-4 -S 60 "
                 "-p 1 -s 30 -n 2",
    "rpar16": "-t random-parallelepiped -g 16 --aspect 1 0.75 -x 2 -m 1.5 0.01 -m 1.7 0.05 --seeds 2 2 "
              "-M bicgstab -e 1e-4 -S 60 -p 1 -s 30 -n 2",
    "nibble16": "-t nibble-ellipsoid -g 16 --aspect 1 1 -x 2 -m 1.5 0.01 --seeds 3 --bites 2 --depth 0.2 "
                "-M bicgstab -e 1e-4 -S 60 -p 1 -s 30 -n 2",
    "hier16": "-t hierarchical -g 16 --aspect 1 1 -x 2 -m 1.5 0.01 -m 1.7 0.05 --seeds 2 3 --levels 2 "
              "-M bicgstab -e 1e-4 -S 60 -p 1 -s 30 -n 2",
    "bisph16": "-t bisphere -g 32 16 16 --radius-ratio 1.5 -x 3 -m 1.5 0.01 -m 1.7 0.05 -M bicgstab -e 1e-4 "
               "-S 60 -p 1 -s 30",
    "hedgehog16": "-t hedgehog -g 16 -x 2 -m 1.5 0.01 --monomer-radius 1.5 --monomer-length-ratio 2 --packing 0.05 "
                  "-M bicgstab -e 1e-4 -S 60 -p 1 -s 30 -n 2",
    "dead16": "-t dead-anthrax -g 16 -x 2 -m 1.5 0.01 --monomer-radius 2 --monomer-length-ratio 1 --packing 0.05 "
              "-M bicgstab -e 1e-4 -S 60 -p 1 -s 30 -n 2",
    "spec16": "-t sphere-spec -g 16 -x 2 -m 1.5 0.01 --packing 0.5 -M bicgstab -e 1e-4 -S 60 -p 1 -s 30",
}


def cli_fixtures(names=None):
    for name, args in CLI_CASES.items():
        if names and name not in names:
            continue
        with tempfile.TemporaryDirectory() as d:
            cmd = [CLI] + args.split() + ["-j", "1", "--seed", "1", "-o", "out.dat", "-D", "out.dst"]
            res = subprocess.run(cmd, cwd=d, capture_output=True, text=True, timeout=300)
            if res.returncode != 0:
                raise RuntimeError("reference run failed: %s\n%s" % (name, res.stdout[-2000:]))
            with open(os.path.join(d, "out.dat")) as f:
                dat = f.read()
            with open(os.path.join(d, "out.dst"), "rb") as f:
                dst = f.read()
        with open(os.path.join(OUT, "%s.out.dat" % name), "w") as f:
            f.write(dat)
        with open(os.path.join(OUT, "%s.out.dst" % name), "wb") as f:
            f.write(dst)
        with open(os.path.join(OUT, "%s.steps.txt" % name), "w") as f:
            f.write("# reference arguments: %s\n" % args)
            for line in res.stdout.splitlines():
                if "discrepancy=" in line or "replaced" in line:
                    f.write(line.split("\ttime=")[0].strip() + "\n")


def main():
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "-s", "ref"], check=True)
    if "--cli" in sys.argv:   # only the named command-line runs
        cli_fixtures(sys.argv[sys.argv.index("--cli") + 1:])
        return
    ref = orc.Reference()
    fft_fixtures(ref)
    lattice_fixtures(ref)
    solve_fixtures(ref)
    physics_fixtures(ref)
    cli_fixtures()
    print("golden fixtures written to", OUT)


if __name__ == "__main__":
    main()
