#!/usr/bin/env python3
"""Benchmark of the DDA-FFT solve path on one MI355X.

Workload (the reference's single-instance 64^3 configuration): homogeneous
oblate spheroid (b/a = 1, c/a = 0.5) in a 64x64x64 lattice (128^3 padded FFT),
x = 10, m = 1.6 + 0.005i, polarizability CM, BiCGStab to a mean per-dipole
residual of 1e-3 (max 300 steps), 2 incident polarizations, 4 scattering
planes, theta 0..180 step 1 (724 directions per polarization).

One step = one sample of that workload: both polarizations solved together as
two right-hand sides sharing Â, including the far field in all 724 directions
and Cabs/Cext, with every input resident in HBM (Â, occupancy and directions
are uploaded once before the timed region).  value = BiCGStab solves/s, one
solve being one (sample, polarization) DDA_FFT solve to the stopping residual.

Prints one JSON line.  `--gpus N` with N > 1 exits: this path runs on one GPU.
"""

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "dda-fft_amd"))

import ddafft  # noqa: E402

HBM_PEAK_GBS = 8000.0
GRID = (64, 64, 64)
X_SIZE, M_INDEX, ASPECT = 10.0, (1.6, 0.005), (1.0, 0.5)
NPLANES, THETA = 4, (0.0, 180.0, 1.0)
TOL, MAX_STEPS = 1e-3, 300


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def build_workload(lib):
    """Occupancy, polarizability and geometry through the library's host code."""
    ni, nj, nl = GRID
    kd = np.float32(2.0 * np.float32(X_SIZE) / np.float32(ni))            # main.cpp:318
    occ = np.zeros(GRID, np.int8)
    lib.ddafft_target_ellipsoid.argtypes = [C.c_int] * 3 + [C.c_double, C.c_double, C.c_void_p]
    ndip = lib.ddafft_target_ellipsoid(ni, nj, nl, ASPECT[0], ASPECT[1], _p(occ))
    lib.ddafft_polarizability.argtypes = [C.c_int, C.c_float, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    a = np.zeros(2, np.float32)
    m = np.array(M_INDEX, np.float32)
    lib.ddafft_polarizability(1, C.c_float(kd), _p(m), None, None, _p(a))
    alpha = np.array([0, a[0] + 1j * a[1]], np.complex64)
    nang = int(round((THETA[1] - THETA[0]) / THETA[2] + 1))
    dirs = np.zeros((NPLANES, nang, 3), np.float32)
    normals = np.zeros((NPLANES, 3), np.float32)
    kinc = np.zeros(3, np.float32)
    einc = np.zeros((2, 3), np.complex64)
    lib.ddafft_geometry_fixed.argtypes = [C.c_int, C.c_int] + [C.c_double] * 5 + [C.c_void_p] * 4
    lib.ddafft_geometry_fixed(NPLANES, nang, THETA[0], THETA[2], 0.0, 0.0, 0.0, _p(dirs), _p(normals), _p(kinc),
                              _p(einc))
    return dict(kd=kd, occ=occ, ndip=ndip, alpha=alpha, dirs=dirs.reshape(-1, 3), kinc=kinc, einc=einc)


def matvec_bytes(n, nrhs):
    """Algorithmic HBM bytes of one matvec pass set (complex64 3-vectors, full
    Â of 6 complex64 per padded cell, int8 occupancy), per pass."""
    return {
        "zf": nrhs * (24 + 48) * n,
        "yf": nrhs * (48 + 96) * n,
        "x": nrhs * (96 + 96) * n + 384 * n,
        "yi": nrhs * (96 + 48) * n,
        "zi": nrhs * (48 + 24 + 24) * n + n,
    }


# Batched Monte-Carlo workload: 128 random realizations on uneven lattices,
# (ni, nj, nl, count), largest first so the batch driver starts the long
# instances early; each an agglomerate of radius-3 sphere monomers at packing
# 0.3 in the inscribed ball, x ~ U(6, 12) (own kd and A-hat), m = 1.55+0.02i,
# CM, 8 Monte-Carlo orientations x 2 polarizations, the reference's default
# scattering geometry (1 plane, theta 0..180 step 10), BiCGStab to 1e-3.
MC_LATTICES = ((64, 64, 64, 16), (32, 32, 64, 48), (32, 32, 32, 64))
MC_ORIENT, MC_M, MC_X, MC_SEED = 8, (1.55, 0.02), (6.0, 12.0), 20240611


def mc_instances(frac):
    """The realizations (a fraction `frac` of each lattice group) with their
    right-hand sides, drawn from one xorshift64* stream."""
    lib = ddafft.load()
    st = np.array([MC_SEED], np.uint64)
    dirs0, _, kinc0, einc0 = ddafft.geometry_fixed(1, 19, 0.0, 10.0)
    dirs0 = dirs0.reshape(-1, 3)
    m = complex(*MC_M)
    out = []
    for ni, nj, nl, cnt in MC_LATTICES:
        for _ in range(max(1, int(round(cnt * frac)))):
            x = MC_X[0] + (MC_X[1] - MC_X[0]) * lib.ddafft_rng_next(_p(st)) / 2147483647.0
            kd = np.float32(2.0 * np.float32(x) / np.float32(ni))
            occ, ndip, _ = ddafft.target_agglomerate((ni, nj, nl), 3.0, 0.3, st)
            kinc, einc, dirs = [], [], []
            for _ in range(MC_ORIENT):
                R = ddafft.orientation_mc(st)
                k = ddafft.rotate(R, kinc0)
                e = ddafft.rotate(R, np.real(einc0).astype(np.float32))
                d = ddafft.rotate(R, dirs0)
                for p in range(2):
                    kinc.append(k)
                    einc.append(e[p].astype(np.complex64))
                    dirs.append(d)
            out.append(dict(dims=(ni, nj, nl), kd=kd, occ=occ, ndip=ndip,
                            alpha=np.array([0, ddafft.polarizability(1, kd, m)], np.complex64),
                            kinc=np.array(kinc), einc=np.array(einc), dirs=np.array(dirs)))
    return out


def batched_run(frac=1.0, width=8):
    """The batched form against the plain single-instance path on the same
    solves.  Batched: every realization's 16 right-hand sides (8 orientations
    x 2 polarizations) share its A-hat in one solve, and ddafft_solve_batch
    interleaves `width` realizations on their own streams.  Plain: the same
    solves one realization and one orientation at a time, both polarizations
    as two right-hand sides, as the plain workload runs them."""
    insts = mc_instances(frac)
    ctxs = []
    for it in insts:
        c = ddafft.Context(it["dims"], it["kd"], nrhs_max=2 * MC_ORIENT)
        c.set_target(it["occ"], [0, 1])
        ctxs.append(c)
    jobs = [dict(alpha=it["alpha"], kinc=it["kinc"], einc=it["einc"], dirs=it["dirs"]) for it in insts]
    args = dict(max_step=MAX_STEPS, err_max=TOL)
    ddafft.solve_batch(ctxs[:1], jobs[:1], "bicgstab", width=1, **args)      # warm-up (code objects)
    t0 = time.perf_counter()
    out = ddafft.solve_batch(ctxs, jobs, "bicgstab", width=width, **args)
    t_batch = time.perf_counter() - t0
    t0 = time.perf_counter()
    plain = []
    for c, it in zip(ctxs, insts):
        for o in range(MC_ORIENT):
            sl = slice(2 * o, 2 * o + 2)
            plain += c.solve_ex("bicgstab", it["alpha"], it["kinc"][sl], it["einc"][sl], dirs=it["dirs"][sl], **args)
    t_plain = time.perf_counter() - t0
    # the same batched solves with the compact Â (one octant per realization,
    # built outside the timed region)
    for c in ctxs:
        c.set_ahat_mode(1)
    t0 = time.perf_counter()
    outc = ddafft.solve_batch(ctxs, jobs, "bicgstab", width=width, **args)
    t_compact = time.perf_counter() - t0
    for c in ctxs:
        c.close()
    flatc = [r for o in outc for r in o]
    flat = [r for o in out for r in o]
    same = all(a["steps"] == b["steps"] and np.array_equal(a["esc"], b["esc"]) for a, b in zip(flat, plain))
    solves = len(flat)
    its = [r["steps"] for r in flat]
    mv_b = sum(r["matvecs"] for r in flat)
    byts = sum(float(np.prod(it["dims"])) * (2 * MC_ORIENT * 649 + 384) for it in insts)
    return {"value": round(solves / t_batch, 3), "unit": "solves/s", "solves": solves,
            "seconds": round(t_batch, 4), "matvecs_per_s": round(mv_b / t_batch, 1),
            "plain_solves_per_s": round(solves / t_plain, 3), "plain_seconds": round(t_plain, 4),
            "vs_plain": round(t_plain / t_batch, 3), "results_identical_to_plain": bool(same),
            "instances": len(insts), "rhs_per_instance": 2 * MC_ORIENT, "width": width,
            "iterations_min_mean_max": [min(its), round(float(np.mean(its)), 2), max(its)],
            "dipoles_total": int(sum(it["ndip"] for it in insts)),
            "one_sweep_algorithmic_gb": round(byts / 1e9, 3),
            "compact_ahat": {
                "value": round(solves / t_compact, 3), "unit": "solves/s", "seconds": round(t_compact, 4),
                "vs_plain": round(t_plain / t_compact, 3),
                "same_iterations_frac": round(float(np.mean([a["steps"] == b["steps"] for a, b in zip(flat, flatc)])), 4),
                "max_esc_rel_diff": float("%.3g" % max(
                    np.max(np.abs(a["esc"] - b["esc"])) / np.max(np.abs(a["esc"])) for a, b in zip(flat, flatc))),
                "max_esc_rel_diff_same_iterations": float("%.3g" % max(
                    np.max(np.abs(a["esc"] - b["esc"])) / np.max(np.abs(a["esc"]))
                    for a, b in zip(flat, flatc) if a["steps"] == b["steps"])),
            },
            "workload": "%d random agglomerates (radius-3 monomers, packing 0.3): %s lattices, x ~ U(6,12) each "
                        "(own kd and A-hat), m=1.55+0.02i, CM, %d MC orientations x 2 polarizations, 19 rotated "
                        "directions, BiCGStab resid 1e-3" % (len(insts), "/".join(
                            "%dx%s" % (max(1, int(round(c * frac))), "x".join(map(str, (a, b, d))))
                            for a, b, d, c in MC_LATTICES), MC_ORIENT)}


def large_run(cpu=False):
    """The large single target: a sphere in a 256^3 lattice (512^3 padded FFT),
    x = 30, m = 1.33+1e-4i, LDR (alpha per polarization), BiCGStab to 1e-4
    (max 2000 steps), 2 polarizations, 8 planes x 361 angles (2888 directions),
    solved to convergence."""
    grid = (256, 256, 256)
    t0 = time.perf_counter()
    occ, ndip = ddafft.target_ellipsoid(grid, 1.0, 1.0)
    kd = np.float32(2.0 * np.float32(30.0) / np.float32(grid[0]))
    dirs, _, kinc, einc = ddafft.geometry_fixed(8, 361, 0.0, 0.5)
    m = 1.33 + 1e-4j
    alpha = np.array([[0, ddafft.polarizability(5, kd, m, kinc, einc[p])] for p in range(2)], np.complex64)
    ctx = ddafft.Context(grid, kd, nrhs_max=2)
    ctx.set_target(occ, [0, 1])
    ctx.set_directions(dirs.reshape(-1, 3))
    t_setup = time.perf_counter() - t0
    res = ctx.solve(2, alpha, kinc, einc, max_step=2000, err_max=1e-4)   # warm-up (code objects, first touch)
    t0 = time.perf_counter()
    res = ctx.solve(2, alpha, kinc, einc, max_step=2000, err_max=1e-4)
    dt = time.perf_counter() - t0
    km, tot = ctx.time_matvec(2, 5)
    ahat = ctx.ahat() if cpu else None
    ctx.close()
    cpu_ref = cpu_large_baseline(occ, alpha[0], kd, kinc, einc[0], dirs.reshape(-1, 3)[:1], ahat) if cpu else None
    del ahat
    n = int(np.prod(grid))
    by = matvec_bytes(n, 2)
    return {"value": round(2 / dt, 4), "unit": "solves/s", "seconds": round(dt, 4),
            "iterations": [int(r["steps"]) for r in res], "matvecs": [int(r["matvecs"]) for r in res],
            "final_discrepancy": [float(r["err_hist"][-1]) for r in res], "dipoles": int(ndip),
            "setup_seconds_incl_ahat": round(t_setup, 3),
            "matvec_ms_2rhs": round(tot, 4), "matvec_gbps": round(sum(by.values()) / (tot * 1e-3) / 1e9, 1),
            "pass_ms": {k: round(float(v), 4) for k, v in zip(("zf", "yf", "x", "yi", "zi"), km)},
            "workload": "sphere in 256^3 (512^3 padded FFT), x=30, m=1.33+1e-4i, LDR, BiCGStab resid 1e-4 "
                        "(max 2000), 2 polarizations, 8 planes x 361 angles, solved to convergence",
            "cpu_reference": cpu_ref}


def matvec_scaling(reps=50):
    """Average duration of each matvec pass (HIP events, ddafft_time_matvec)
    for 1..16 right-hand sides sharing one operator, with the algorithmic
    GB/s of each pass."""
    print("# lattice nrhs | total_us us_per_rhs | zf yf x yi zi (us) | zf yf x yi zi (GB/s) | matvec GB/s")
    for dims, rmax in (((32, 32, 32), 16), ((64, 64, 64), 16), ((128, 128, 128), 4)):
        n = int(np.prod(dims))
        ctx = ddafft.Context(dims, 0.3, nrhs_max=rmax)
        occ, _ = ddafft.target_ellipsoid(dims, 1.0, 1.0)
        ctx.set_target(occ, [0, 1])
        r = 1
        while r <= rmax:
            km, tot = ctx.time_matvec(r, reps)
            by = matvec_bytes(n, r)
            gbs = [by[k] / (t * 1e-3) / 1e9 for k, t in zip(("zf", "yf", "x", "yi", "zi"), km)]
            print("%s %2d | %8.1f %8.1f | %s | %s | %6.0f" % (
                "x".join(map(str, dims)), r, tot * 1e3, tot * 1e3 / r, " ".join("%6.1f" % (v * 1e3) for v in km),
                " ".join("%5.0f" % g for g in gbs), sum(by.values()) / (tot * 1e-3) / 1e9), flush=True)
            r *= 2
        ctx.close()


def gradient_run(ctx, w, fwd_ms, reps=20):
    """The with-gradient form on the plain workload: forward solve of both
    polarizations, adjoint solve of d(mean Cext)/d(m, chi), assembly.  Each
    call is timed on its own; the figure is the median (about one call in 40
    takes 6 ms longer, a host-side stall seen with every build), the mean is
    reported beside it."""
    nd = w["dirs"].shape[0]
    W = np.zeros((1, 2 + nd), np.float32)
    W[0, 0] = 1.0
    m = complex(M_INDEX[0], M_INDEX[1])
    alpha = np.stack([w["alpha"], w["alpha"]])
    dadm = np.array([[0, ddafft.polarizability_deriv(1, w["kd"], m)]] * 2, np.complex64)
    # grad_chi lands in a page-locked buffer that the caller reuses across calls
    gchi = ddafft.pinned_empty((1,) + GRID + (2,), np.float32)
    args = dict(dirs=None, normals=None, max_step=MAX_STEPS, err_max=TOL, ndirs=nd, grad_chi_out=gchi)
    res = ctx.gradient("bicgstab", alpha, dadm, [1.0, m], w["kinc"], w["einc"], W, **args)
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        res = ctx.gradient("bicgstab", alpha, dadm, [1.0, m], w["kinc"], w["einc"], W, **args)
        ts.append(1e3 * (time.perf_counter() - t0))
    ms = float(np.median(ts))
    mvs = int(res["info"][:, 3].sum())
    return {"ms_per_gradient": round(ms, 4), "ms_mean": round(float(np.mean(ts)), 4), "calls": reps,
            "forward_ms": round(fwd_ms, 4), "ratio_to_forward": round(ms / fwd_ms, 3),
            "solves_per_s": round(2 / (ms * 1e-3), 3), "matvecs": mvs,
            "objective": "d(mean Cext)/d(Re m, Im m) and d/d(chi) per dipole, both polarizations"}


def compact_run(ctx, step, res0, steps, n):
    """The plain workload with the compact Â (ddafft_set_ahat_mode 1: one
    octant of the transform of A with its Nyquist planes zeroed): solves/s,
    the X pass and matvec from profiled steps, and how far the observables
    move from the reference-Â solve (float rounding only)."""
    ctx.set_ahat_mode(1)
    for _ in range(3):
        res = step()
    t0 = time.perf_counter()
    for _ in range(steps):
        res = step()
    dt = time.perf_counter() - t0
    ctx.profile(True)
    for _ in range(5):
        step()
    pass_ms, nmv = ctx.profile(False)
    ctx.set_ahat_mode(0)
    avg_ms = pass_ms / max(nmv, 1)
    rel = lambda a, b: float("%.3g" % (np.max(np.abs(np.asarray(a) - np.asarray(b))) / np.max(np.abs(np.asarray(b)))))
    return {
        "value": round(2 * steps / dt, 3), "unit": "solves/s",
        "ahat_mb": round(6 * 8 * (GRID[0] + 1) * (GRID[1] + 1) * (GRID[2] + 1) / 1e6, 2),
        "x_ms": round(float(avg_ms[2]), 5), "matvec_ms": round(float(np.sum(avg_ms)), 5),
        "iterations": [int(r["steps"]) for r in res],
        "rel_diff_vs_reference_ahat": {
            "cext": max(rel(r["cext"], r0["cext"]) for r, r0 in zip(res, res0)),
            "cabs": max(rel(r["cabs"], r0["cabs"]) for r, r0 in zip(res, res0)),
            "esc": max(rel(r["esc"], r0["esc"]) for r, r0 in zip(res, res0)),
        },
    }


def host_cores():
    """Cores this process may run on (the GPU's share of the host)."""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_large_baseline(occ, alpha, kd, kinc, einc, dirs, ahat):
    """The reference's own DDA_FFT (oracle/_ref) on the 256^3 large target,
    one core (it runs one sample), A-hat taken from the GPU (bit-identical to
    the reference's): BiCGStab step 0 and two iterations (7 matvecs), far
    field in one direction.  s/iteration = 2/7 of the run (an iteration is two
    of its seven matvecs; the reductions and updates are a few percent)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    try:
        ref = oracle.Reference()
    except (FileNotFoundError, OSError):
        return None
    print("bench: reference DDA_FFT on the 256^3 target, 2 steps (one core)...", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    ref.dda_fft(2, occ, np.array([0, 1], np.int32), np.asarray(alpha, np.complex64), kd, kinc, einc, ahat, dirs,
                max_step=2, err_max=1e-12, log_path=os.devnull)
    dt = time.perf_counter() - t0
    return {"seconds": round(dt, 2), "matvecs": 7, "s_per_iteration": round(2 * dt / 7, 2), "cores": 1,
            "kind": "reference", "sample": "BiCGStab step 0 + 2 iterations of one polarization (DDA_FFT, -S 2), "
                                            "A-hat prebuilt, 1 far-field direction"}


def cpu_batched_baseline(frac=0.125):
    """The reference program itself (oracle/_ref/dda_fft.out) on a labelled
    subset of the batched Monte-Carlo workload: each realization a target-11
    shape file run with 8 Monte-Carlo orientations (-n 8 -O mc, 16 solves),
    one process per realization and as many processes at once as this process
    has cores (at most 16, the GPU's share of the host).  Wall time from the
    first start to the last exit."""
    import subprocess
    import tempfile

    exe = os.path.join(ROOT, "oracle", "_ref", "dda_fft.out")
    if not os.path.exists(exe):
        return None
    lib = ddafft.load()
    workers = max(1, min(16, host_cores()))
    insts = mc_instances(frac)
    tmp = tempfile.mkdtemp(prefix="ddafft_cpu_")
    cmds = []
    for k, it in enumerate(insts):
        d = os.path.join(tmp, "r%d" % k)
        os.makedirs(d)
        ni, nj, nl = it["dims"]
        lib.ddafft_target_file_write.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_void_p]
        lib.ddafft_target_file_write(os.path.join(d, "shape.txt").encode(), ni, nj, nl, _p(it["occ"]))
        x = float(it["kd"]) * ni / 2.0
        cmds.append(([exe, "-t", "file", "-f", "shape.txt", "-g", str(ni), str(nj), str(nl), "-x", "%.9g" % x,
                      "-m", str(MC_M[0]), str(MC_M[1]), "-P", "cm", "-M", "bicgstab", "-e", str(TOL), "-S",
                      str(MAX_STEPS), "-O", "mc", "-n", str(MC_ORIENT), "-p", "1", "-a", "0", "-b", "180", "-s",
                      "10", "-j", "1", "--seed", str(MC_SEED + k), "-o", "out.dat", "-D", "out.dst"], d))
    print("bench: reference program on %d realizations, %d at once..." % (len(cmds), workers), file=sys.stderr,
          flush=True)
    running, pending, failed = [], list(cmds), 0
    t0 = time.perf_counter()
    while pending or running:
        while pending and len(running) < workers:
            cmd, d = pending.pop(0)
            running.append(subprocess.Popen(cmd, cwd=d, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL))
        time.sleep(0.02)
        for pr in list(running):
            if pr.poll() is not None:
                failed += pr.returncode != 0
                running.remove(pr)
    dt = time.perf_counter() - t0
    solves = 2 * MC_ORIENT * len(cmds)
    return {"value": round(solves / dt, 3), "unit": "solves/s", "seconds": round(dt, 2), "solves": solves,
            "cores": workers, "kind": "reference", "failed_runs": failed,
            "sample": "%d of the 128 realizations (%s), each the reference program with -t file -n 8 -O mc, "
                      "%d processes at once" % (len(cmds), "/".join(
                          "%dx%s" % (max(1, int(round(c * frac))), "x".join(map(str, (a, b, e))))
                          for a, b, e, c in MC_LATTICES), workers)}


def cpu_baseline(w, budget_s=10.0):
    """The reference's own DDA_FFT (oracle/_ref, compiled from its sources) on
    this host for the same sample; falls back to the oracle restatement."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    kind = "reference"
    try:
        impl = oracle.Reference()
    except (FileNotFoundError, OSError):
        impl = oracle.Oracle("f32")
        kind = "port"
    ahat = impl.build_ahat(GRID, w["kd"])
    mv = np.array([0, 1], np.int32)
    solves, t_total, reps = 0, 0.0, 0
    while reps < 3 and (t_total < budget_s or reps == 0):
        for pol in range(2):
            kw = {"log_path": os.devnull} if kind == "reference" else {}
            t0 = time.perf_counter()
            impl.dda_fft(2, w["occ"], mv, w["alpha"], w["kd"], w["kinc"], w["einc"][pol], ahat, w["dirs"],
                         max_step=MAX_STEPS, err_max=TOL, **kw)
            t_total += time.perf_counter() - t0
            solves += 1
        reps += 1
    return {
        "value": round(solves / t_total, 4),
        "unit": "solves/s",
        "cores": 1,
        "host_cores": host_cores(),
        "kind": kind,
        "sample": "%d BiCGStab solve(s) of the same 64^3 workload (DDA_FFT incl. far field, A-hat prebuilt), "
                  "single worker as the reference runs one sample" % solves,
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched multi-instance form")
    ap.add_argument("--batched-frac", type=float, default=1.0,
                    help="fraction of each lattice group of the 128-realization batched workload")
    ap.add_argument("--batched-width", type=int, default=0,
                    help="realizations in flight in the batch driver (0: the autotuner's choice)")
    ap.add_argument("--no-gradient", action="store_true", help="skip the forward+adjoint form")
    ap.add_argument("--no-large", action="store_true", help="skip the 256^3 single-target solve")
    ap.add_argument("--matvec-scaling", action="store_true",
                    help="print the per-pass matvec times against the number of right-hand sides and exit")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="first run one plain device copy of this many MiB (torch), the reference kernel "
                         "against which rocprofv3 FETCH_SIZE / WRITE_SIZE counters are scaled")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this DDA-FFT path runs on one GPU; a multi-GPU run is not part of it "
              "(independent solves would shard by instance, one lattice by slabs along i).", file=sys.stderr)
        return 2

    if args.calib_copy_mib > 0:
        import torch

        src = torch.ones(args.calib_copy_mib * 2**20 // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    lib = ddafft.load()
    if args.matvec_scaling:
        matvec_scaling()
        return 0
    w = build_workload(lib)
    ctx = ddafft.Context(GRID, w["kd"], nrhs_max=2)
    ctx.set_target(w["occ"], [0, 1])
    ctx.set_directions(w["dirs"])
    alpha2 = np.stack([w["alpha"], w["alpha"]])

    def step():
        return ctx.solve(2, alpha2, w["kinc"], w["einc"], max_step=MAX_STEPS, err_max=TOL)

    for _ in range(args.warmup):
        res = step()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        res = step()
    t1 = time.perf_counter()
    dt = t1 - t0
    # per-pass kernel durations: HIP events around every pass on the context's
    # stream, recorded in separate steps so that they do not touch the timed ones
    ctx.profile(True)
    for _ in range(max(5, args.steps // 5)):
        step()
    pass_ms, nmv = ctx.profile(False)

    solves = 2 * args.steps
    matvecs = sum(r["matvecs"] for r in res) // 2 * args.steps     # batched matvecs (2 RHS each)
    n = int(np.prod(GRID))
    by = matvec_bytes(n, 2)
    avg_ms = pass_ms / max(nmv, 1)
    x_gbs = by["x"] / (avg_ms[2] * 1e-3) / 1e9
    mv_ms = float(np.sum(avg_ms))
    mv_gbs = sum(by.values()) / (mv_ms * 1e-3) / 1e9

    out = {
        "metric": "BiCGStab solves/s & FFT-matvec GB/s (%HBM peak) at 64³, resid 1e-3",
        "value": round(solves / dt, 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * dt / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "64^3 oblate spheroid (b/a=1, c/a=0.5), x=10, m=1.6+0.005i, CM, BiCGStab resid 1e-3 "
                        "(max 300), 2 polarizations batched as 2 RHS, 4 planes x 181 angles",
            "lattice": list(GRID),
            "padded_fft": [2 * g for g in GRID],
            "dipoles": int(w["ndip"]),
            "rhs_per_step": 2,
            "iterations": [int(r["steps"]) for r in res],
            "discrepancy": [[round(float(v), 7) for v in r["err_hist"]] for r in res],
        },
        "matvecs_per_s": round(matvecs * 2 / dt, 1),
        "matvec_ms": round(mv_ms, 5),
        "matvec_gbps": round(mv_gbs, 1),
        "matvec_hbm_frac": round(mv_gbs / HBM_PEAK_GBS, 4),
        "pass_ms": {k: round(float(v), 5) for k, v in zip(("zf", "yf", "x", "yi", "zi"), avg_ms)},
        "roofline": {
            "bound": "hbm",
            "kernel": "k_x (fused x-forward FFT, A-hat product, x-inverse FFT), 2 RHS",
            "achieved": round(x_gbs, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(x_gbs / HBM_PEAK_GBS, 4),
            "traffic": traffic_from_profiles(),
        },
    }
    if not args.no_gradient:
        out["gradient"] = gradient_run(ctx, w, 1e3 * dt / args.steps)
    out["compact_ahat"] = compact_run(ctx, step, res, max(10, args.steps // 2), n)
    ctx.close()
    if not args.no_batched:
        out["batched"] = batched_run(args.batched_frac, args.batched_width)
    if not args.no_large:
        out["large"] = large_run(cpu=not args.no_cpu_baseline)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(w)
        out["cpu_baseline"]["batched"] = cpu_batched_baseline()
        if "large" in out:
            out["cpu_baseline"]["large"] = out["large"].pop("cpu_reference")
    print(json.dumps(out), flush=True)
    return 0


def traffic_from_profiles():
    """HBM bytes per launch of the dominant kernel from the committed rocprof
    counter summary (FETCH_SIZE + WRITE_SIZE, copy-kernel calibrated)."""
    path = os.path.join(ROOT, "profiles", "k_x_spheroid64_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("bytes_per_launch")


if __name__ == "__main__":
    sys.exit(main())
