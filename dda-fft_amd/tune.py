"""Autotuner of the launch schedule of libddafft_amd.

Measures, on the GPU it runs on, for each (lattice, right-hand sides) case:

- the X pass (the fused x-forward FFT, A-hat product and x-inverse FFT) under
  every grouping of the right-hand sides: 1, 2 or 4 per tile sharing each
  A-hat read, and runs of groups per block (one group per block, or runs
  long enough to cut the launch to 8192 / 2048 blocks);
- the z- and y-forward transforms as one kernel per i-plane against two
  passes (planes up to 64 x 64);
- a whole BiCGStab solve with its iterations replayed from captured HIP
  graphs against launched kernel by kernel;
- for the batched Monte-Carlo workload, the number of realizations in
  flight in ddafft_solve_batch.

and writes the winners as the table the library compiles in
(csrc/tuned_table.hpp; rebuild with `make -C dda-fft_amd` afterwards) plus a
text record of every measurement.  The results of every schedule are the
same bits; only the time differs.

    python dda-fft_amd/tune.py [--out-dir DIR] [--quick]
"""

import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import ddafft  # noqa: E402

# (lattice, right-hand sides): the plain solve (2 polarizations) on the
# benchmark lattices, the batched workload's 16 right-hand sides per
# realization, and the long-axis sizes
CASES = [((32, 32, 32), 2), ((64, 64, 64), 2), ((128, 128, 128), 2), ((256, 256, 256), 2),
         ((32, 32, 32), 16), ((32, 32, 64), 16), ((64, 64, 64), 16), ((32, 32, 64), 2), ((64, 64, 64), 4),
         ((128, 128, 128), 4)]
X_BLOCK_RUNS = (1 << 30, 8192, 2048)


def x_candidates(dims, nrhs):
    out = []
    for rb in (1, 2, 4):
        if rb > nrhs or (dims[0] > 256 and rb > 1):
            continue
        groups = (nrhs + rb - 1) // rb
        for xb in X_BLOCK_RUNS if groups > 1 else X_BLOCK_RUNS[:1]:
            out.append((rb, xb))
    return out


def time_x(ctx, nrhs, reps):
    km, _ = ctx.time_matvec(nrhs, reps)
    return float(km[2])


def time_zy(ctx, nrhs, reps):
    km, _ = ctx.time_matvec(nrhs, reps)
    return float(km[0] + km[1])


def target(dims, x=5.0, m=1.5 + 0.01j):
    kd = np.float32(2.0 * x / dims[0])
    occ, _ = ddafft.target_ellipsoid(dims, min(1.0, dims[1] / dims[0]), min(1.0, dims[2] / dims[0]))
    alpha = np.array([0, ddafft.polarizability(1, kd, m)], np.complex64)
    return kd, occ, alpha


def tune_case(dims, nrhs, reps, log):
    kd, occ, alpha = target(dims)
    ctx = ddafft.Context(dims, kd, nrhs_max=nrhs)
    ctx.set_target(occ, [0, 1])
    best = None
    for rb, xb in x_candidates(dims, nrhs):
        ctx.set_tuning(x_rb=rb, x_blocks=xb)
        time_x(ctx, nrhs, 2)                            # code objects, first touch
        t = min(time_x(ctx, nrhs, reps) for _ in range(3))
        log("%-14s R=%-2d X pass x_rb=%d x_blocks=%-10d %9.1f us" % ("x".join(map(str, dims)), nrhs, rb, xb,
                                                                    t * 1e3))
        if best is None or t < best[0]:
            best = (t, rb, xb)
    _, rb, xb = best
    ctx.set_tuning(x_rb=rb, x_blocks=xb)
    fuse = 0
    if dims[1] <= 64 and dims[2] <= 64:
        tz = {}
        for f in (0, 1, 0, 1):
            ctx.set_tuning(fuse_zy=f)
            time_zy(ctx, nrhs, 2)
            tz[f] = min(tz.get(f, 1e30), min(time_zy(ctx, nrhs, reps) for _ in range(2)))
        fuse = 1 if tz[1] < tz[0] else 0
        log("%-14s R=%-2d z/y forward: two passes %.1f us, fused %.1f us -> fuse=%d" % (
            "x".join(map(str, dims)), nrhs, tz[0] * 1e3, tz[1] * 1e3, fuse))
    ctx.set_tuning(fuse_zy=fuse)
    # graphs against launches on a whole solve of nrhs right-hand sides
    kinc = np.array([0, 0, 1], np.float32)
    einc = np.array([np.eye(3)[r % 2] for r in range(nrhs)], np.complex64)
    th = np.arange(0, 181, 10.0)
    dirs = np.stack([np.sin(np.radians(th)), 0 * th, np.cos(np.radians(th))], 1).astype(np.float32)
    steps = 8 if np.prod(dims) <= 64 ** 3 else 3
    tg = {}
    for g in (0, 1, 0, 1):
        ctx.set_tuning(graph=g)
        ctx.solve("bicgstab", alpha, kinc, einc, dirs=dirs, max_step=steps, err_max=1e-12)
        n = max(3, reps // 4)
        t0 = time.perf_counter()
        for _ in range(n):
            ctx.solve("bicgstab", alpha, kinc, einc, dirs=dirs, max_step=steps, err_max=1e-12)
        dt = (time.perf_counter() - t0) / n
        tg[g] = min(tg.get(g, 1e30), dt)
    graph = 1 if tg[1] <= tg[0] else 0
    log("%-14s R=%-2d solve (%d BiCGStab steps): launched %.3f ms, graphs %.3f ms -> graph=%d" % (
        "x".join(map(str, dims)), nrhs, steps, tg[0] * 1e3, tg[1] * 1e3, graph))
    ctx.close()
    return dict(dims=dims, nrhs=nrhs, x_rb=rb, x_blocks=xb, graph=graph, fuse=fuse, x_us=best[0] * 1e3,
                solve_ms=min(tg.values()) * 1e3)


def tune_width(log, frac=0.25):
    """Realizations in flight for the batched Monte-Carlo workload (bench.py's
    mc_instances, a fraction of each lattice group)."""
    sys.path.insert(0, os.path.dirname(HERE))
    import bench
    insts = bench.mc_instances(frac)
    ctxs = []
    for it in insts:
        c = ddafft.Context(it["dims"], it["kd"], nrhs_max=16)
        c.set_target(it["occ"], [0, 1])
        ctxs.append(c)
    jobs = [dict(alpha=it["alpha"], kinc=it["kinc"], einc=it["einc"], dirs=it["dirs"]) for it in insts]
    ddafft.solve_batch(ctxs[:2], jobs[:2], "bicgstab", max_step=bench.MAX_STEPS, err_max=bench.TOL, width=2)
    best = None
    for w in (2, 4, 8, 12, 16):
        t0 = time.perf_counter()
        ddafft.solve_batch(ctxs, jobs, "bicgstab", max_step=bench.MAX_STEPS, err_max=bench.TOL, width=w)
        dt = time.perf_counter() - t0
        log("batched workload (%d realizations): width %2d  %.4f s" % (len(insts), w, dt))
        if best is None or dt < best[0]:
            best = (dt, w)
    for c in ctxs:
        c.close()
    return best[1]


def write_table(entries, width, path):
    lines = ["// Launch schedule per (lattice, right-hand sides): written by",
             "// dda-fft_amd/tune.py from measurements on an MI355X.  x_rb: right-hand",
             "// sides per X-pass tile; x_blocks: X-pass launch size up to which groups of",
             "// right-hand sides are spread over blocks (1 << 30: each group its own",
             "// block); graph: iterations replayed as captured HIP graphs; fuse: z- and",
             "// y-forward transforms of an i-plane in one kernel.",
             "#pragma once", "", "namespace ddafft {", "",
             "struct TunedEntry {", "    int ni, nj, nl, nrhs, x_rb, x_blocks, graph, fuse;", "};", "",
             "// realizations in flight in ddafft_solve_batch for the batched Monte-Carlo",
             "// workload (the call's width argument 0 stands for it)",
             "constexpr int kTunedBatchWidth = %d;" % width, "",
             "constexpr TunedEntry kTuned[] = {"]
    for e in entries:
        xb = "1 << 30" if e["x_blocks"] == 1 << 30 else str(e["x_blocks"])
        lines.append("    {%d, %d, %d, %d, %d, %s, %d, %d},   // X pass %.1f us, solve %.3f ms" % (
            e["dims"][0], e["dims"][1], e["dims"][2], e["nrhs"], e["x_rb"], xb, e["graph"], e["fuse"],
            e["x_us"], e["solve_ms"]))
    lines += ["};", "", "}  // namespace ddafft", ""]
    with open(path, "w") as f:
        f.write("\n".join(lines))


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out-dir", default=os.path.join(HERE, "csrc"),
                    help="where tuned_table.hpp and tuning_record.txt go")
    ap.add_argument("--quick", action="store_true", help="fewer repetitions, no 256^3 case")
    args = ap.parse_args()
    os.makedirs(args.out_dir, exist_ok=True)
    rec = open(os.path.join(args.out_dir, "tuning_record.txt"), "w")

    def log(msg):
        print(msg, flush=True)
        rec.write(msg + "\n")
        rec.flush()

    ddafft.load()
    reps = 5 if args.quick else 20
    entries = []
    for dims, nrhs in CASES:
        if args.quick and np.prod(dims) > 128 ** 3:
            continue
        entries.append(tune_case(dims, nrhs, reps, log))
    width = tune_width(log)
    write_table(entries, width, os.path.join(args.out_dir, "tuned_table.hpp"))
    log("wrote %s (batch width %d)" % (os.path.join(args.out_dir, "tuned_table.hpp"), width))


if __name__ == "__main__":
    main()
