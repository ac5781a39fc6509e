#!/usr/bin/env python3
"""Benchmark: pressure projections per second to a 1e-6 relative true
residual on the 256^3 two-phase FLIP dam break (liquid:air 1000:1), one GPU.

A step is one cold-started projection of the frame's system on the GPU:
updateDirichletFields + build system, coarse-level refresh, MGPCG until
||b - A x||_2 <= 1e-6 ||b||_2 (FP64 true residual), circuit breaker and
velocity correction (createGridPressureSolver3.project, grid_pressure_solver3.js:339-558).
The pressure is reset to zero before each step so every step solves the same
system from the same start; inputs are resident in HBM before timing starts.

Prints one JSON line (rank 0).  `--gpus N` with N > 1 is refused: this build
runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--levels", type=int, default=6)
    ap.add_argument("--kernel", default="relax", help="kernel family timed for the roofline")
    ap.add_argument("--storage", default="f32", choices=["f32", "f16", "bf16"],
                    help="V-cycle storage: f32 is the reference's plain method, f16/bf16 the mixed form")
    ap.add_argument("--smoother", default="rbgs", choices=["rbgs", "jacobi"],
                    help="V-cycle smoother: rbgs is the reference's plain method, jacobi its damped-Jacobi form")
    ap.add_argument("--nu", type=int, default=2, help="smoothing iterations down / up / final")
    ap.add_argument("--jacobi-omega", type=float, default=2.0 / 3.0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-iters", type=int, default=2, help="PCG iterations in the CPU sample")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run a plain 512 MiB device-to-device copy three times (the known-size kernel that scales "
                         "FETCH_SIZE / WRITE_SIZE counter readings to bytes)")
    return ap.parse_args()


# algorithmic bytes per finest-level cell and launch, per kernel family
# (FP32 fields, u8 mask, three FP32 face-weight arrays ~ 12 B/cell)
ALGO_BYTES_PER_CELL = {
    # one colour of red-black GS: x read (4) + b on the colour (2) + mask (1)
    # + face weights (12) + x written on the colour (2)
    "relax": 21.0,
    # one fused red-black sweep (both colours): x in (4) + b (4) + mask (1)
    # + face weights (12) + x out (4)
    "sweep": 25.0,
    # p = z + beta p (z 4 + p 4 read, p 4 written), Ap written (4), mask 1, weights 12
    "apply": 29.0,
    # x64 read+write (16), p 4, b 4, r written 4, mask 1, weights 12
    "update": 41.0,
    # x 4 + b 4 + mask 1 + weights 12 read, r 4 written
    "residual": 25.0,
    # one damped-Jacobi sweep: x read 4, b 4, mask 1, weights 12, x' written 4
    "jacobi": 25.0,
    # coarse x (1/8 of 4) + mask 1 read, fine x read+write 8
    "correct": 9.5,
}


# fine-level kernel of each family, as rocprofv3 names it
KERNEL_SYMBOL = {"relax": "k_relax3<1, 0,", "sweep": "k_rbgs_fused<1,", "jacobi": "k_jacobi3<1,", "apply": "k_pcg_apply3",
                 "update": "k_pcg_update3", "residual": "k_residual3<1,", "correct": "k_correct<1,"}


def measured_traffic(kernel, storage):
    """HBM bytes per dispatch of the family's fine-level kernel from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary (copy-calibrated), or
    None when no summary covers it."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic_dambreak256_%s.json" % storage)
    sym = KERNEL_SYMBOL.get(kernel)
    if not sym or not os.path.exists(path):
        return None
    best = None
    for name, grids in json.load(open(path))["kernels"].items():
        if sym not in name:
            continue
        for grid, rec in grids.items():
            if best is None or int(grid) > best[0]:
                best = (int(grid), rec["hbm_bytes"])
    return None if best is None else best[1]


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(sc, levels, gpu_iters, iters, threads):
    """The oracle's restatement of the same projection (f32 vectors, f64
    solution accumulator: the HIP path's arithmetic) on the host cores:
    setup plus `iters` PCG iterations are timed and scaled to a full
    projection at the GPU's iteration count; run on `threads` OpenMP
    threads and on one."""
    one = cpu_sample(sc, levels, gpu_iters, 1, 1)  # single-threaded: a shorter sample
    res = cpu_sample(sc, levels, gpu_iters, iters, threads)
    res["single_thread_value"] = one["value"]
    res["cpu_model"] = cpu_model()
    res["nproc"] = os.cpu_count()
    return res


def cpu_sample(sc, levels, gpu_iters, iters, threads):
    from oracle import pyoracle as O
    O.set_threads(threads)
    n, h = sc["shape"], sc["spacing"]
    t0 = time.perf_counter()
    A = O.pressure3_assemble(n, h, sc["vel"], active=sc["active"], target=sc["target"], beta=sc["beta"], suffix="f32")
    mg = O.Multigrid(n, h, levels=levels, mask=A["mask"], coarse_mask=A["vent"], weights=A["weights"], suffix="f32")
    t1 = time.perf_counter()
    O.pcg_solve_mixed(n, h, A["b"], x0=A["pressure"], mask=A["mask"], weights=A["weights"], mg=mg, tol=1e-6,
                      maxiter=1)
    t2 = time.perf_counter()
    O.pcg_solve_mixed(n, h, A["b"], x0=A["pressure"], mask=A["mask"], weights=A["weights"], mg=mg, tol=1e-6,
                      maxiter=1 + iters)
    t3 = time.perf_counter()
    per_iter = (t3 - t2) / iters
    setup = (t1 - t0) + (t2 - t1) - per_iter
    proj_s = setup + gpu_iters * per_iter
    return {
        "value": 1.0 / proj_s,
        "unit": "projections/s",
        "cores": threads,
        "kind": "port",
        "sample": "%d^3 system: assembly + V-cycle setup + %d PCG iterations timed (%.1f s), scaled to %d iterations"
                  % (n[0], 1 + iters, t3 - t0, gpu_iters),
        "s_per_iteration": per_iter,
    }


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only; --gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)
    import tsl_array_n_amd as F
    from tsl_array_n_amd import scenes

    F.init(int(os.environ.get("LOCAL_RANK", "0")))
    if args.copy_calibration:
        # device-to-device copies of a 512 MiB field (the runtime's blit kernel)
        src = F.array_n("float", (128 << 20,))
        src.fill(1.0)
        dst = F.array_n("float", (128 << 20,))
        for _ in range(3):
            dst.copy_from(src)
        F.sync()
        print("bench: copy calibration: 3 x %d bytes read + written" % ((128 << 20) * 4), file=sys.stderr, flush=True)
        del src, dst
    dev, cus = F.device_info()
    n = args.n
    print("bench: building the %d^3 dam-break system on %s" % (n, dev), file=sys.stderr, flush=True)
    sc = scenes.dam_break_system(n)
    shape = sc["shape"]
    vin = [F.array_n("float", F.face_shape(shape, a), sc["vel"][a]) for a in range(3)]
    vout = [F.array_n("float", F.face_shape(shape, a)) for a in range(3)]
    beta = [F.array_n("float", F.face_shape(shape, a), sc["beta"][a]) for a in range(3)]
    act = F.array_n("u8", shape, sc["active"])
    tgt = F.array_n("float", shape, sc["target"])
    ps = F.create_grid_pressure_solver3(resolution=shape, grid_spacing=sc["spacing"], dirichlet=(act, tgt),
                                        face_weights=beta,
                                        multigrid={"levels": args.levels, "storage": args.storage,
                                                   "smoother": args.smoother, "nu_down": args.nu, "nu_up": args.nu,
                                                   "n_final": args.nu, "jacobi_omega": args.jacobi_omega},
                                        tolerance=1e-6,
                                        relative_tolerance=True)

    def step():
        ps.pressure.fill(0.0)
        return ps.project(vin, vout)

    for w in range(args.warmup):
        d = step()
        print("bench: warmup %d: %s" % (w, d), file=sys.stderr, flush=True)

    F.kprof_enable(args.kernel)
    F.sync()
    t0 = time.perf_counter()
    iters = []
    for s in range(args.steps):
        d = step()
        iters.append(d["iterations"])
        if not d["converged"]:
            print("bench: step %d did not converge: %s" % (s, d), file=sys.stderr, flush=True)
    F.sync()
    t1 = time.perf_counter()
    k_ms, k_launches = F.kprof_read()
    F.kprof_enable("")
    elapsed = t1 - t0
    print("bench: %d steps in %.3f s, iterations %s, last %s" % (args.steps, elapsed, iters, d), file=sys.stderr,
          flush=True)

    cells = float(np.prod(shape))
    roof = None
    if k_launches:
        avg_s = k_ms / 1e3 / k_launches
        algo = ALGO_BYTES_PER_CELL.get(args.kernel, 0.0) * cells
        achieved = algo / avg_s / 1e9
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(args.kernel, args.storage),
                "kernel": args.kernel, "avg_us": round(avg_s * 1e6, 2), "launches": k_launches,
                "algorithmic_bytes": algo}

    cpu = None
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        print("bench: CPU baseline on %d threads" % threads, file=sys.stderr, flush=True)
        cpu = cpu_baseline(sc, args.levels, int(np.median(iters)), args.cpu_iters, threads)

    line = {
        "metric": "projections/sec to 1e-6 rel residual at 256^3",
        "value": args.steps / elapsed,
        "unit": "projections/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (f64 accumulator, residual and dots)",
        "data": "synthetic: seeded 3-D two-phase dam break, frame-0 system",
        "config": {"workload": "3D FLIP two-phase dam break %d^3, liquid:air 1000:1, pinned top-centre, "
                               "MGPCG %d levels 2/2 RBGS, V-cycle storage %s, cold start" % (n, args.levels, args.storage),
                   "cells": int(cells), "particles": sc["particles"], "iterations": int(np.median(iters)),
                   "parallelism": "single GPU", "device": dev},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
