#!/usr/bin/env python3
"""Benchmark: pressure projections per second to a 1e-6 relative true
residual on the 256^3 two-phase FLIP dam break (liquid:air 1000:1), one GPU.

A step is one cold-started projection of the frame's system on the GPU:
updateDirichletFields + build system, coarse-level refresh, MGPCG until
||b - A x||_2 <= 1e-6 ||b||_2 (FP64 true residual), circuit breaker and
velocity correction (createGridPressureSolver3.project, grid_pressure_solver3.js:339-558).
The pressure is reset to zero before each step so every step solves the same
system from the same start; inputs are resident in HBM before timing starts.

Every form of the solver is run on the same system, each timed on its own:
  plain_rbgs          FP32 V-cycle, red-black Gauss-Seidel 2/2 (multigrid.js
                      defaults): the reference's method
  mixed_f16           coarse levels and every level's face weights stored in
                      FP16 (finest-level x, b FP32), FP32 arithmetic, FP32
                      outer CG, FP64 dots and true residual
  mixed_bf16          the same in BF16
  mixed_f16_cs        mixed_f16 plus the additive coarse-space correction of
                      the preconditioner along the free-cell indicator (the
                      near-null mode of the system pinned at one cell)
  mixed_f16_cs_exact  mixed_f16_cs with the exact solution as a second coarse
                      vector: the bound of what a learned coarse vector could
                      add
  mixed_f16_brick_cs  (headline) mixed_f16_cs with the brick smoother: every
                      smoothing is two passes over memory, each running 12
                      red-black sweeps brick by brick (64x8x8 cells, the
                      second pass on bricks shifted by half a brick) in
                      registers and LDS; post-smoothing the adjoint
  mixed_f16_brick_cs_nu6, mixed_f16_brick (no correction), plain_brick (FP32
                      storage), mixed_f16_brick_cs_exact (+ the exact
                      solution as a coarse vector)
  plain_jacobi        (on request) FP32 V-cycle, damped Jacobi 4/4, omega 6/7
                      (sandbox/jacobi-smoother/multigrid_jacobi.js)
The headline form (`--headline`, default mixed_f16_brick_cs) gives `value`,
`ms_per_step` and the roofline; `forms` carries every form's time,
iterations and FP64 relative residual, and `speedup_vs_plain` is each form's
time against plain_rbgs.  The CPU leg times the oracle's restatement of the
plain form on a bounded sample (extrapolated, and marked so) and solves the
same system in FP64 to report each form's error against it.

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

FORMS = {
    # the reference's method with its default 2/2 red-black sweeps
    "plain_rbgs": {"storage": "f32", "smoother": "rbgs", "nu": 2, "omega": 1.0},
    # the same method with 4/4 sweeps (separates the smoothing count's share)
    "plain_rbgs_nu4": {"storage": "f32", "smoother": "rbgs", "nu": 4, "omega": 1.0},
    "mixed_f16": {"storage": "f16", "smoother": "rbgs", "nu": 2, "omega": 1.0},
    "mixed_bf16": {"storage": "bf16", "smoother": "rbgs", "nu": 2, "omega": 1.0},
    # FP16 V-cycle + the additive coarse-space correction along the
    # free-cell indicator, with 4/4 sweeps (measured at 256^3 on the FLIP
    # system: 2/2 54 iterations 74.9 ms, 3/3 37 / 64.3, 4/4 29 / 60.5, 5/5
    # 24 / 58.6, 6/6 21 / 58.6)
    "mixed_f16_cs": {"storage": "f16", "smoother": "rbgs", "nu": 4, "omega": 1.0, "cs": "free"},
    "mixed_f16_cs_nu2": {"storage": "f16", "smoother": "rbgs", "nu": 2, "omega": 1.0, "cs": "free"},
    # the bound of any learned coarse vector: the exact solution as the
    # second coarse vector (taken from a plain solve before timing)
    "mixed_f16_cs_exact": {"storage": "f16", "smoother": "rbgs", "nu": 4, "omega": 1.0, "cs": "free+exact"},
    "plain_jacobi": {"storage": "f32", "smoother": "jacobi", "nu": 4, "omega": 6.0 / 7.0},
    # brick-wise red-black: every smoothing is two passes over memory (bricks
    # of 64x8x8 cells, then bricks shifted by half a brick), each running all
    # of its sweeps on the brick in registers and LDS.  Sweeps per pass chosen
    # by measurement at 256^3 on the FLIP system with the free-cell
    # correction (iterations / ms per projection: 1 -> 58 / 67.3, 2 -> 32 /
    # 39.0, 4 -> 20 / 26.4, 6 -> 16 / 22.6, 8 -> 15 / 22.5, 10 -> 14 / 22.2,
    # 12 -> 13 / 21.7); fewer sweeps on the coarse levels than on the finest
    # cost iterations (12 fine / 4 coarse: 17 iterations, 26.0 ms)
    "mixed_f16_brick_cs": {"storage": "f16", "smoother": "brick", "nu": 12, "omega": 1.0, "cs": "free"},
    "mixed_f16_brick_cs_nu4": {"storage": "f16", "smoother": "brick", "nu": 4, "omega": 1.0, "cs": "free"},
    "mixed_f16_brick_cs_nu6": {"storage": "f16", "smoother": "brick", "nu": 6, "omega": 1.0, "cs": "free"},
    # without the free-cell correction (the system is a closed box pinned at
    # one cell, nearly singular), 6 sweeps stop on the reference's
    # residual-growth guard near 6e-4 after 15 iterations; 4 sweeps converge
    "mixed_f16_brick": {"storage": "f16", "smoother": "brick", "nu": 4, "omega": 1.0},
    "plain_brick": {"storage": "f32", "smoother": "brick", "nu": 4, "omega": 1.0},
    # the bound of any learned coarse vector with the brick smoother
    "mixed_f16_brick_cs_exact": {"storage": "f16", "smoother": "brick", "nu": 12, "omega": 1.0, "cs": "free+exact"},
}
DEFAULT_FORMS = ("plain_rbgs,plain_rbgs_nu4,mixed_f16,mixed_bf16,mixed_f16_cs,mixed_f16_cs_exact,mixed_f16_brick_cs_nu6,"
                 "mixed_f16_brick,plain_brick,mixed_f16_brick_cs_exact")
# the single custom form of --storage/--smoother/--nu/--coarse-space sweeps


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50, help="timed projections of the headline form (about a second)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--form-steps", type=int, default=5, help="timed projections of each other form")
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--system", default="flip", choices=["flip", "frame0"],
                    help="flip: the pressure system the two-phase FLIP dam break solves at step --flip-steps "
                         "(captured from the FLIP solver); frame0: the synthetic frame-0 system (gravity on "
                         "particles at rest)")
    ap.add_argument("--flip-steps", type=int, default=3)
    ap.add_argument("--levels", type=int, default=6)
    ap.add_argument("--headline", default="mixed_f16_brick_cs", choices=sorted(FORMS))
    ap.add_argument("--forms", default=DEFAULT_FORMS,
                    help="comma-separated forms to run (the headline always runs); plain_jacobi is not in the default "
                         "set: at 256^3 its V-cycle is not a usable preconditioner for this system (the residual-growth "
                         "guard stops it near iteration 29, on the oracle too)")
    ap.add_argument("--kernel", default=None, help="kernel family timed for the roofline (default: the headline "
                                                    "form's smoother)")
    # a single custom form (overrides --headline) for sweeps
    ap.add_argument("--storage", choices=["f32", "f16", "bf16"])
    ap.add_argument("--smoother", choices=["rbgs", "jacobi", "brick"])
    ap.add_argument("--nu", type=int)
    ap.add_argument("--jacobi-omega", type=float)
    ap.add_argument("--coarse-space", choices=["none", "free"])
    ap.add_argument("--levels-custom", type=int, dest="levels_custom")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU legs (baseline and FP64 reference)")
    ap.add_argument("--cpu-iters", type=int, default=2, help="PCG iterations in the CPU sample")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run a plain 512 MiB device-to-device copy three times (the known-size kernel that scales "
                         "FETCH_SIZE / WRITE_SIZE counter readings to bytes)")
    return ap.parse_args()


def store_bytes(storage):
    return 4 if storage == "f32" else 2


def fine_storage():
    """FLUX_FINE_STORAGE=1 stores the finest level's x and b in the V-cycle's
    storage type too (opt-in); otherwise they are FP32"""
    return os.environ.get("FLUX_FINE_STORAGE") == "1"


def algo_bytes_per_cell(kernel, storage):
    """Algorithmic bytes per finest-level cell and launch: the unique bytes
    the launch must move (u8 neighbour-mask byte; face weights one value per
    face, three per cell)."""
    s = store_bytes(storage)
    x = s if fine_storage() else 4
    if kernel == "relax":
        # one red-black colour launch: x read, b on the colour (half), x
        # written on the colour (half), mask 1, weights 3 x storage
        # (21 B in FP32; 15 B with FP16/BF16 weights; 11 B with FP16 x, b too)
        return 2 * x + 1 + 3 * s
    if kernel == "jacobi":
        # one steady-state damped-Jacobi sweep: x read, b read, x' written in
        # the fine level's storage (FP32 for f32; FP16/BF16 in the mixed
        # form), mask 1, weights 3 x storage
        return 3 * x + 1 + 3 * s
    if kernel == "brick":
        # one brick pass that reads the field: x read, b read, x written
        # (FP32 on the finest level), mask 1, weights 3 x storage; the halo
        # of the brick (re-read from the neighbouring bricks' cells) is not
        # counted (25 B in FP32; 19 B with FP16/BF16 weights)
        return 3 * x + 1 + 3 * s
    if kernel == "update":
        # x64 read + write (16), p 4, b 4, r written 4, mask 1, weights 12
        return 41.0
    if kernel == "apply":
        # z 4 + p 4 read, p 4 written, mask 1, weights 12
        return 25.0
    if kernel == "residual":
        return 3 * s + 1 + 3 * s if storage != "f32" else 25.0
    raise KeyError(kernel)


# fine-level kernel of each family as rocprofv3 names it
KERNEL_SYMBOL = {("relax", "f32"): "k_relax3<1, 0, 0, float, float>",
                 ("relax", "f16"): "k_relax3<1, 0, 0, float, __half>",
                 ("relax", "bf16"): "k_relax3<1, 0, 0, float, __hip_bfloat16>",
                 ("jacobi", "f32"): "k_jacobi3<1, 0, 0, float, float>",
                 ("jacobi", "f16"): "k_jacobi3<1, 0, 0, float, __half>",
                 ("brick", "f32"): "k_brick3<0, 0, 1, float, float>",
                 ("brick", "f16"): "k_brick3<0, 0, 1, float, __half>",
                 ("brick", "bf16"): "k_brick3<0, 0, 1, float, __hip_bfloat16>",
                 ("update", "f32"): "k_pcg_update3", ("apply", "f32"): "k_pcg_apply3"}
if fine_storage():
    KERNEL_SYMBOL.update({("relax", "f16"): "k_relax3m<0, 0, __half>",
                          ("relax", "bf16"): "k_relax3m<0, 0, __hip_bfloat16>",
                          ("jacobi", "f16"): "k_jacobi3m<0, 0, __half>",
                          ("jacobi", "bf16"): "k_jacobi3m<0, 0, __hip_bfloat16>"})


def measured_traffic(kernel, storage):
    """HBM bytes per dispatch of the family's fine-level kernel from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary (copy-calibrated), or
    None when no summary covers it."""
    # the brick smoother's summary is its own (headline form mixed_f16_brick_cs)
    name = "pmc_traffic_dambreak256_%s%s.json" % (storage, "_brick" if kernel == "brick" else "")
    path = os.path.join(ROOT, "profiles", name)
    sym = KERNEL_SYMBOL.get((kernel, storage)) or KERNEL_SYMBOL.get((kernel, "f32"))
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


def cpu_sample(sc, levels, gpu_iters, iters, threads):
    """The oracle's restatement of the plain projection (f32 vectors, f64
    solution accumulator: the HIP plain path's arithmetic) on `threads` host
    threads: assembly, V-cycle setup and 1 + `iters` PCG iterations are
    timed; the projection time is extrapolated to the GPU's iteration count."""
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
        "extrapolated": True,
        "sample": "plain form (RBGS 2/2, FP32 V-cycle) on the %d^3 system: assembly + V-cycle setup + %d PCG iterations "
                  "timed (%.1f s); the projection time is extrapolated to the GPU's %d iterations"
                  % (n[0], 1 + iters, t3 - t0, gpu_iters),
        "s_per_iteration": per_iter,
    }


FP64_TOL = 1e-9


def fp64_reference(sc, levels, threads):
    """The same system solved on the host in FP64 (the oracle's f64 PCG with
    its FP64 V-cycle) from the cold start to a relative residual of 1e-9.
    (A warm start from a GPU answer trips the reference's pAp-growth guard,
    whose base is the initial residual.)"""
    from oracle import pyoracle as O
    O.set_threads(threads)
    n, h = sc["shape"], sc["spacing"]
    t0 = time.perf_counter()
    A = O.pressure3_assemble(n, h, sc["vel"], active=sc["active"], target=sc["target"], beta=sc["beta"], suffix="f64")
    mg = O.Multigrid(n, h, levels=levels, mask=A["mask"], coarse_mask=A["vent"], weights=A["weights"], suffix="f64")
    x, st = O.pcg_solve(n, h, A["b"], x0=A["pressure"], mask=A["mask"], weights=A["weights"], mg=mg, tol=FP64_TOL,
                        maxiter=400)
    return x, st, time.perf_counter() - t0


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only; --gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)
    import tsl_array_n_amd as F
    from tsl_array_n_amd import scenes

    forms = dict(FORMS)
    headline = args.headline
    if args.storage or args.smoother or args.nu or args.jacobi_omega or args.coarse_space:
        base = dict(FORMS[headline])
        for k, v in (("storage", args.storage), ("smoother", args.smoother), ("nu", args.nu),
                     ("omega", args.jacobi_omega)):
            if v is not None:
                base[k] = v
        if args.coarse_space:
            base["cs"] = None if args.coarse_space == "none" else args.coarse_space
        forms = {"custom": base}
        headline = "custom"
        run = ["custom"]
    else:
        run = [f for f in args.forms.split(",") if f in FORMS and f != headline]
        run = [headline] + run
    hform = forms[headline]
    kernel = args.kernel or {"jacobi": "jacobi", "brick": "brick"}.get(hform["smoother"], "relax")

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
    if args.system == "flip":
        from tsl_array_n_amd import flip as FL
        print("bench: running the %d^3 two-phase FLIP dam break for %d steps on %s" % (n, args.flip_steps, dev),
              file=sys.stderr, flush=True)
        fsol, parts = FL.two_phase_dam_break(n, levels=args.levels)
        fsol.capture_projection_input(True)
        for k in range(args.flip_steps):
            d = fsol.on_advance_time_step()
            print("bench: FLIP step %d: %s" % (k, d), file=sys.stderr, flush=True)
        sc = fsol.captured_system()
        sc["particles"] = parts["count"]
        sc["liquid_particles"] = parts["liquid"]
        del fsol, parts
        F.sync()
    else:
        print("bench: building the %d^3 dam-break system on %s" % (n, dev), file=sys.stderr, flush=True)
        sc = scenes.dam_break_system(n)
    shape = sc["shape"]
    cells = float(np.prod(shape))
    vin = [F.array_n("float", F.face_shape(shape, a), sc["vel"][a]) for a in range(3)]
    vout = [F.array_n("float", F.face_shape(shape, a)) for a in range(3)]
    beta = [F.array_n("float", F.face_shape(shape, a), sc["beta"][a]) for a in range(3)]
    act = F.array_n("u8", shape, sc["active"])
    tgt = F.array_n("float", shape, sc["target"])

    results, pressures = {}, {}
    line_time = None
    roof = None
    for name in run:
        fm = forms[name]
        ps = F.create_grid_pressure_solver3(
            resolution=shape, grid_spacing=sc["spacing"], dirichlet=(act, tgt), face_weights=beta,
            multigrid={"levels": args.levels, "storage": fm["storage"], "smoother": fm["smoother"],
                       "nu_down": fm["nu"], "nu_up": fm["nu"], "n_final": fm["nu"], "jacobi_omega": fm["omega"],
                       "omega": 1.0},
            tolerance=1e-6, relative_tolerance=True)
        cs = fm.get("cs")
        if cs == "free":
            ps.set_coarse_space(True, [])
        elif cs == "free+exact":
            exact = pressures.get("plain_rbgs")
            if exact is None:
                ps.pressure.fill(0.0)
                ps.project(vin, vout)
                exact = ps.pressure.to_array()
            exact_field = F.array_n("float", shape, exact)
            ps.set_coarse_space(True, [exact_field])

        def step():
            ps.pressure.fill(0.0)
            return ps.project(vin, vout)

        for w in range(args.warmup):
            d = step()
            print("bench: %s warmup %d: %s" % (name, w, d), file=sys.stderr, flush=True)
        steps = args.steps if name == headline else min(args.steps, args.form_steps)
        if name == headline:
            F.kprof_enable(kernel)
        F.sync()
        t0 = time.perf_counter()
        iters, last = [], None
        for s in range(steps):
            last = step()
            iters.append(last["iterations"])
            if not last["converged"]:
                print("bench: %s step %d did not converge: %s" % (name, s, last), file=sys.stderr, flush=True)
        F.sync()
        t1 = time.perf_counter()
        if name == headline:
            k_ms, k_launches = F.kprof_read()
            F.kprof_enable("")
            line_time = (t1 - t0, steps)
            if k_launches:
                avg_s = k_ms / 1e3 / k_launches
                algo = algo_bytes_per_cell(kernel, fm["storage"]) * cells
                achieved = algo / avg_s / 1e9
                roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 4),
                        "traffic": measured_traffic(kernel, fm["storage"]),
                        "kernel": "%s (%s)" % (KERNEL_SYMBOL.get((kernel, fm["storage"]), kernel), fm["storage"]),
                        "avg_us": round(avg_s * 1e6, 2), "launches": k_launches, "algorithmic_bytes": algo,
                        "algorithmic_bytes_per_cell": algo / cells}
        ms = (t1 - t0) / steps * 1e3
        results[name] = {"ms_per_projection": round(ms, 3), "projections_per_s": round(1e3 / ms, 3),
                         "iterations": int(np.median(iters)), "converged": bool(last["converged"]),
                         "fp64_relative_residual": last["residual"] / last["bNorm"] if last["bNorm"] else None,
                         "storage": fm["storage"], "smoother": fm["smoother"], "nu": fm["nu"],
                         "omega": round(fm["omega"], 4), "coarse_space": fm.get("cs")}
        pressures[name] = ps.pressure.to_array()
        print("bench: %s: %d steps, %.3f ms/projection, iterations %s, last %s" % (name, steps, ms, iters, last),
              file=sys.stderr, flush=True)
        del ps

    plain = results.get("plain_rbgs")
    if plain:
        for r in results.values():
            r["speedup_vs_plain"] = round(plain["ms_per_projection"] / r["ms_per_projection"], 3)

    cpu = None
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        gpu_iters = (plain or results[headline])["iterations"]
        print("bench: CPU baseline on %d threads" % threads, file=sys.stderr, flush=True)
        cpu = cpu_sample(sc, args.levels, gpu_iters, args.cpu_iters, threads)
        one = cpu_sample(sc, args.levels, gpu_iters, 1, 1)
        cpu["single_thread_value"] = one["value"]
        cpu["cpu_model"] = cpu_model()
        cpu["nproc"] = os.cpu_count()
        print("bench: FP64 reference solve on %d threads" % threads, file=sys.stderr, flush=True)
        x64, st64, t64 = fp64_reference(sc, args.levels, threads)
        nx = np.linalg.norm(x64)
        for name, p in pressures.items():
            results[name]["error_vs_fp64_cpu_solve"] = (float(np.linalg.norm(p.astype(np.float64) - x64) / nx)
                                                        if st64["converged"] else None)
        cpu["fp64_reference"] = {"tolerance": FP64_TOL, "iterations": st64["iterations"],
                                 "converged": st64["converged"], "stopped_by": st64["stoppedBy"],
                                 "seconds": round(t64, 1), "threads": threads}

    elapsed, steps = line_time
    hf = results[headline]
    line = {
        "metric": "projections/sec to 1e-6 rel residual at 256^3",
        "value": steps / elapsed,
        "unit": "projections/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"f32": "f32", "f16": "f16 storage / f32 arithmetic", "bf16": "bf16 storage / f32 arithmetic"}[
            hform["storage"]] + " V-cycle; f32 CG vectors; f64 accumulator, residual and dots",
        "data": ("synthetic: the projection system of step %d of the seeded two-phase FLIP dam break "
                 "(%d particles, %d liquid), captured from the FLIP solver" % (args.flip_steps, sc["particles"],
                                                                               sc["liquid_particles"])
                 if args.system == "flip" else "synthetic: seeded 3-D two-phase dam break, frame-0 system"),
        "config": {"workload": "3D FLIP two-phase dam break %d^3, liquid:air 1000:1, pinned top-centre, %s, "
                               "MGPCG %d levels, form %s (%s V-cycle, %s %d/%d), cold start"
                               % (n, "system of FLIP step %d" % args.flip_steps if args.system == "flip"
                                  else "frame-0 system", args.levels, headline, hform["storage"], hform["smoother"],
                                  hform["nu"], hform["nu"]),
                   "cells": int(cells), "particles": sc["particles"], "iterations": hf["iterations"],
                   "parallelism": "single GPU", "device": dev},
        "speedup_vs_plain": hf.get("speedup_vs_plain"),
        "forms": results,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
