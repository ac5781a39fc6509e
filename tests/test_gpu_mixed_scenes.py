"""Mixed-precision MGPCG on 64 seeded random 3-D scenes at 128^3 (the
mixed-precision configuration, scenes.mixed_precision_scene): random unions
of spheres and boxes of denser fluid (radius 6-30 cells, density ratio
log-uniform in [1, 1000]), the divergence of a band-limited Gaussian-noise
velocity as the RHS, closed walls and an open top layer; the V-cycle in
FP32, FP16 and BF16 storage, the outer CG with FP32 vectors and FP64
reductions, to a 1e-6 relative true residual.  Each GPU pressure is compared
with the oracle's FP64 solve of the same system at 1e-9 and the per-scene
table is printed."""
import numpy as np
import pytest

from oracle import pyoracle as O
from tsl_array_n_amd import scenes

pytestmark = pytest.mark.gpu

N, LEVELS = 128, 5
# the 64 scenes of the mixed-precision workload, four tests of 16
SEED0, SCENES, CHUNK = 11, 64, 16


def _fp64_reference(seed):
    """scene + the oracle's FP64 MGPCG solve at 1e-9 (runs in a worker
    thread: the oracle's C calls release the interpreter lock)"""
    sc = scenes.mixed_precision_scene(N, seed)
    shape, h = sc["shape"], sc["spacing"]
    A = O.pressure3_assemble(shape, h, sc["vel"], active=sc["active"], target=sc["target"], beta=sc["beta"],
                             suffix="f64")
    mg = O.Multigrid(shape, h, levels=LEVELS, mask=A["mask"], coarse_mask=A["vent"], weights=A["weights"],
                     suffix="f64")
    ref, st = O.pcg_solve(shape, h, A["b"], mask=A["mask"], weights=A["weights"], mg=mg, tol=1e-9, maxiter=400)
    del mg, A
    return sc, ref, st


@pytest.mark.parametrize("chunk", range(SCENES // CHUNK))
def test_mixed_precision_random_scenes_against_fp64(flux, chunk):
    import os
    from concurrent.futures import ThreadPoolExecutor
    F = flux
    O.lib()  # load and declare the oracle library once, before the worker threads
    seeds = [SEED0 + chunk * CHUNK + q for q in range(CHUNK)]
    rows = []
    with ThreadPoolExecutor(max_workers=max(1, min(CHUNK, (os.cpu_count() or 2) - 1))) as pool:
        for seed, fut in [(sd, pool.submit(_fp64_reference, sd)) for sd in seeds]:
            sc, ref, st = fut.result()
            assert st["converged"], (seed, st)
            shape, h = sc["shape"], sc["spacing"]
            row = {"seed": seed, "ratio": sc["ratio"], "ref_iterations": st["iterations"]}
            beta = [F.array_n("float", F.face_shape(shape, a), sc["beta"][a]) for a in range(3)]
            act = F.array_n("u8", shape, sc["active"])
            tgt = F.array_n("float", shape, sc["target"])
            for storage in ("f32", "f16", "bf16"):
                ps = F.create_grid_pressure_solver3(resolution=shape, grid_spacing=h, dirichlet=(act, tgt),
                                                    face_weights=beta,
                                                    multigrid={"levels": LEVELS, "storage": storage},
                                                    tolerance=1e-6)
                # project() corrects the velocity in place: every solve gets the original one
                vel = [F.array_n("float", F.face_shape(shape, a), sc["vel"][a]) for a in range(3)]
                d = ps.project(vel)
                assert d["converged"] and not d["rejected"], (seed, storage, d)
                p = ps.pressure.to_array().astype(np.float64)
                err = np.linalg.norm(p - ref) / np.linalg.norm(ref)
                row[storage] = (d["iterations"], err)
                del ps, vel
            rows.append(row)
            del sc, ref
    print("\nseed  ratio   ref_it | f32 it  err      | f16 it  err      | bf16 it err")
    for r in rows:
        print("%4d %7.1f %6d | %6d %.2e | %6d %.2e | %6d %.2e" % (
            r["seed"], r["ratio"], r["ref_iterations"], *r["f32"], *r["f16"], *r["bf16"]), flush=True)
    for r in rows:
        f32_it = r["f32"][0]
        # every storage stops at the same true-residual tolerance, so each
        # pressure is the FP64 solution to that tolerance's conditioning
        # (1e-7 .. 1.4e-5 over the 64 scenes); reduced storage preconditions
        # a little worse and takes a few more iterations
        for storage in ("f32", "f16", "bf16"):
            assert r[storage][1] <= 1e-4, (r["seed"], storage, r[storage])
        assert r["f16"][0] <= f32_it + 4
        assert r["bf16"][0] <= int(1.6 * f32_it) + 2
