"""Mixed-precision MGPCG on seeded random 3-D scenes at 128^3 (the
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

N, LEVELS, SEEDS = 128, 5, (11, 12, 13, 14)


def test_mixed_precision_random_scenes_against_fp64(flux):
    F = flux
    rows = []
    for seed in SEEDS:
        sc = scenes.mixed_precision_scene(N, seed)
        shape, h = sc["shape"], sc["spacing"]
        A = O.pressure3_assemble(shape, h, sc["vel"], active=sc["active"], target=sc["target"], beta=sc["beta"],
                                 suffix="f64")
        mg = O.Multigrid(shape, h, levels=LEVELS, mask=A["mask"], coarse_mask=A["vent"], weights=A["weights"],
                         suffix="f64")
        ref, st = O.pcg_solve(shape, h, A["b"], mask=A["mask"], weights=A["weights"], mg=mg, tol=1e-9, maxiter=400)
        assert st["converged"], st
        row = {"seed": seed, "ratio": sc["ratio"], "ref_iterations": st["iterations"]}
        beta = [F.array_n("float", F.face_shape(shape, a), sc["beta"][a]) for a in range(3)]
        act = F.array_n("u8", shape, sc["active"])
        tgt = F.array_n("float", shape, sc["target"])
        for storage in ("f32", "f16", "bf16"):
            ps = F.create_grid_pressure_solver3(resolution=shape, grid_spacing=h, dirichlet=(act, tgt),
                                                face_weights=beta, multigrid={"levels": LEVELS, "storage": storage},
                                                tolerance=1e-6)
            # project() corrects the velocity in place: every solve gets the original one
            vel = [F.array_n("float", F.face_shape(shape, a), sc["vel"][a]) for a in range(3)]
            d = ps.project(vel)
            assert d["converged"] and not d["rejected"], (seed, storage, d)
            p = ps.pressure.to_array().astype(np.float64)
            err = np.linalg.norm(p - ref) / np.linalg.norm(ref)
            row[storage] = (d["iterations"], err)
        rows.append(row)
    print("\nseed  ratio   ref_it | f32 it  err      | f16 it  err      | bf16 it err")
    for r in rows:
        print("%4d %7.1f %6d | %6d %.2e | %6d %.2e | %6d %.2e" % (
            r["seed"], r["ratio"], r["ref_iterations"], *r["f32"], *r["f16"], *r["bf16"]))
    for r in rows:
        f32_it, f32_err = r["f32"]
        # the FP32 V-cycle's solution is the reference's to the tolerance's
        # conditioning; reduced storage preconditions a little worse but
        # converges to the same true-residual tolerance
        assert f32_err <= 1e-3
        assert r["f16"][0] <= f32_it + 4 and r["f16"][1] <= 10 * f32_err + 1e-6
        assert r["bf16"][0] <= int(1.6 * f32_it) + 2 and r["bf16"][1] <= 10 * f32_err + 1e-6
