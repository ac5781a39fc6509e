"""The damped-Jacobi smoother on the GPU against the oracle (row N1,
sandbox/jacobi-smoother/multigrid_jacobi.js:622-642, 1271-1324).

Tolerances as for red-black: one V-cycle relative L2 <= 1e-5; PCG solves
relative L2 <= 1e-5 with iteration counts within 1 of the oracle's mixed arm.
"""
import numpy as np
import pytest

from oracle import pyoracle as O
from tsl_array_n_amd import scenes

from test_gpu_solver import rel, upload, weights_fields, random_case

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("levels", [1, 2, 4])
def test_jacobi_vcycle_matches_oracle(flux, frame110, levels):
    F, c = flux, frame110
    sh = c["shape"]
    m = upload(F, "u8", sh, c["mask"])
    v = upload(F, "u8", sh, c["vent"])
    w = weights_fields(F, sh, c["weights"])
    mg = F.create_multigrid_preconditioner(sh, c["spacing"], dirichlet_mask=m, coarse_dirichlet_mask=v,
                                           face_weights=w, levels=levels, smoother="jacobi")
    ref = O.Multigrid(sh, c["spacing"], levels=levels, mask=c["mask"], coarse_mask=c["vent"], weights=c["weights"],
                      suffix="f32", smoother="jacobi")
    z = F.array_n("float", sh)
    mg(upload(F, "float", sh, c["b"]), z)
    assert rel(z.to_array(), ref.apply(c["b"])) <= 1e-5


@pytest.mark.parametrize("omega", [2.0 / 3.0, 0.8])
def test_jacobi_vcycle_dam_break_64(flux, omega):
    """Multi-block pair kernels (64^3, z-chunks) and the V-cycle tail."""
    F = flux
    sc = scenes.dam_break_system(64)
    sh, h = sc["shape"], sc["spacing"]
    A = O.pressure3_assemble(sh, h, sc["vel"], active=sc["active"], target=sc["target"], beta=sc["beta"],
                             suffix="f32")
    m = upload(F, "u8", sh, A["mask"])
    v = upload(F, "u8", sh, A["vent"])
    w = [upload(F, "float", F.face_shape(sh, a), A["weights"][a]) for a in range(3)]
    for nu in (2, 4):
        mg = F.create_multigrid_preconditioner(sh, h, dirichlet_mask=m, coarse_dirichlet_mask=v, face_weights=w,
                                               levels=4, nu_down=nu, nu_up=nu, n_final=nu, smoother="jacobi",
                                               jacobi_omega=omega)
        ref = O.Multigrid(sh, h, levels=4, nu_down=nu, nu_up=nu, n_final=nu, mask=A["mask"], coarse_mask=A["vent"],
                          weights=A["weights"], suffix="f32", smoother="jacobi", omega=omega)
        r = np.random.default_rng(nu).standard_normal(int(np.prod(sh))).astype(np.float32)
        z = F.array_n("float", sh)
        mg(upload(F, "float", sh, r), z)
        assert rel(z.to_array(), ref.apply(r)) <= 1e-5, nu


def test_jacobi_vcycle_2d_and_odd_rows(flux):
    F = flux
    for sh, sp in (((64, 32), (1.0, 1.0)), ((36, 20, 12), (0.5, 1.0, 1.0))):
        mask, ws = random_case(sh, 9)
        m = upload(F, "u8", sh, mask)
        w = weights_fields(F, sh, ws)
        mg = F.create_multigrid_preconditioner(sh, sp, dirichlet_mask=m, face_weights=w, levels=3, smoother="jacobi")
        r = np.random.default_rng(10).standard_normal(int(np.prod(sh))).astype(np.float32)
        z = F.array_n("float", sh)
        mg(upload(F, "float", sh, r), z)
        ref = O.Multigrid(sh, sp, levels=3, mask=mask, weights=ws, suffix="f32", smoother="jacobi").apply(r)
        assert rel(z.to_array(), ref) <= 1e-5, sh


def test_jacobi_pcg_matches_oracle(flux, frame110, frame112):
    F = flux
    for c in (frame110, frame112):
        sh = c["shape"]
        m = upload(F, "u8", sh, c["mask"])
        v = upload(F, "u8", sh, c["vent"])
        w = weights_fields(F, sh, c["weights"])
        op = F.create_laplacian_operator(sh, c["spacing"], m, w)
        mg = F.create_multigrid_preconditioner(sh, c["spacing"], dirichlet_mask=m, coarse_dirichlet_mask=v,
                                               face_weights=w, levels=4, smoother="jacobi", nu_down=4, nu_up=4,
                                               n_final=4)
        b = upload(F, "float", sh, c["b"])
        x = F.array_n("float", sh)
        s = F.create_pcg_solver(op, "multigrid", b, x, multigrid=mg)
        ok = s.solve(1e-6, 200, relative_tolerance=True)
        ref_mg = O.Multigrid(sh, c["spacing"], levels=4, nu_down=4, nu_up=4, n_final=4, mask=c["mask"],
                             coarse_mask=c["vent"], weights=c["weights"], suffix="f32", smoother="jacobi")
        xr, _, st = O.pcg_solve_mixed(sh, c["spacing"], c["b"], mask=c["mask"], weights=c["weights"], mg=ref_mg,
                                      tol=1e-6, maxiter=200)
        print("gpu", s.state, "oracle", st)
        assert ok and st["converged"]
        assert abs(s.state["iterations"] - st["iterations"]) <= 1
        assert rel(x.to_array(), xr) <= 1e-5


def test_jacobi_odd_sweeps_rejected(flux):
    F = flux
    with pytest.raises(F.FluxError, match="must be even"):
        F.create_multigrid_preconditioner((8, 8, 8), (1, 1, 1), levels=2, nu_down=3, smoother="jacobi")

