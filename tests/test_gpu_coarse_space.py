"""The additive coarse-space correction of the preconditioner on the GPU
(flux_pcg_set_coarse_space) against its numpy restatement
(pyoracle.pcg_coarse_space, pinned in test_oracle_coarse_space.py): the
64^3 dam break, Phi = {free-cell indicator} and {free-cell indicator, a
smooth field}; iterations within 1, solutions within 1e-5 relative L2."""
import numpy as np
import pytest

from oracle import pyoracle as O
from tsl_array_n_amd import scenes

from test_gpu_solver import rel, upload

pytestmark = pytest.mark.gpu


def _smooth(n):
    k, j, i = np.meshgrid(*[np.arange(n)] * 3, indexing="ij")
    return (np.cos(np.pi * (i + 0.5) / n) * (1 + 0.3 * np.sin(np.pi * (j + 0.5) / n))).ravel().astype(np.float32)


@pytest.mark.parametrize("extra", [0, 1])
def test_coarse_space_pcg_matches_restatement(flux, extra):
    F = flux
    n = 64
    sc = scenes.dam_break_system(n)
    sh, h = sc["shape"], sc["spacing"]
    A = O.pressure3_assemble(sh, h, sc["vel"], active=sc["active"], target=sc["target"], beta=sc["beta"], suffix="f32")
    m = upload(F, "u8", sh, A["mask"])
    v = upload(F, "u8", sh, A["vent"])
    w = [upload(F, "float", F.face_shape(sh, a), A["weights"][a]) for a in range(3)]
    op = F.create_laplacian_operator(sh, h, m, w)
    mg = F.create_multigrid_preconditioner(sh, h, dirichlet_mask=m, coarse_dirichlet_mask=v, face_weights=w, levels=4)
    b = upload(F, "float", sh, A["b"])
    x = F.array_n("float", sh)
    s = F.create_pcg_solver(op, "multigrid", b, x, multigrid=mg)
    vecs = [_smooth(n)] if extra else []
    s.set_coarse_space(True, [upload(F, "float", sh, q) for q in vecs])
    ok = s.solve(1e-6, 200, relative_tolerance=True)
    ref_mg = O.Multigrid(sh, h, levels=4, mask=A["mask"], coarse_mask=A["vent"], weights=A["weights"], suffix="f32")
    xr, st = O.pcg_coarse_space(sh, h, A["b"], A["mask"], A["weights"], ref_mg, vecs, free_indicator=True, tol=1e-6)
    s0 = F.create_pcg_solver(op, "multigrid", b, F.array_n("float", sh), multigrid=mg)
    s0.solve(1e-6, 200, relative_tolerance=True)
    print("gpu", s.state, "restatement", st, "without the correction", s0.state["iterations"])
    assert ok and st["converged"]
    assert abs(s.state["iterations"] - st["iterations"]) <= 1
    assert s.state["iterations"] < s0.state["iterations"]
    assert rel(x.to_array(), xr) <= 1e-5
    # removing the correction restores the plain solver
    s.set_coarse_space(False, [])
    x.fill(0.0)
    s.solve(1e-6, 200, relative_tolerance=True)
    assert s.state["iterations"] == s0.state["iterations"]
