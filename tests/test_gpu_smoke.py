"""createGridSmokeSolver3 on the GPU against the oracle's composed step (GPU tests).

A 24^3 closed box with a hot, dense ball re-injected every step (the smoke
plume scene at reduced size), four steps with order-1 and with MacCormack
advection: velocity, density and
temperature agree to 1e-4 of their norms (the projections agree to the PCG
tolerance; everything else is the same f32 arithmetic in the same order).
"""
import numpy as np
import pytest

from oracle import pyoracle as O
from tsl_array_n_amd.smoke import GridSmokeSolver3

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("order", [1, 2])
def test_smoke_steps_match_oracle(flux, order):
    n, h, dt = (24, 24, 24), (1.0, 1.0, 1.0), 0.25
    kw = dict(a=-0.05, b=0.1, order=order)
    s = GridSmokeSolver3(resolution=n, grid_spacing=h, dt=dt, buoyancy_smoke_density_factor=-0.05,
                         buoyancy_temperature_factor=0.1, pressure={"multigrid": {"levels": 3}},
                         advection={"order": order})
    cells = int(np.prod(n))
    st = {"dens": np.zeros(cells, np.float32), "temp": np.zeros(cells, np.float32),
          "pressure": np.zeros(cells, np.float32), "p_accepted": np.zeros(cells, np.float32)}
    for a, c in enumerate("uvw"):
        st[c] = np.zeros(int(np.prod(O.face_shape(n, a))), np.float32)
    src = (12.0, 2.5, 12.0)
    for step in range(4):
        s.source_sphere(src, 3.0, 1.0, 1.0)
        st["dens"] = O.source_sphere3(st["dens"], n, h, (0, 0, 0), src, 3.0, 1.0)
        st["temp"] = O.source_sphere3(st["temp"], n, h, (0, 0, 0), src, 3.0, 1.0)
        dg = s.on_advance_time_step()
        do = O.smoke3_step(st, n, h, (0, 0, 0), dt, levels=3, **kw)
        print("step", step, dg, do)
        assert dg["converged"] and do["converged"]
        vel = np.concatenate([st[c] for c in "uvw"])
        got = np.concatenate([f.to_array() for f in s.velocity_grid])
        assert np.linalg.norm(got - vel) <= 1e-4 * np.linalg.norm(vel)
        for name, key in (("density", "dens"), ("temperature", "temp")):
            g = s.field(name).to_array()
            assert np.linalg.norm(g - st[key]) <= 1e-4 * np.linalg.norm(st[key])
    assert np.max(st["v"]) > 0.01   # the plume moves up
