"""createGridFlipSolver3 on the GPU against the oracle's composed step (GPU tests).

Three steps of a 16^3 dam break with resampling on: the fluid mask must match
exactly; particle positions to 1e-5 of the domain; particle velocities and
each grid velocity component to 1e-4 of the whole velocity's norm (the projections agree to the PCG tolerance,
1e-6 relative, and nothing else differs in arithmetic).
"""
import numpy as np
import pytest

import tsl_array_n_amd as F
from oracle import pyoracle as O
from tsl_array_n_amd.flip import GridFlipSolver3

pytestmark = pytest.mark.gpu


def initial_state(n, h, block):
    pos = []
    for k in range(block[2]):
        for j in range(block[1]):
            for i in range(block[0]):
                for sk in range(2):
                    for sj in range(2):
                        for si in range(2):
                            pos.append([(i + (si + 0.5) / 2) * h[0], (j + (sj + 0.5) / 2) * h[1],
                                        (k + (sk + 0.5) / 2) * h[2]])
    return np.array(pos, np.float32).ravel()


def rel(a, b):
    return np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30)


def test_flip_steps_match_oracle(flux):
    n, h, dt = (16, 16, 16), (1 / 16,) * 3, 1 / 60
    pos = initial_state(n, h, (6, 10, 16))
    count = pos.size // 3
    s = GridFlipSolver3(resolution=n, grid_spacing=h, max_particles=count, dt=dt, gravity=(0.0, -9.81, 0.0),
                        pressure={"multigrid": {"levels": 3}})
    s.positions.from_array(pos)
    st = {"pos": pos.copy(), "vel": np.zeros_like(pos), "pressure": np.zeros(16 ** 3, np.float32),
          "p_accepted": np.zeros(16 ** 3, np.float32)}
    for a, c in enumerate("uvw"):
        st[c] = np.zeros(int(np.prod(F.face_shape(n, a))), np.float32)
    for step in range(3):
        dg = s.on_advance_time_step()
        do = O.flip3_step(st, n, h, (0, 0, 0), dt, levels=3)
        print("step", step, "gpu", dg, "oracle", do)
        assert dg["converged"] and do["converged"] and not dg["rejected"]
        np.testing.assert_array_equal(s.fluid_mask.to_array(), st["mask"])
        assert np.max(np.abs(s.positions.to_array() - st["pos"])) <= 1e-5
        assert rel(s.velocities.to_array(), st["vel"]) <= 1e-4
        # per component against the norm of the whole grid velocity (the
        # dam break's W is rounding noise around 0)
        scale = np.linalg.norm(np.concatenate([st[c] for c in "uvw"]))
        for a, c in enumerate("uvw"):
            assert np.linalg.norm(s.velocity_grid[a].to_array() - st[c]) <= 1e-4 * scale
        assert rel(s.pressure.to_array(), st["pressure"]) <= 1e-4


@pytest.mark.parametrize("push", [(0.0, 0.0), (0.05, 0.01)])
def test_flip_steps_with_a_collider_match_oracle(flux, push):
    """the dam break falling onto a static ball: the particle trace stops at
    the ball, particles inside are pushed out (colliderPushThresh / Shift),
    constrainVelocity projects against the ball; three steps against the
    oracle's composed step at the same bars as above"""
    from tsl_array_n_amd import boundary as B
    n, h, dt = (16, 16, 16), (1 / 16,) * 3, 1 / 60
    pos = initial_state(n, h, (10, 10, 16))
    count = pos.size // 3
    sdf = O.collider_sdf3([B.sphere((0.55, 0.2, 0.5), 0.16)], n, h, (0.0, 0.0, 0.0))
    col = F.array_n("float", n, sdf)
    s = GridFlipSolver3(resolution=n, grid_spacing=h, max_particles=count, dt=dt, gravity=(0.0, -9.81, 0.0),
                        pressure={"multigrid": {"levels": 3}}, collider_sdf=col, friction=0.5,
                        collider_push_thresh=push[0], collider_push_shift=push[1])
    s.positions.from_array(pos)
    st = {"pos": pos.copy(), "vel": np.zeros_like(pos), "pressure": np.zeros(16 ** 3, np.float32),
          "p_accepted": np.zeros(16 ** 3, np.float32)}
    for a, c in enumerate("uvw"):
        st[c] = np.zeros(int(np.prod(F.face_shape(n, a))), np.float32)
    for step in range(3):
        dg = s.on_advance_time_step()
        do = O.flip3_step(st, n, h, (0, 0, 0), dt, levels=3, collider_sdf=sdf, friction=0.5, push_thresh=push[0],
                          push_shift=push[1])
        print("step", step, "gpu", dg, "oracle", do)
        assert dg["converged"] and do["converged"] and not dg["rejected"]
        np.testing.assert_array_equal(s.fluid_mask.to_array(), st["mask"])
        assert np.max(np.abs(s.positions.to_array() - st["pos"])) <= 1e-5
        assert rel(s.velocities.to_array(), st["vel"]) <= 1e-4
        scale = np.linalg.norm(np.concatenate([st[c] for c in "uvw"]))
        for a, c in enumerate("uvw"):
            assert np.linalg.norm(s.velocity_grid[a].to_array() - st[c]) <= 1e-4 * scale
