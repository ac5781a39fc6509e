"""Pressure-system fixtures in the reference's on-disk format
(export_system.mjs:65-126): the reference's own captured frames load with
the documented shapes; export -> load round trips on CPU; on the GPU a
solver's assembled system is exported, reloaded and re-solved to the same
pressure, and the reference's frame-110 capture solves in its recorded
iteration budget."""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN
from tsl_array_n_amd import fixtures


def test_reference_capture_loads():
    s = fixtures.load_system(os.path.join(GOLDEN, "stalled_system", "frame-110"))
    assert s["shape"] == (48, 24, 24)
    assert s["b"].size == 48 * 24 * 24 and s["weights"][0].size == 49 * 24 * 24
    assert s["mask"].dtype == np.uint8 and set(np.unique(s["mask"])) <= {0, 1}
    assert s["meta"]["diagnostics"]["iterations"] == 15
    # the capture's |b| (meta.json of frame 110 records 3.888)
    assert abs(np.linalg.norm(s["b"].astype(np.float64)) - 3.888) < 5e-3


class _View:
    def __init__(self, a):
        self.a = a

    def to_array(self):
        return self.a


class _FakeSolver:
    """the fields export_system reads, held on the host"""

    def __init__(self, s):
        self.resolution = s["shape"]
        self.b = _View(s["b"])
        self.pressure = _View(s["pressure"])
        self.dirichlet_mask = _View(s["mask"])
        self.vent_mask = _View(s["vent"])
        self.weights = [_View(w) for w in s["weights"]]
        self.diagnostics = {"converged": True, "iterations": 15}


def test_export_load_round_trip(tmp_path):
    s = fixtures.load_system(os.path.join(GOLDEN, "stalled_system", "frame-110"))
    meta = fixtures.export_system(tmp_path / "copy", _FakeSolver(s), frame=110, tolerance=1e-6)
    assert meta["dataSizeW"] == [48, 24, 25]
    t = fixtures.load_system(tmp_path / "copy")
    for k in ("b", "pressure", "mask", "vent"):
        np.testing.assert_array_equal(t[k], s[k])
    for a in range(3):
        np.testing.assert_array_equal(t["weights"][a], s["weights"][a])
    m = json.load(open(tmp_path / "copy" / "meta.json"))
    assert m["resolution"] == [48, 24, 24] and m["frame"] == 110


@pytest.mark.gpu
def test_exported_gpu_system_resolves_identically(flux, tmp_path):
    F = flux
    from tsl_array_n_amd import scenes
    sc = scenes.dam_break_system(32)
    shape = sc["shape"]
    vel = [F.array_n("float", F.face_shape(shape, a), sc["vel"][a]) for a in range(3)]
    beta = [F.array_n("float", F.face_shape(shape, a), sc["beta"][a]) for a in range(3)]
    act = F.array_n("u8", shape, sc["active"])
    tgt = F.array_n("float", shape, sc["target"])
    ps = F.create_grid_pressure_solver3(resolution=shape, grid_spacing=sc["spacing"], dirichlet=(act, tgt),
                                        face_weights=beta, multigrid={"levels": 3}, tolerance=1e-6)
    ps.assemble(vel)
    fixtures.export_system(tmp_path / "sys", ps, grid_spacing=sc["spacing"], multigrid={"levels": 3})
    d = ps.solve(vel)
    assert d["converged"]
    s = fixtures.load_system(tmp_path / "sys")
    p, st = fixtures.solve_system(F, s, levels=3, tol=1e-6)
    assert st["converged"] and st["iterations"] == d["iterations"]
    np.testing.assert_array_equal(p, ps.pressure.to_array())


@pytest.mark.gpu
def test_reference_capture_solves(flux):
    s = fixtures.load_system(os.path.join(GOLDEN, "stalled_system", "frame-110"))
    p, st = fixtures.solve_system(flux, s, levels=4, tol=1e-6, maxiter=100,
                                  x0=np.zeros(int(np.prod(s["shape"])), np.float32))
    assert st["converged"]
    print("frame 110:", st)
