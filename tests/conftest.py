import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libfluxhip.so")


def load_capture(name):
    """A captured pressure system (export_system.mjs:65-126 format)."""
    d = os.path.join(GOLDEN, "stalled_system", name)
    meta = json.load(open(os.path.join(d, "meta.json")))
    rd = lambda n: np.fromfile(os.path.join(d, n + ".bin"), dtype="<f4")
    cap = {
        "meta": meta,
        "shape": tuple(meta["resolution"]),
        "spacing": tuple(float(h) for h in meta["gridSpacing"]),
        "b": rd("b"),
        "pressure": rd("pressure"),
        "mask": (rd("dirichletMask") > 0.5).astype(np.uint8),
        "vent": (rd("ventMask") > 0.5).astype(np.uint8),
        "weights": [rd("weightU"), rd("weightV"), rd("weightW")],
    }
    return cap


@pytest.fixture(scope="session")
def frame110():
    return load_capture("frame-110")


@pytest.fixture(scope="session")
def frame112():
    return load_capture("frame-112")


@pytest.fixture(scope="session")
def flux():
    """The HIP library, initialised on device 0 (GPU tests only)."""
    import tsl_array_n_amd as F
    F.init(0)
    yield F
