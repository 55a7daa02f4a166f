import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import Oracle

    return Oracle()


@pytest.fixture(scope="session")
def gpu_lib():
    """The HIP library; a GPU test fails (never falls back) when it is missing."""
    from emma_amd import rt

    lib = rt.load()
    if rt.device_count() < 1:
        pytest.fail("no GPU visible to the HIP runtime")
    return lib


def max_rel(a, b, floor=1e-300):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    d = np.abs(a - b) / np.maximum(np.abs(b), floor)
    d[a == b] = 0.0
    return float(d.max()) if d.size else 0.0


def compare_fields(A, B, fields=None, floor=None):
    """Max relative difference per field; `floor` maps a field to the
    magnitude below which differences are measured absolutely."""
    from emma_amd.synth import Fields

    out = {}
    for k in fields or Fields.NAMES:
        a, b = getattr(A, k), getattr(B, k)
        fl = (floor or {}).get(k, 1e-300)
        out[k] = max_rel(a, b, fl)
    return out
