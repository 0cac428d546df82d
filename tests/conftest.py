import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (os.path.join(ROOT, "hydrograd.jl_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def load_golden_case(name, **kw):
    """Load one of the reference's example cases copied under tests/golden."""
    from hgswe import cases

    d = os.path.join(GOLDEN, name)
    hyd = [f for f in os.listdir(d) if f.endswith(".srhhydro")][0]
    return cases.load_srh_case(d, hyd, **kw)


@pytest.fixture(scope="session")
def golden_dir():
    return GOLDEN


def rel_err(a, b, scale=None):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    s = np.max(np.abs(b)) if scale is None else scale
    return float(np.max(np.abs(a - b)) / max(s, 1e-300))
