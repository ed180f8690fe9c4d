import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


def relerr(a, b) -> float:
    """Normwise relative difference ||a - b||_2 / ||b||_2 (0 when both vanish).

    NaN entries must sit at the same positions in both (the reference itself
    produces them in some configurations); they are then left out.  A NaN
    mismatch returns inf.
    """
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    na, nbm = np.isnan(a), np.isnan(b)
    if not np.array_equal(na, nbm):
        return float("inf")
    a, b = a[~na], b[~nbm]
    nb = np.linalg.norm(b)
    d = np.linalg.norm(a - b)
    if nb == 0.0:
        return float(d)
    return float(d / nb)


@pytest.fixture
def tiny_mesh():
    from unsflow2d_amd import meshgen
    return meshgen.bump_channel(12, 6, 7, h=0.08)
