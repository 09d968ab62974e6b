import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")

# Parity bar for FP32 fields: max|a-b| / max|b| over the field.
NORM_REL_TOL = 1e-5


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X / gfx950)")


def norm_rel(a, b):
    """max|a - b| / max|b| (norm-wise relative error)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.abs(b).max() if b.size else 0.0
    num = np.abs(a - b).max() if b.size else 0.0
    if den == 0.0:
        return float(num)
    return float(num / den)


def load_golden(name):
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle
    oracle.build(with_reference=False)
    return oracle.load_oracle()


@pytest.fixture(scope="session")
def orc64():
    from oracle import oracle
    oracle.build(with_reference=False)
    return oracle.Oracle64()


@pytest.fixture(scope="session")
def ref():
    from oracle import oracle
    return oracle.load_reference()


@pytest.fixture(scope="session")
def ctx():
    import gridfluidsim3d_amd as g
    c = g.Context(0)
    yield c
    c.close()
