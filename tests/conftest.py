import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def golden(name):
    """A committed fixture written by tests/golden/make_fixtures.py from the
    reference's own programs."""
    return np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)


MOLECULES = ["co2", "heh", "h2o", "h2o_2", "h2o_4", "mol0", "mol1", "mol2", "mol3", "mol4", "mol5"]
CONVERGED = ["co2", "heh", "h2o", "h2o_2", "h2o_4", "mol2", "mol3"]


@pytest.fixture(scope="session")
def ctx():
    from myqc_amd.hip import Context
    c = Context(0)
    yield c
    c.close()
