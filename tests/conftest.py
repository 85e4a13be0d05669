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


def load_golden(name):
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as d:
        return {k: d[k] for k in d.files}


@pytest.fixture(scope="session")
def g6lib():
    """The HIP library; fails (does not skip) when it is not built."""
    from amuse_amd.g6 import G6
    return G6(0)


@pytest.fixture
def g6(g6lib):
    """An opened unit in the default state (FP64, Taylor predictor)."""
    if g6lib.lib.get_device_count() <= 0:
        pytest.fail("no GPU visible to libg6.so")
    g6lib.open()
    g6lib.set_precision(0)
    g6lib.set_predictor_compat(False)
    g6lib.set_tiles(0, 0)
    yield g6lib
    g6lib.set_tiles(0, 0)
    g6lib.set_precision(0)
    g6lib.set_predictor_compat(False)
    g6lib.close()


def print_case(msg):
    print(msg, flush=True)
