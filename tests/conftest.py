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


def oracle_threads():
    """Threads for oracle sums in GPU tests: the CPUs this process may use,
    at most 16 (one GPU's share of the host)."""
    return max(1, min(16, len(os.sched_getaffinity(0))))


def par_oracle(fn, id_i, xi, vi, *rest):
    """Runs an oracle pair sum (oracle.force / oracle.force_ld) over i-chunks
    on several threads (the ctypes call releases the GIL) and concatenates
    the per-i outputs."""
    from concurrent.futures import ThreadPoolExecutor
    ni = len(id_i)
    nt = min(oracle_threads(), max(1, ni))
    cuts = np.linspace(0, ni, nt + 1).astype(int)
    with ThreadPoolExecutor(nt) as ex:
        parts = list(ex.map(lambda k: fn(id_i[cuts[k]:cuts[k + 1]], xi[cuts[k]:cuts[k + 1]],
                                         vi[cuts[k]:cuts[k + 1]], *rest), range(nt)))
    return tuple(np.concatenate([p[q] for p in parts]) for q in range(len(parts[0])))
