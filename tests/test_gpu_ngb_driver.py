"""Neighbour lists and the block-step driver on the GPU library.

Neighbour lists (lib/sapporo_light/sapporo.cpp:234-272 interface): for every
i whose count fits the 256 kept ids, the sorted list equals the oracle's
brute-force radius query exactly (r^2 rounds like the reference, so boundary
cases agree); counts are exact; the overflow flag is set iff some count
exceeds 256, and an overflowing list holds 256 true neighbours.

Driver: the same block-step Hermite code on libg6.so and on the CPU oracle
behind the ABI follows the same trajectory (forces agree to ~1e-16, so the
states agree to far better than the integration error over a short run),
and in predictor-compat mode it reproduces the run on the compiled
reference lib/g6lib/g6lib.c.
"""
import os

import numpy as np
import pytest

import oracle
from amuse_amd import ic
from amuse_amd.g6 import LIBPATH
from amuse_amd.hermite import Hermite4
from conftest import print_case

pytestmark = pytest.mark.gpu

ORACLE_ABI = os.path.join(os.path.dirname(oracle.__file__), "libg6oracle.so")


@pytest.mark.parametrize("prec", [0, 1])
def test_neighbour_lists(g6, prec):
    n = 4096
    m, x, v = ic.plummer(n, 31)
    ids = np.arange(1, n + 1, dtype=np.int32)
    z = np.zeros((n, 3))
    g6.set_j_particles(np.arange(n), ids, np.zeros(n), 0.0, m, z, z, z, v, x)
    g6.set_ti(0.0)
    g6.set_precision(prec)
    rng = np.random.Generator(np.random.PCG64(1))
    sel = rng.choice(n, 300, replace=False)
    h2 = rng.uniform(0, 0.05, sel.size) ** 2
    h2[:5] = 0.5        # large radius: several hundred neighbours -> overflow
    h2[5:10] = 0.0      # no neighbours
    g6.calc(n, ids[sel], x[sel], v[sel], 1e-4, h2=h2)
    flag = g6.read_neighbour_list()
    counts, lists = oracle.ngb(ids[sel], x[sel], h2, ids, x, cap=n)
    assert flag == int((counts > 256).any()) == 1
    for q in range(sel.size):
        rc, cnt, got = g6.get_neighbour_list(q, 1000)
        assert cnt == counts[q], q
        truth = np.sort(lists[q, :counts[q]])
        if counts[q] <= 256:
            assert rc == 0
            assert np.array_equal(got, truth), q
        else:
            assert rc == 1 and got.size == 256
            assert np.all(np.isin(got, truth)) and np.all(np.diff(got) > 0)
    # maxlength shorter than the list: truncated copy, flag 1
    q = int(np.argmax((counts > 3) & (counts <= 256)))
    rc, cnt, got = g6.get_neighbour_list(q, 3)
    assert rc == 1 and cnt == counts[q] and got.size == 3
    from amuse_amd.g6 import G6Error
    with pytest.raises(G6Error):
        g6.get_neighbour_list(sel.size, 10)


def run_driver(libpath, m, x, v, t_end, compat=False, npipes=0, eta=0.01):
    old = os.environ.get("G6_PREDICTOR_COMPAT")
    if compat:
        os.environ["G6_PREDICTOR_COMPAT"] = "g6lib"
    else:
        os.environ.pop("G6_PREDICTOR_COMPAT", None)
    try:
        h = Hermite4(libpath, npipes=npipes, compat=compat)
        h.commit(m, x, v, eps2=1e-4, eta=eta)
        e0 = sum(h.energy())
        h.evolve(t_end)
        e1 = sum(h.energy())
        xs, vs = h.state()
        cnt = h.counters()
        h.close()
    finally:
        if old is None:
            os.environ.pop("G6_PREDICTOR_COMPAT", None)
        else:
            os.environ["G6_PREDICTOR_COMPAT"] = old
    return dict(e0=e0, e1=e1, x=xs, v=vs, cnt=cnt)


def test_driver_gpu_matches_oracle():
    m, x, v = ic.plummer(256, 41)
    print_case("driver gpu vs oracle N=256 t=1/16")
    a = run_driver(LIBPATH, m, x, v, 0.0625)
    b = run_driver(ORACLE_ABI, m, x, v, 0.0625)
    assert a["cnt"]["block_steps"] == b["cnt"]["block_steps"]
    assert a["cnt"]["particle_steps"] == b["cnt"]["particle_steps"]
    assert np.abs(a["x"] - b["x"]).max() < 1e-11
    assert np.abs(a["v"] - b["v"]).max() < 1e-10
    assert abs(a["e1"] - b["e1"]) < 1e-12 * abs(b["e1"])


@pytest.mark.skipif(not os.path.exists(oracle.REF_SO), reason="reference g6lib not built")
def test_driver_gpu_compat_matches_reference_g6lib():
    m, x, v = ic.plummer(128, 42)
    print_case("driver gpu compat vs reference g6lib N=128 t=1/16")
    a = run_driver(LIBPATH, m, x, v, 0.0625, compat=True)
    b = run_driver(oracle.REF_SO, m, x, v, 0.0625, compat=True, npipes=128)
    assert a["cnt"]["block_steps"] == b["cnt"]["block_steps"]
    assert np.abs(a["x"] - b["x"]).max() < 1e-11
    assert abs(a["e1"] - b["e1"]) < 1e-12 * abs(b["e1"])


def test_driver_mixed_energy():
    """Mixed-precision forces in the same integration, energies measured with
    the FP64 path: the energy error stays within a small factor of the FP64
    run's over the same interval."""
    m, x, v = ic.plummer(1024, 43)
    print_case("driver fp64 vs mixed N=1024 t=1/8")
    de = {}
    for prec in (0, 1):
        h = Hermite4(LIBPATH, precision=prec)
        h.commit(m, x, v, eps2=1e-4, eta=0.01)
        h.set_precision(0)
        e0 = sum(h.energy())
        h.set_precision(prec)
        h.evolve(0.125)
        h.set_precision(0)
        e1 = sum(h.energy())
        h.close()
        de[prec] = abs((e1 - e0) / e0)
    print("relative energy error fp64 / mixed:", de)
    assert de[0] < 1e-9
    assert de[1] < 100 * de[0] + 1e-9
