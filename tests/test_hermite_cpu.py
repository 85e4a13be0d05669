"""The block-step Hermite driver (amuse_amd/csrc/hermite4.cpp) on CPU
GRAPE-6 implementations: the compiled reference lib/g6lib/g6lib.c and the
oracle behind the same ABI.

* on the reference library and on the oracle in predictor-compat mode the
  trajectories are bit-identical (same forces, same driver);
* block-step invariants of ph4 (idata.cc:460-510, jdata.cc:531-561): every
  step a power of two <= 1/32, every particle time a multiple of its step;
* energy is conserved to the level a 4th-order scheme gives at eta = 0.01.
"""
import os

import numpy as np
import pytest

import oracle
from amuse_amd import ic
from amuse_amd.hermite import Hermite4

ORACLE_ABI = os.path.join(os.path.dirname(oracle.__file__), "libg6oracle.so")


def small_plummer(n=64, seed=3):
    return ic.plummer(n, seed)


def run(libpath, m, x, v, t_end, compat_env=False, npipes=0):
    old = os.environ.get("G6_PREDICTOR_COMPAT")
    if compat_env:
        os.environ["G6_PREDICTOR_COMPAT"] = "g6lib"
    else:
        os.environ.pop("G6_PREDICTOR_COMPAT", None)
    try:
        h = Hermite4(libpath, npipes=npipes)
        h.commit(m, x, v, eps2=1e-4, eta=0.01)
        e0 = sum(h.energy())
        h.evolve(t_end)
        e1 = sum(h.energy())
        xs, vs = h.state()
        t, dt = h.timesteps()
        cnt = h.counters()
        tnow = h.time
        h.close()
    finally:
        if old is None:
            os.environ.pop("G6_PREDICTOR_COMPAT", None)
        else:
            os.environ["G6_PREDICTOR_COMPAT"] = old
    return dict(e0=e0, e1=e1, x=xs, v=vs, t=t, dt=dt, cnt=cnt, time=tnow)


@pytest.mark.skipif(not os.path.exists(oracle.REF_SO), reason="reference g6lib not compiled")
def test_driver_on_reference_equals_oracle_compat():
    oracle.build()
    m, x, v = small_plummer()
    a = run(oracle.REF_SO, m, x, v, 0.125, npipes=256)
    b = run(ORACLE_ABI, m, x, v, 0.125, compat_env=True)
    assert a["cnt"]["block_steps"] == b["cnt"]["block_steps"] > 10
    assert np.array_equal(a["x"], b["x"]) and np.array_equal(a["v"], b["v"])
    assert a["e1"] == b["e1"]


def test_block_step_invariants_and_energy():
    oracle.build()
    m, x, v = small_plummer(128, 4)
    r = run(ORACLE_ABI, m, x, v, 0.25)
    assert r["time"] == 0.25
    dt, t = r["dt"], r["t"]
    mant, _ = np.frexp(dt)
    assert np.all(mant == 0.5), "steps are powers of two"
    assert dt.max() <= 0.03125
    assert np.all(np.fmod(t, dt) == 0), "times commensurate with steps"
    assert np.all(t == 0.25)   # dyadic end time: everyone synchronised
    de = abs((r["e1"] - r["e0"]) / r["e0"])
    assert de < 1e-8, de
    assert r["cnt"]["particle_steps"] > r["cnt"]["block_steps"]


def test_initial_energy_matches_direct_sum():
    oracle.build()
    m, x, v = small_plummer(100, 5)
    h = Hermite4(ORACLE_ABI)
    h.commit(m, x, v, eps2=1e-4, eta=0.01)
    k, u = h.energy()
    h.close()
    assert k == pytest.approx(ic.kinetic_energy(m, v), rel=1e-14)
    assert u == pytest.approx(ic.potential_energy(m, x, 1e-4), rel=1e-12)


def test_evolve_block_budget():
    """evolve() reports whether t_end was reached or the block budget ran
    out first; a run continued without a budget ends where an uninterrupted
    run ends, bit for bit."""
    oracle.build()
    m, x, v = small_plummer(64, 6)
    h = Hermite4(ORACLE_ABI)
    h.commit(m, x, v, eps2=1e-4, eta=0.01)
    assert h.evolve(0.0625, max_blocks=5) is False
    assert h.counters()["block_steps"] == 5
    assert h.evolve(0.0625) is True
    xa, va = h.state()
    ba = h.counters()["block_steps"]
    h.close()
    h = Hermite4(ORACLE_ABI)
    h.commit(m, x, v, eps2=1e-4, eta=0.01)
    assert h.evolve(0.0625) is True
    xb, vb = h.state()
    assert h.counters()["block_steps"] == ba
    h.close()
    assert np.array_equal(xa, xb) and np.array_equal(va, vb)
