"""Device-resident block-step driver (amuse_amd/blocks.py, g6_blocks.hip).

* Plummer N = 1024, eps2 = 1e-4, eta = 0.01 (configuration c0) to t = 1/2:
  the energy error stays within 2x that of the reference ph4 integrator's run
  on the CPU oracle (tests/golden/ph4_c0_energy.npz) and the block / particle
  step counts within 1 %;
* against the host-side driver over the GRAPE-6 ABI (same library, same
  arithmetic per particle, force kernels of other launch shapes): identical
  block structure and states within 1e-12 over the first 1/32 time unit;
* a run split by a block budget ends bit for bit where an uninterrupted run
  ends (the block order inside a step does not matter);
* mixed precision at ph4's default eta = 0.14: energy error within 2x of
  FP64's;
* energy and state at t = 0 equal the direct sums of the initial conditions.
"""
import os

import numpy as np
import pytest

from amuse_amd import ic
from conftest import print_case

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ph4_c0_energy.npz")


def driver(precision=0):
    from amuse_amd.blocks import DeviceHermite4
    return DeviceHermite4(precision=precision, steps_per_graph=64)


def test_initial_energy_and_state():
    m, x, v = ic.plummer(512, 3)
    d = driver()
    d.commit(m, x, v, eps2=1e-4, eta=0.01)
    k, u = d.energy()
    xs, vs = d.state()
    d.close()
    assert np.array_equal(xs, x) and np.array_equal(vs, v)
    assert k == pytest.approx(ic.kinetic_energy(m, v), rel=1e-14)
    assert u == pytest.approx(ic.potential_energy(m, x, 1e-4), rel=1e-12)


def test_c0_energy_follows_reference_ph4():
    g = np.load(GOLD)
    m, x, v = ic.plummer(int(g["n"]), int(g["seed"]))
    d = driver()
    d.commit(m, x, v, eps2=float(g["eps2"]), eta=float(g["eta"]))
    E = [sum(d.energy())]
    print_case("device driver c0 t=0..0.5")
    assert d.evolve(0.5)
    E.append(sum(d.energy()))
    c = d.counters()
    d.close()
    de = abs((E[1] - E[0]) / E[0])
    dref = abs((g["E"][1] - g["E"][0]) / g["E"][0])
    print("device driver |dE/E|", de, "reference", dref, "blocks", c["block_steps"],
          g["blocks"][1], "steps", c["particle_steps"], g["steps"][1], flush=True)
    assert E[0] == pytest.approx(g["E"][0], rel=1e-14)
    assert de <= 2 * dref + 1e-13
    assert abs(c["block_steps"] - g["blocks"][1]) <= 0.01 * g["blocks"][1]
    assert abs(c["particle_steps"] - g["steps"][1]) <= 0.01 * g["steps"][1]


def test_matches_host_driver():
    from amuse_amd.hermite import Hermite4
    m, x, v = ic.plummer(1024, 7)
    print_case("device vs host driver N=1024 t=1/32")
    h = Hermite4(precision=0)
    h.commit(m, x, v, eps2=1e-4, eta=0.01)
    d = driver()
    d.commit(m, x, v, eps2=1e-4, eta=0.01)
    th, dth = h.timesteps()
    td, dtd = d.timesteps()
    assert np.array_equal(dth, dtd)  # first steps: same forces to 1e-16, same rule
    for t in (1 / 128, 1 / 64, 1 / 32):
        h.evolve(t)
        d.evolve(t)
        ch, cd = h.counters(), d.counters()
        xh, vh = h.state()
        xd, vd = d.state()
        err = np.abs(xd - xh).max()
        print(f"t={t}: blocks host {ch['block_steps']} device {cd['block_steps']}, "
              f"max |dx| {err:.2e}", flush=True)
        assert ch["block_steps"] == cd["block_steps"]
        assert ch["particle_steps"] == cd["particle_steps"]
        assert err < 1e-12 and np.abs(vd - vh).max() < 1e-11
    h.close()
    d.close()


@pytest.mark.parametrize("n", [7, 513, 1001])
def test_sizes_not_multiples_of_32(n):
    """Systems whose state arrays do not fill whole 256-byte rows (the
    driver pads every array to a 256-byte boundary) follow the host driver
    like N = 1024 does."""
    from amuse_amd.hermite import Hermite4
    m, x, v = ic.plummer(n, 70 + n)
    print_case(f"device vs host driver N={n} t=1/64")
    h = Hermite4(precision=0)
    h.commit(m, x, v, eps2=1e-4, eta=0.01)
    d = driver()
    d.commit(m, x, v, eps2=1e-4, eta=0.01)
    h.evolve(1 / 64)
    d.evolve(1 / 64)
    ch, cd = h.counters(), d.counters()
    assert ch["block_steps"] == cd["block_steps"]
    assert ch["particle_steps"] == cd["particle_steps"]
    xh, vh = h.state()
    xd, vd = d.state()
    assert np.abs(xd - xh).max() < 1e-12 and np.abs(vd - vh).max() < 1e-11
    h.close()
    d.close()


def test_budget_split_is_exact():
    m, x, v = ic.plummer(256, 8)
    runs = []
    for split in (False, True):
        d = driver()
        d.commit(m, x, v, eps2=1e-4, eta=0.01)
        if split:
            assert d.evolve(1 / 16, max_blocks=7) is False
            assert d.counters()["block_steps"] == 7
        assert d.evolve(1 / 16) is True
        runs.append((d.state(), d.counters()["block_steps"]))
        d.close()
    (sa, ba), (sb, bb) = runs
    assert ba == bb
    assert np.array_equal(sa[0], sb[0]) and np.array_equal(sa[1], sb[1])


def test_mixed_energy_error_at_default_eta():
    m, x, v = ic.plummer(1024, 1024)
    de = {}
    for prec in (0, 1):
        print_case(f"device driver prec={prec} eta=0.14 t=1")
        d = driver(prec)
        d.commit(m, x, v, eps2=1e-4, eta=0.14)
        e0 = sum(d.energy())
        assert d.evolve(1.0, max_blocks=200000)
        de[prec] = abs((sum(d.energy()) - e0) / e0)
        d.close()
    print("device driver |dE/E| fp64", de[0], "mixed", de[1], flush=True)
    assert de[1] <= 2 * de[0] + 1e-10
