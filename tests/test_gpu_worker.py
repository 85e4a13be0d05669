"""The device-resident integrator driven out of process through AMUSE's
socket protocol (amuse_amd/worker.py <-> amuse_amd/lib/g6_worker) equals the
same integrator in process:

* Plummer N = 1024, eps2 = 1e-4, eta = 0.01: after commit and two
  evolve_model calls, get_state (ph4's raw state, interface.cc:485-502),
  get_time, the block and particle step counts and the energies are
  bit-identical to amuse_amd.blocks.DeviceHermite4 run in this process;
* the seven bodies on a line of src/amuse_ph4/tests/test_ph4.py:970-1004,
  through ph4's calls (set_eps2, new_particle, commit_particles,
  evolve_model, get_particle_timestep, energies): the test's smallest-step
  answers and the golden ph4 run's energies within 1e-9.
"""
import os

import numpy as np
import pytest

from amuse_amd import ic
from amuse_amd.worker import Ph4Worker
from conftest import print_case

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def load(w, m, x, v, eps2, eta):
    assert w.call("initialize_code")[0] == 0
    assert w.call("set_eps2", eps2)[0] == 0
    assert w.call("set_eta", eta)[0] == 0
    assert w.call("commit_parameters")[0] == 0
    r, out = w.call("new_particle", m, *x.T, *v.T, np.zeros(m.size), -1)
    assert (r == 0).all()
    assert w.call("commit_particles")[0] == 0
    return out["index_of_the_particle"]


def test_worker_equals_in_process_driver():
    from amuse_amd.blocks import DeviceHermite4
    m, x, v = ic.plummer(1024, 5)
    print_case("worker N=1024 vs in-process driver")
    w = Ph4Worker()
    try:
        ids = load(w, m, x, v, 1e-4, 0.01)
        assert ids.tolist() == list(range(1, 1025))
        d = DeviceHermite4(precision=0, steps_per_graph=64)
        d.commit(m, x, v, eps2=1e-4, eta=0.01)
        for t in (1 / 64, 1 / 32):
            assert w.call("evolve_model", t)[0] == 0
            assert d.evolve(t)
            r, st = w.call("get_state", ids)
            assert (r == 0).all()
            xd, vd = d.state_raw()
            assert np.array_equal(np.stack([st["x"], st["y"], st["z"]], 1), xd)
            assert np.array_equal(np.stack([st["vx"], st["vy"], st["vz"]], 1), vd)
            c = d.counters()
            assert w.call("get_time")[1]["time"] == c["system_time"]
            assert w.call("get_block_steps")[1]["block_steps"] == c["block_steps"]
            assert w.call("get_total_steps")[1]["total_steps"] == c["particle_steps"]
            k, u = d.energy()
            assert w.call("get_kinetic_energy")[1]["kinetic_energy"] == k
            assert w.call("get_potential_energy")[1]["potential_energy"] == u
            r, ts = w.call("get_particle_timestep", ids)
            assert np.array_equal(ts["timestep"], d.timesteps()[1])
        d.close()
        # adding particles after commit is refused, the run goes on
        assert w.call("new_particle", 1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, -1)[0] == -2
        assert w.call("evolve_model", 3 / 64)[0] == 0
    finally:
        assert w.stop() == 0


def test_worker_line7_known_answers():
    import sys
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from ph4_run import line7
    g = np.load(os.path.join(ROOT, "tests", "golden", "ph4_line7.npz"))
    m, x, v = line7()
    print_case("worker line7")
    w = Ph4Worker()
    try:
        ids = load(w, m, x, v, 1e-10, 0.14)
        E, DT = [], []
        for t in g["t"]:
            assert w.call("evolve_model", float(t))[0] == 0
            r, ts = w.call("get_particle_timestep", ids)
            DT.append(float(ts["timestep"].min()))
            E.append(w.call("get_kinetic_energy")[1]["kinetic_energy"]
                     + w.call("get_potential_energy")[1]["potential_energy"])
        assert DT == [0.5 ** 10, 0.5 ** 11, 0.5 ** 20, 0.5 ** 10, 0.5 ** 10]
        assert np.allclose(E, g["E"], rtol=1e-9, atol=0)
    finally:
        assert w.stop() == 0
