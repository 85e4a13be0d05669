"""The device-resident integrator driven out of process through AMUSE's
socket protocol (amuse_amd/worker.py <-> amuse_amd/lib/g6_worker) equals the
same integrator in process:

* Plummer N = 1024, eps2 = 1e-4, eta = 0.01: after commit and two
  evolve_model calls, get_state (ph4's raw state, interface.cc:485-502),
  get_time, the block and particle step counts and the energies are
  bit-identical to amuse_amd.blocks.DeviceHermite4 run in this process;
* edits after commit (synchronize, add, delete, set, recommit) equal the
  same edits on an in-process driver synchronised and committed anew;
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
    finally:
        assert w.stop() == 0


def test_worker_edits_after_commit():
    """ph4's edit cycle after a run (interface.cc:349-379, 424-580):
    evolve_model to an output time off the step grid, synchronize_model,
    get_state, then new_particle (twice), delete_particle, set_velocity and
    set_mass, recommit_particles and evolve_model again.  In process the same
    is: the driver synchronised (g6d_synchronize), its raw state edited the
    same way, a new driver committed from it, evolved by the remaining
    time.  States, times, step counts (carried over) and energies agree bit
    for bit."""
    from amuse_amd.blocks import DeviceHermite4
    m, x, v = ic.plummer(512, 8)
    t1, t2 = 0.0371, 0.0703
    print_case("worker edits after commit, N=512")
    w = Ph4Worker()
    try:
        ids = load(w, m, x, v, 1e-4, 0.01)
        d = DeviceHermite4(precision=0, steps_per_graph=64)
        d.commit(m, x, v, eps2=1e-4, eta=0.01)
        assert w.call("evolve_model", t1)[0] == 0
        assert d.evolve(t1)
        assert w.call("synchronize_model")[0] == 0
        d.synchronize()
        c1 = d.counters()
        tq, _ = d.timesteps()
        assert (tq == c1["system_time"]).all()   # every particle at the system time
        r, st = w.call("get_state", ids)
        xs, vs = d.state_raw()
        assert np.array_equal(np.stack([st["x"], st["y"], st["z"]], 1), xs)
        assert np.array_equal(np.stack([st["vx"], st["vy"], st["vz"]], 1), vs)
        assert w.call("get_time")[1]["time"] == c1["system_time"]
        assert w.call("get_block_steps")[1]["block_steps"] == c1["block_steps"]
        # the same edits on both sides
        newm = np.array([1.0 / 512, 2.0 / 512])
        newx = np.array([[0.31, -0.2, 0.05], [-0.4, 0.11, -0.27]])
        newv = np.array([[0.01, 0.2, -0.1], [0.0, -0.15, 0.3]])
        r, out = w.call("new_particle", newm, *newx.T, *newv.T, np.zeros(2), -1)
        assert (r == 0).all()
        new_ids = out["index_of_the_particle"]
        assert new_ids.tolist() == [513, 514]
        assert w.call("delete_particle", 77)[0] == 0
        assert w.call("set_velocity", 5, 0.1, -0.2, 0.3)[0] == 0
        assert w.call("set_mass", 9, 3.0 / 512)[0] == 0
        assert w.call("get_number_of_particles")[1]["number_of_particles"] == 513
        assert w.call("recommit_particles")[0] == 0
        ids2 = np.concatenate([np.delete(ids, 76), new_ids]).astype(np.int32)
        m2 = np.concatenate([m, newm])
        x2 = np.concatenate([xs, newx])
        v2 = np.concatenate([vs, newv])
        v2[4] = [0.1, -0.2, 0.3]
        m2[8] = 3.0 / 512
        keep = np.arange(514) != 76
        d2 = DeviceHermite4(precision=0, steps_per_graph=64)
        d2.commit(m2[keep], x2[keep], v2[keep], eps2=1e-4, eta=0.01, ids=ids2)
        d.close()
        assert w.call("evolve_model", t2)[0] == 0
        assert d2.evolve(t2 - c1["system_time"])
        c2 = d2.counters()
        r, st = w.call("get_state", ids2)
        assert (r == 0).all()
        xd, vd = d2.state_raw()
        assert np.array_equal(np.stack([st["x"], st["y"], st["z"]], 1), xd)
        assert np.array_equal(np.stack([st["vx"], st["vy"], st["vz"]], 1), vd)
        assert np.array_equal(st["mass"], m2[keep])
        assert w.call("get_time")[1]["time"] == c1["system_time"] + c2["system_time"]
        assert w.call("get_block_steps")[1]["block_steps"] == c1["block_steps"] + c2["block_steps"]
        assert (w.call("get_total_steps")[1]["total_steps"]
                == c1["particle_steps"] + c2["particle_steps"])
        k, u = d2.energy()
        assert w.call("get_kinetic_energy")[1]["kinetic_energy"] == k
        assert w.call("get_potential_energy")[1]["potential_energy"] == u
        r, ts = w.call("get_particle_timestep", ids2)
        assert np.array_equal(ts["timestep"], d2.timesteps()[1])
        d2.close()
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
