"""Seven bodies on a line (src/amuse_ph4/tests/test_ph4.py:970-1004): unit
line of m = 0.1 bodies, eps2 = 1e-10, ph4's default eta = 0.14; bodies 1
and 2 meet head on at t = 1/4, so the smallest step falls to 2^-20 and
returns to 2^-10.  Output at t = 0, 0.23, 0.24, 0.26, 0.27.

Known answers from that test: the smallest particle step at those times is
2^-10, 2^-11, 2^-20, 2^-10, 2^-10, and the energy at t = 0 equals the
unsoftened direct sum to 10 digits.

Golden run (tests/golden/ph4_line7.npz): the reference ph4 integrator built
unchanged with -DGPU, on the CPU oracle behind the GRAPE-6 ABI.  It meets the
known step answers; its energy changes between outputs are 1.08e-7, 3.46e-5,
2.59e-6, 1.56e-8 relative (the test's 7/4/5/5-digit bounds hold except the
first, 1.08e-7 against 1e-7, for ph4 itself too).

* CPU: our driver (amuse_amd/csrc/hermite4.cpp) on the oracle reproduces the
  golden run bit for bit;
* GPU: our host driver on libg6.so and the device-resident driver give the
  same step answers and block counts, energies within 1e-9 relative.
"""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import print_case

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden", "ph4_line7.npz")
ORACLE_ABI = os.path.join(ROOT, "oracle", "libg6oracle.so")
G6 = os.path.join(ROOT, "amuse_amd", "lib", "libg6.so")
TOUTS = [0.0, 0.23, 0.24, 0.26, 0.27]
# test_ph4.py:992-993
KNOWN_DTMIN = [0.5 ** 10, 0.5 ** 11, 0.5 ** 20, 0.5 ** 10, 0.5 ** 10]


def line7():
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from ph4_run import line7 as f
    return f()


def own(lib):
    res = subprocess.run([sys.executable, os.path.join(ROOT, "tests", "ph4_run.py"),
                          "--integrator", "own", "--lib", lib, "--ic", "line7", "--eps2", "1e-10",
                          "--eta", "0.14", "--touts", ",".join(str(t) for t in TOUTS)],
                         capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stderr[-2000:]
    line = [ln for ln in res.stdout.splitlines() if ln.startswith("RESULT ")][-1]
    return json.loads(line[7:])


def test_golden_meets_known_answers():
    g = np.load(GOLD)
    assert np.allclose(g["t"], TOUTS)
    assert g["dtmin"].tolist() == KNOWN_DTMIN
    m, x, v = line7()
    k = 0.5 * (m * (v * v).sum(1)).sum()
    u = 0.0
    for i in range(m.size):
        for j in range(i + 1, m.size):
            u -= m[i] * m[j] / np.sqrt(((x[i] - x[j]) ** 2).sum())
    assert abs(g["E"][0] - (k + u)) < 1e-10 * abs(k + u)
    rel = np.abs(np.diff(g["E"])) / np.abs(g["E"][:-1])
    assert (rel[1:] < [1e-4, 1e-5, 1e-5]).all()
    assert rel[0] < 1.1e-7


def test_own_driver_on_oracle_equals_reference_ph4():
    g = np.load(GOLD)
    r = own(ORACLE_ABI)
    assert r["blocks"] == g["blocks"].tolist() and r["steps"] == g["steps"].tolist()
    assert r["dtmin"] == KNOWN_DTMIN
    assert r["E"] == g["E"].tolist()


@pytest.mark.gpu
def test_host_driver_on_libg6():
    g = np.load(GOLD)
    print_case("line7 host driver on libg6.so")
    r = own(G6)
    print("blocks", r["blocks"], "E", r["E"], flush=True)
    assert r["dtmin"] == KNOWN_DTMIN
    assert r["blocks"] == g["blocks"].tolist() and r["steps"] == g["steps"].tolist()
    assert np.allclose(r["E"], g["E"], rtol=1e-9, atol=0)


@pytest.mark.gpu
def test_device_driver():
    from amuse_amd.blocks import DeviceHermite4
    g = np.load(GOLD)
    m, x, v = line7()
    print_case("line7 device driver")
    d = DeviceHermite4(precision=0, steps_per_graph=64)
    d.commit(m, x, v, eps2=1e-10, eta=0.14)
    E, B, S, DT = [], [], [], []
    for t in TOUTS:
        assert d.evolve(t)
        assert d.time >= t
        E.append(sum(d.energy()))
        c = d.counters()
        B.append(c["block_steps"])
        S.append(c["particle_steps"])
        DT.append(float(d.timesteps()[1].min()))
    d.close()
    print("blocks", B, "E", E, flush=True)
    assert DT == KNOWN_DTMIN
    assert B == g["blocks"].tolist() and S == g["steps"].tolist()
    assert np.allclose(E, g["E"], rtol=1e-9, atol=0)
