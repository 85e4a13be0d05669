"""The reference ph4 integrator, built unchanged from src/amuse_ph4/src with
-DGPU (`make -C oracle ph4run`), runs on a GRAPE-6 library of our choosing.

CPU: on the CPU oracle behind the ABI, our driver
amuse_amd/csrc/hermite4.cpp reproduces ph4 bit for bit — the same block and
particle step counts and the same energies to the last bit — which pins the
driver to the reference integrator.

GPU: ph4 itself on libg6.so (MI355X) agrees with ph4 on the CPU oracle
(energies to 1e-12 relative, identical step counts), and our driver on
libg6.so reproduces ph4 on libg6.so bit for bit.  With every particle
synchronised to the system time after each output (AMUSE's
synchronize_model = jdata::synchronize_all) at output times off the step
grid, so that the synchronising steps are real ones, the device-resident
driver (g6d_synchronize) reproduces ph4 on libg6.so bit for bit too.
"""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_DIR = os.path.join(ROOT, "oracle", "_ref")
PH4_ORACLE = os.path.join(REF_DIR, "libph4run_oracle.so")
PH4_G6 = os.path.join(REF_DIR, "libph4run_g6.so")
ORACLE_ABI = os.path.join(ROOT, "oracle", "libg6oracle.so")
G6 = os.path.join(ROOT, "amuse_amd", "lib", "libg6.so")


def run(integrator, lib, n=256, t=0.125, touts=None, sync=False):
    extra = (["--touts", touts] if touts else []) + (["--sync"] if sync else [])
    res = subprocess.run([sys.executable, os.path.join(ROOT, "tests", "ph4_run.py"),
                          "--integrator", integrator, "--lib", lib, "--n", str(n), "--t", str(t),
                          *extra], capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stderr[-2000:]
    line = [ln for ln in res.stdout.splitlines() if ln.startswith("RESULT ")][-1]
    return json.loads(line[7:])


@pytest.fixture(scope="module")
def built():
    if os.path.exists("/root/reference/src/amuse_ph4/src/gpu.cc"):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ph4run"], check=True,
                       capture_output=True)
    if not (os.path.exists(PH4_ORACLE) and os.path.exists(PH4_G6)):
        pytest.skip("reference ph4 not built (needs the reference checkout)")


def test_own_driver_equals_reference_ph4_on_oracle(built):
    a = run("ph4", PH4_ORACLE)
    b = run("own", ORACLE_ABI)
    assert a["used_library"] == 1
    assert a["blocks"] == b["blocks"] and a["steps"] == b["steps"]
    assert a["E"] == b["E"]
    assert a["K"] == b["K"]


@pytest.mark.gpu
def test_reference_ph4_on_libg6(built):
    a = run("ph4", PH4_G6)
    b = run("ph4", PH4_ORACLE)
    print("ph4 on libg6:", a, "\nph4 on oracle:", b, flush=True)
    assert a["blocks"] == b["blocks"] and a["steps"] == b["steps"]
    for ea, eb in zip(a["E"], b["E"]):
        assert abs(ea - eb) <= 1e-12 * abs(eb)
    c = run("own", G6)
    assert c["blocks"] == a["blocks"] and c["steps"] == a["steps"]
    assert c["E"] == a["E"]


@pytest.mark.gpu
def test_device_synchronize_equals_reference_ph4(built):
    touts = "0,0.011,0.023,0.05,0.0713"
    a = run("ph4", PH4_G6, touts=touts, sync=True)
    plain = run("ph4", PH4_G6, touts=touts)
    d = run("device", G6, touts=touts, sync=True)
    print("ph4 sync:", a, "\ndevice sync:", d, flush=True)
    # the synchronising steps are real: more particle steps than without
    assert a["steps"][-1] > plain["steps"][-1]
    assert d["blocks"] == a["blocks"] and d["steps"] == a["steps"]
    assert d["E"] == a["E"] and d["K"] == a["K"]
