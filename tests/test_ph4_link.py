"""An unchanged GRAPE-6 consumer links against libg6.so.

The reference ph4 core (src/amuse_ph4/src, the objects of libph4gpu.a,
built with -DGPU exactly as src/amuse_ph4/src/Makefile:8-27 does, plus
-DNOMPI) is compiled from its sources where they lie by `make -C oracle
ph4link` and linked into a shared object against amuse_amd/lib/libg6.so with
--no-undefined.  Every GRAPE-6 entry point ph4 calls must resolve from
libg6.so, and ph4's own by-value wrappers (grape.cc:8-146) must win over the
library's weak ones.  Runs only where the reference checkout is present.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference/src/amuse_ph4/src/gpu.cc"
OUT = os.path.join(ROOT, "oracle", "_ref", "libph4gpu_link.so")

pytestmark = pytest.mark.skipif(not os.path.exists(REF), reason="reference ph4 not present")


@pytest.fixture(scope="module")
def linked():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ph4link"], check=True,
                   capture_output=True)
    assert os.path.exists(OUT)
    return OUT


def _nm(path, *flags):
    out = subprocess.run(["nm", "-D", *flags, path], capture_output=True, text=True,
                         check=True).stdout
    syms = {}
    for line in out.splitlines():
        parts = line.split()
        syms[parts[-1]] = parts[-2]
    return syms


def test_fortran_entry_points_resolve_from_libg6(linked):
    und = _nm(linked, "--undefined-only")
    g6 = _nm(os.path.join(ROOT, "amuse_amd", "lib", "libg6.so"), "--defined-only")
    needed = sorted(s for s in und if s.startswith("g6"))
    assert {"g6_open_", "g6_set_j_particle_", "g6calc_firsthalf_", "g6calc_lasthalf_",
            "g6calc_lasthalf2_", "g6_set_ti_", "g6_npipes_"} <= set(needed)
    for s in needed:
        assert s in g6, s


def test_ph4_wrappers_win_over_weak_library_names(linked):
    own = _nm(linked, "--defined-only")
    for s in ("g6_open", "g6_set_j_particle", "g6calc_firsthalf", "g6calc_lasthalf2"):
        assert own.get(s) == "T", s


def test_dynamic_dependency_is_libg6(linked):
    out = subprocess.run(["ldd", linked], capture_output=True, text=True).stdout
    assert "libg6.so => " + os.path.join(ROOT, "amuse_amd", "lib", "libg6.so") in out
