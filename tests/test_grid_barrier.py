"""The persistent ensemble kernel's grid barrier (amuse_amd/csrc/g6_gridbar.h),
its counting code compiled for the CPU with one thread per workgroup
(tests/gridbar_check.cpp), plain and under ThreadSanitizer: every workgroup
sees every other's writes of the phase after each barrier, the counters end
at zero and the generation at the number of barriers, for grids that do and
do not fill the last group of 32; and with a workgroup that never arrives,
every other one gives up after the spin limit, raises the abort word and
returns false."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "gridbar_check.cpp")


def build(tmp_path, flags):
    exe = os.path.join(str(tmp_path), "gridbar_check")
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-pthread", *flags, SRC, "-o", exe],
                   check=True, capture_output=True, text=True)
    return exe


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
@pytest.mark.parametrize("flags", [[], ["-fsanitize=thread"]], ids=["plain", "tsan"])
def test_grid_barrier_counting(tmp_path, flags):
    exe = build(tmp_path, flags)
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1 exitcode=66")
    res = subprocess.run([exe], capture_output=True, text=True, timeout=600, env=env)
    assert res.returncode == 0, res.stdout + res.stderr[-3000:]
    assert res.stdout.count("ok") == 9, res.stdout
