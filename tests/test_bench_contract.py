"""The benchmark's own contract, without a GPU: more than one GPU is refused
with a message and a non-zero status for every workload (one GPU is what
this build measures), and the host details the CPU baselines record are
read from the host they run on."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("extra", [[], ["--workload", "blocks"], ["--workload", "clusters"]])
def test_more_than_one_gpu_is_refused(extra):
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2", *extra],
                       capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_host_details_and_thread_count():
    sys.path.insert(0, ROOT)
    import bench
    info = bench.host_info()
    for key in ("nproc", "affinity", "cgroup_cpu_quota", "cpu_model", "cpu_mhz_now_mean",
                "cpu_mhz_max"):
        assert key in info
    assert info["affinity"] == len(os.sched_getaffinity(0))
    n = bench.usable_cores(info)
    assert 1 <= n <= info["affinity"]
    # a cgroup quota below the affinity mask caps the threads; none leaves the mask
    assert bench.usable_cores({"affinity": 16, "cgroup_cpu_quota": 4.5}) == 4
    assert bench.usable_cores({"affinity": 16, "cgroup_cpu_quota": None}) == 16
    assert bench.usable_cores({"affinity": 2, "cgroup_cpu_quota": 0.5}) == 1
