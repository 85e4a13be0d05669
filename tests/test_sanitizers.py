"""The CPU code under AddressSanitizer and UndefinedBehaviorSanitizer.

The oracle (oracle/g6_oracle.c, g6_oracle_abi.c, cpu_best.c), the host
block-step driver (amuse_amd/csrc/hermite4.cpp) and the socket worker
(amuse_amd/csrc/g6_worker.cpp, which parses headers and lengths sent over
a socket) are built with -fsanitize=address,undefined
(tests/sanitize/Makefile, host code only) into a copy of the package tree.
The CPU tests that drive them then run in that copy, in a child process
with the ASan runtime preloaded (the interpreter is not instrumented): the
oracle against its fixtures and closed forms, the host driver's
trajectories and block-step invariants, and the worker protocol including
malformed and oversized headers.  A sanitizer report stops the child, so
the test passes only when all of those pass with no report.
"""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CHILD_TESTS = ["test_oracle.py", "test_oracle_ngb.py", "test_oracle_ensemble.py",
               "test_hermite_cpu.py", "test_worker_protocol.py"]


def _gcc_runtime(name):
    r = subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True, text=True)
    p = r.stdout.strip()
    return p if os.path.isabs(p) and os.path.exists(p) else None


def _copy_tree(dst):
    skip = shutil.ignore_patterns("*.so", "__pycache__", "_ref", "lib", "csrc", "*.tmp",
                                  "sanitize")
    for d in ("amuse_amd", "oracle", "tests"):
        shutil.copytree(os.path.join(ROOT, d), os.path.join(dst, d), ignore=skip)
    shutil.copy2(os.path.join(ROOT, "pytest.ini"), dst)
    os.makedirs(os.path.join(dst, "amuse_amd", "lib"))
    # the GPU library (not instrumented) that the worker links
    shutil.copy2(os.path.join(ROOT, "amuse_amd", "lib", "libg6.so"),
                 os.path.join(dst, "amuse_amd", "lib"))
    ref = os.path.join(ROOT, "oracle", "_ref", "libg6ref.so")
    if os.path.exists(ref):    # the compiled reference g6lib, as built
        os.makedirs(os.path.join(dst, "oracle", "_ref"))
        shutil.copy2(ref, os.path.join(dst, "oracle", "_ref"))


def test_cpu_code_under_asan_and_ubsan(tmp_path):
    asan = _gcc_runtime("libasan.so")
    if asan is None:
        pytest.skip("gcc has no AddressSanitizer runtime")
    if not os.path.exists(os.path.join(ROOT, "amuse_amd", "lib", "libg6.so")):
        pytest.fail("amuse_amd/lib/libg6.so is not built")
    tree = str(tmp_path / "tree")
    _copy_tree(tree)
    jobs = str(max(1, min(8, len(os.sched_getaffinity(0)))))
    r = subprocess.run(["make", "-s", "-j", jobs, "-C", os.path.join(ROOT, "tests", "sanitize"),
                        f"OUT={tree}"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    env = dict(os.environ, LD_PRELOAD=asan,
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    env.pop("PYTHONPATH", None)
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-m", "not gpu", "-p",
                        "no:cacheprovider", *[os.path.join("tests", t) for t in CHILD_TESTS]],
                       cwd=tree, env=env, capture_output=True, text=True, timeout=1200)
    out = r.stdout + r.stderr
    assert "AddressSanitizer" not in out and "runtime error:" not in out, out[-6000:]
    assert r.returncode == 0, out[-6000:]
    assert " passed" in r.stdout and " failed" not in r.stdout, r.stdout[-2000:]
    # the child ran the instrumented builds (none was rebuilt plain on the way)
    for rel in ("oracle/liboracle.so", "oracle/libg6oracle.so", "oracle/libcpubest.so",
                "amuse_amd/lib/libhermite4.so", "amuse_amd/lib/g6_worker"):
        syms = subprocess.run(["nm", "-D", os.path.join(tree, rel)], capture_output=True,
                              text=True).stdout
        assert "__asan_report" in syms and "__ubsan_handle" in syms, rel
