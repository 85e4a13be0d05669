"""libg6.so loads and exports every function include/*.h declares; the C
by-value names are weak (ph4's grape.cc defines them strongly); error paths
that need no GPU return negative codes instead of exiting."""
import os
import re
import subprocess

import pytest

from amuse_amd import g6 as g6mod

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", h) for h in ("g6lib.h", "g6x.h")]


def declared_functions(path):
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    names = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", text)
    skip = {"if", "defined", "extern"}
    return sorted({n for n in names if n not in skip and (n.startswith("g6") or n in (
        "get_device_count", "force_j_particle_send", "get_j_part_data"))})


def dynamic_symbols():
    out = subprocess.run(["nm", "-D", "--defined-only", g6mod.LIBPATH], capture_output=True,
                         text=True, check=True).stdout
    syms = {}
    for line in out.splitlines():
        parts = line.split()
        if len(parts) == 3:
            syms[parts[2]] = parts[1]
    return syms


def test_library_built():
    assert os.path.exists(g6mod.LIBPATH), "run amuse_amd.build"


@pytest.mark.parametrize("header", HEADERS)
def test_every_declared_function_is_exported(header):
    syms = dynamic_symbols()
    missing = [f for f in declared_functions(header) if f not in syms]
    assert not missing, missing


def test_by_value_names_are_weak_and_fortran_names_strong():
    syms = dynamic_symbols()
    for name in ("g6_open", "g6_close", "g6_npipes", "g6_set_ti", "g6_set_j_particle",
                 "g6calc_firsthalf", "g6calc_lasthalf", "g6calc_lasthalf2", "get_device_count"):
        assert syms[name] == "W", (name, syms[name])
    for name in ("g6_open_", "g6calc_firsthalf_", "g6calc_lasthalf2_", "g6_set_j_particle_"):
        assert syms[name] == "T", (name, syms[name])


def test_python_binding_lists_match_exports():
    syms = dynamic_symbols()
    for name in g6mod.ABI_SYMBOLS + g6mod.EXT_SYMBOLS:
        assert name in syms, name


def test_no_reference_or_oracle_dependency():
    out = subprocess.run(["ldd", g6mod.LIBPATH], capture_output=True, text=True).stdout
    assert "oracle" not in out and "g6ref" not in out
    nm = subprocess.run(["nm", "-D", "--undefined-only", g6mod.LIBPATH], capture_output=True,
                        text=True).stdout
    assert "orc_" not in nm


def test_calls_before_open_fail_cleanly_without_gpu():
    lib = g6mod.load_library()
    if lib.get_device_count() > 0:
        pytest.skip("a GPU is visible; covered by the gpu tests")
    assert lib.g6_open(0) < 0
    assert lib.g6_set_ti(0, 1.0) == -2           # G6X_ERR_NOT_OPEN
    assert b"not open" in lib.g6x_last_error()
    assert lib.g6x_set_precision(7) == -3
    assert lib.g6x_set_tiles(3, 1, 0, 0) == -3
    assert lib.g6x_set_tiles(2, 1, 0, 1) == -3


def test_device_driver_rejects_bad_handles_without_gpu():
    """g6d_* entry points check their handle and arguments before touching
    the GPU; creation without a GPU returns NULL instead of exiting."""
    import ctypes as C
    import numpy as np
    from amuse_amd.blocks import _bind
    lib = _bind(C.CDLL(g6mod.LIBPATH))
    k, u = C.c_double(), C.c_double()
    assert lib.g6d_energy(None, C.byref(k), C.byref(u)) == -3
    assert lib.g6d_run(None, 1.0, 0, 8, None) == -3
    assert lib.g6d_set_precision(None, 0) == -3
    assert lib.g6d_last_error(None) == b"null driver handle"
    lib.g6d_free(None)
    if not os.path.exists("/dev/kfd"):
        m = np.ones(4) / 4
        x = np.eye(4, 3)
        D = C.POINTER(C.c_double)
        h = lib.g6d_create(4, None, m.ctypes.data_as(D), x.ctypes.data_as(D),
                           x.ctypes.data_as(D), 1e-4, 0.01, 0.0625, 0.03125, 8.0, 0, 1, 0)
        assert h is None


def test_tuning_file_round_trip(tmp_path):
    """g6x_load_tuning keeps the well-formed rows of supported shapes and
    drops the rest; g6x_save_tuning writes back what was kept.  Neither
    needs a GPU (the table is host state)."""
    lib = g6mod.load_library()
    src = tmp_path / "tune.txt"
    src.write_text("# ni_lo ni_hi njs ipl nchunk smem\n"
                   "1 64 1 2 0 1\n"         # kept
                   "65 4096 4 1 3 0\n"      # kept
                   "1 2 3 1 0 0\n"          # no shape with 3 lanes per i
                   "1 2 2 1 0 1\n"          # LDS-resident j needs one lane per i
                   "10 5 1 1 0 0\n"         # empty ni range
                   "1 2 1 1 -1 0\n"         # negative chunk count
                   "garbage\n")
    empty = tmp_path / "empty.txt"
    empty.write_text("")
    try:
        assert lib.g6x_load_tuning(str(src).encode()) == 2
        out = tmp_path / "out.txt"
        assert lib.g6x_save_tuning(str(out).encode()) == 2
        rows = [ln.split() for ln in out.read_text().splitlines() if not ln.startswith("#")]
        assert rows == [["1", "64", "1", "2", "0", "1"], ["65", "4096", "4", "1", "3", "0"]]
        assert lib.g6x_load_tuning(str(tmp_path / "missing.txt").encode()) == -3
        assert lib.g6x_save_tuning(str(tmp_path / "no" / "such" / "dir.txt").encode()) == -3
    finally:
        assert lib.g6x_load_tuning(str(empty).encode()) == 0


def test_ensemble_rejects_bad_counts():
    """Cluster sizes are checked before anything reaches the library."""
    import numpy as np
    from amuse_amd.ensemble import Ensemble
    m = np.ones(4) / 4
    x = np.zeros((4, 3))
    for counts in ([2, 1], [2, 0, 2], [[2, 2]], [5, -1]):
        with pytest.raises(ValueError):
            Ensemble(counts, m, x, x)
