"""Python view of the GRAPE-6 C ABI exported by amuse_amd/lib/libg6.so.

The method names and argument meanings follow the reference ABI
(lib/g6lib/g6lib.h:8-128, src/amuse_ph4/src/grape.h:6-120): ``open``,
``set_ti``, ``set_j_particle``, ``calc_firsthalf``, ``calc_lasthalf``,
``calc_lasthalf2``, ``read_neighbour_list``, ``get_neighbour_list``.
Arrays are numpy; [n][3] vectors are C-contiguous float64 (n, 3).

There is no CPU fallback: constructing :class:`G6` without the built library
raises, and every call returning a negative code raises :class:`G6Error`
with the library's message.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "lib", "libg6.so")

FP64 = 0
MIXED = 1
NGB_CAP = 256   # neighbour ids kept per i-particle

_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)

# Every exported symbol, for load checks.
ABI_SYMBOLS = [
    "g6_open_", "g6_close_", "g6_npipes_", "g6_set_tunit_", "g6_set_xunit_", "g6_set_ti_",
    "g6_set_j_particle_", "g6calc_firsthalf_", "g6calc_lasthalf_", "g6calc_lasthalf2_",
    "g6_initialize_jp_buffer_", "g6_flush_jp_buffer_", "g6_reset_", "g6_reset_fofpga_",
    "g6_read_neighbour_list_", "g6_get_neighbour_list_",
    "g6_open", "g6_close", "g6_npipes", "g6_set_tunit", "g6_set_xunit", "g6_set_ti",
    "g6_set_j_particle", "g6calc_firsthalf", "g6calc_lasthalf", "g6calc_lasthalf2",
    "g6_initialize_jp_buffer", "g6_flush_jp_buffer", "g6_reset", "g6_reset_fofpga",
    "g6_read_neighbour_list", "g6_get_neighbour_list", "g6_set_neighbour_list_sort_mode",
    "g6_get_neighbour_list_sort_mode", "g6_set_overflow_flag_test_mode",
    "g6_get_number_of_pipelines", "g6_reinitialize", "get_device_count",
    "force_j_particle_send", "get_j_part_data",
]
EXT_SYMBOLS = [
    "g6x_set_precision", "g6x_get_precision", "g6x_set_predictor_compat", "g6x_set_npipes",
    "g6x_last_error", "g6x_set_j_particles", "g6x_calc_self", "g6x_fetch", "g6x_sync",
    "g6x_timing", "g6x_get_timing", "g6x_set_tiles", "g6x_autotune", "g6x_save_tuning",
    "g6x_load_tuning", "g6x_set_spatial",
]


class G6Error(RuntimeError):
    pass


def _share_hip_runtime() -> None:
    """Imports torch, when it is installed, before libg6.so is loaded.

    PyTorch-ROCm ships its own libamdhip64.so under the same soname as
    ROCm's.  Whichever process loads first provides the HIP runtime to both:
    torch first, and libg6.so binds to torch's copy and both see the GPU;
    libg6.so first, and torch later finds no GPU through ROCm's copy.  A
    process that uses only libg6.so (no torch installed) is unaffected."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def load_library(path: str = LIBPATH) -> C.CDLL:
    if not os.path.exists(path):
        raise G6Error(f"{path} is not built; run amuse_amd.build (hipcc, gfx950)")
    _share_hip_runtime()
    lib = C.CDLL(path)
    v = C.c_void_p
    sig = {
        "g6_open": (C.c_int, [C.c_int]),
        "g6_close": (C.c_int, [C.c_int]),
        "g6_npipes": (C.c_int, []),
        "g6_set_ti": (C.c_int, [C.c_int, C.c_double]),
        "g6_set_j_particle": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                        C.c_double, _D, _D, _D, _D, _D]),
        "g6calc_firsthalf": (None, [C.c_int, C.c_int, C.c_int, _I, _D, _D, _D, _D, _D,
                                    C.c_double, _D]),
        "g6calc_lasthalf": (C.c_int, [C.c_int, C.c_int, C.c_int, _I, _D, _D, C.c_double, _D,
                                      _D, _D, _D]),
        "g6calc_lasthalf2": (C.c_int, [C.c_int, C.c_int, C.c_int, _I, _D, _D, C.c_double, _D,
                                       _D, _D, _D, _I]),
        "g6_reset": (C.c_int, [C.c_int]),
        "g6_read_neighbour_list": (C.c_int, [C.c_int]),
        "g6_get_neighbour_list": (C.c_int, [C.c_int, C.c_int, C.c_int, _I, _I]),
        "get_device_count": (C.c_int, []),
        "g6x_set_precision": (C.c_int, [C.c_int]),
        "g6x_get_precision": (C.c_int, []),
        "g6x_set_predictor_compat": (C.c_int, [C.c_int]),
        "g6x_set_npipes": (C.c_int, [C.c_int]),
        "g6x_last_error": (C.c_char_p, []),
        "g6x_set_j_particles": (C.c_int, [C.c_int, _I, _I, _D, _D, _D, _D, _D, _D, _D, _D]),
        "g6x_calc_self": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_double, C.c_int]),
        "g6x_fetch": (C.c_int, [C.c_int, _D, _D, _D, _I]),
        "g6x_calc_self_noalloc": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_double, v, v, v, v]),
        "g6x_sync": (C.c_int, []),
        "g6x_set_spatial": (C.c_int, [C.c_int]),
        "g6x_timing": (C.c_int, [C.c_int]),
        "g6x_get_timing": (C.c_int, [_D, _D, _D, C.POINTER(C.c_longlong), _D]),
        "g6x_set_tiles": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int]),
        "g6x_autotune": (C.c_int, [C.c_int, C.c_int, C.c_double]),
        "g6x_save_tuning": (C.c_int, [C.c_char_p]),
        "g6x_load_tuning": (C.c_int, [C.c_char_p]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    del v
    return lib


def _d(a: np.ndarray):
    return a.ctypes.data_as(_D)


def _i(a: np.ndarray):
    return a.ctypes.data_as(_I)


def _vec(a, n: int) -> np.ndarray:
    out = np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(n, 3))
    return out


class G6:
    """One GRAPE-6 unit (the library keeps a single unit per process)."""

    def __init__(self, cluster_id: int = 0, path: str = LIBPATH):
        self.lib = load_library(path)
        self.cluster = cluster_id
        self.is_open = False

    # -- helpers --------------------------------------------------------
    def _check(self, rc: int, what: str) -> int:
        if rc is not None and rc < 0:
            msg = self.lib.g6x_last_error().decode()
            raise G6Error(f"{what} failed ({rc}): {msg}")
        return rc

    # -- lifecycle (g6lib.c:190-258) ------------------------------------
    def open(self) -> None:
        self._check(self.lib.g6_open(self.cluster), "g6_open")
        self.is_open = True

    def close(self) -> None:
        self._check(self.lib.g6_close(self.cluster), "g6_close")
        self.is_open = False

    def npipes(self) -> int:
        return self.lib.g6_npipes()

    def reset(self) -> None:
        self._check(self.lib.g6_reset(self.cluster), "g6_reset")

    def set_ti(self, ti: float) -> None:
        self._check(self.lib.g6_set_ti(self.cluster, float(ti)), "g6_set_ti")

    # -- j-store (g6lib.c:260-315) --------------------------------------
    def set_j_particle(self, address, index, tj, dtj, mass, k18, j6, a2, v, x) -> None:
        arrs = [np.ascontiguousarray(np.asarray(q, dtype=np.float64).reshape(3))
                for q in (k18, j6, a2, v, x)]
        self._check(self.lib.g6_set_j_particle(self.cluster, int(address), int(index), float(tj),
                                               float(dtj), float(mass), *[_d(q) for q in arrs]),
                    "g6_set_j_particle")

    def set_j_particles(self, address, index, tj, dtj, mass, k18, j6, a2, v, x) -> None:
        addr = np.ascontiguousarray(address, dtype=np.int32)
        n = addr.size
        idx = np.ascontiguousarray(index, dtype=np.int32).reshape(n)
        sc = [np.ascontiguousarray(np.broadcast_to(np.asarray(q, dtype=np.float64), (n,)))
              for q in (tj, dtj, mass)]
        vs = [_vec(np.broadcast_to(np.asarray(q, dtype=np.float64), (n, 3)), n)
              for q in (k18, j6, a2, v, x)]
        self._check(self.lib.g6x_set_j_particles(n, _i(addr), _i(idx), *[_d(q) for q in sc],
                                                 *[_d(q) for q in vs]), "g6x_set_j_particles")

    # -- calc (g6lib.c:317-441) -----------------------------------------
    def calc_firsthalf(self, nj, index, xi, vi, eps2, h2=None) -> None:
        idx = np.ascontiguousarray(index, dtype=np.int32)
        ni = idx.size
        xi = _vec(xi, ni)
        vi = _vec(vi, ni)
        # one radius per i (reshape raises on a wrong length: the library
        # reads ni of them)
        h2 = np.zeros(ni) if h2 is None else np.ascontiguousarray(h2, dtype=np.float64).reshape(ni)
        dummy = np.zeros((max(ni, 1), 3))
        self._last_i = (idx, xi, vi, h2)
        self.lib.g6calc_firsthalf(self.cluster, int(nj), ni, _i(idx), _d(xi), _d(vi), _d(dummy),
                                  _d(dummy), _d(dummy[:, 0].copy()), float(eps2), _d(h2))
        err = self.lib.g6x_last_error().decode()
        self._first_err = err

    def _lasthalf(self, nj, ni, eps2, nn: bool):
        idx, xi, vi, h2 = self._last_i
        idx = idx.copy()
        acc = np.zeros((ni, 3))
        jerk = np.zeros((ni, 3))
        pot = np.zeros(ni)
        if nn:
            nnb = np.zeros(ni, dtype=np.int32)
            rc = self.lib.g6calc_lasthalf2(self.cluster, int(nj), ni, _i(idx), _d(xi), _d(vi),
                                           float(eps2), _d(h2), _d(acc), _d(jerk), _d(pot),
                                           _i(nnb))
            self._check(rc, "g6calc_lasthalf2")
            return acc, jerk, pot, nnb, idx
        rc = self.lib.g6calc_lasthalf(self.cluster, int(nj), ni, _i(idx), _d(xi), _d(vi),
                                      float(eps2), _d(h2), _d(acc), _d(jerk), _d(pot))
        self._check(rc, "g6calc_lasthalf")
        return acc, jerk, pot, idx

    def calc_lasthalf(self, nj, ni, eps2):
        """Returns acc (ni,3), jerk (ni,3), pot (ni,), written-back index."""
        return self._lasthalf(nj, ni, eps2, False)

    def calc_lasthalf2(self, nj, ni, eps2):
        """Returns acc, jerk, pot, nn ids, written-back index."""
        return self._lasthalf(nj, ni, eps2, True)

    def calc(self, nj, index, xi, vi, eps2, nn=True, h2=None):
        """firsthalf + lasthalf[2] in one go."""
        idx = np.ascontiguousarray(index, dtype=np.int32)
        self.calc_firsthalf(nj, idx, xi, vi, eps2, h2)
        return self._lasthalf(nj, idx.size, eps2, nn)

    # -- neighbour lists (sapporo.cpp:234-272) ----------------------------
    def read_neighbour_list(self) -> int:
        return self._check(self.lib.g6_read_neighbour_list(self.cluster), "g6_read_neighbour_list")

    def get_neighbour_list(self, ipipe: int, maxlength: int):
        n = C.c_int(0)
        buf = np.zeros(max(maxlength, 1), dtype=np.int32)
        rc = self.lib.g6_get_neighbour_list(self.cluster, int(ipipe), int(maxlength),
                                            C.byref(n), _i(buf))
        self._check(rc, "g6_get_neighbour_list")
        # ids copied: the full count, capped by the 256 kept and by maxlength
        return rc, n.value, buf[:min(n.value, NGB_CAP, maxlength)].copy()

    # -- extensions (include/g6x.h) ---------------------------------------
    def set_precision(self, mode: int) -> None:
        self._check(self.lib.g6x_set_precision(int(mode)), "g6x_set_precision")

    def set_predictor_compat(self, on: bool) -> None:
        self._check(self.lib.g6x_set_predictor_compat(int(bool(on))), "g6x_set_predictor_compat")

    def set_npipes(self, n: int) -> None:
        self._check(self.lib.g6x_set_npipes(int(n)), "g6x_set_npipes")

    def set_tiles(self, njs: int, ipl: int, nchunk: int = 0, smem: bool = False) -> None:
        self._check(self.lib.g6x_set_tiles(int(njs), int(ipl), int(nchunk), int(bool(smem))),
                    "g6x_set_tiles")

    def calc_self(self, nj: int, i0: int, ni: int, eps2: float, nn: bool = True) -> None:
        self._check(self.lib.g6x_calc_self(int(nj), int(i0), int(ni), float(eps2), int(nn)),
                    "g6x_calc_self")

    def calc_self_noalloc(self, nj: int, i0: int, ni: int, eps2: float, acc: int, jerk: int,
                          pot: int, nn: int | None) -> None:
        """g6x_calc_self_noalloc: acc, jerk, pot, nn are device addresses
        (e.g. torch tensors' data_ptr()) of [ni][3], [ni][3], [ni] float64
        and [ni] int32 arrays; nn None for no nearest neighbour.  Runs on the
        unit's stream: sync() before reading the arrays."""
        self._check(self.lib.g6x_calc_self_noalloc(int(nj), int(i0), int(ni), float(eps2), acc,
                                                   jerk, pot, nn), "g6x_calc_self_noalloc")

    def fetch(self, ni: int, nn: bool = True):
        acc = np.zeros((ni, 3))
        jerk = np.zeros((ni, 3))
        pot = np.zeros(ni)
        nnb = np.zeros(ni, dtype=np.int32) if nn else None
        self._check(self.lib.g6x_fetch(int(ni), _d(acc), _d(jerk), _d(pot),
                                       _i(nnb) if nn else None), "g6x_fetch")
        return acc, jerk, pot, nnb

    def set_spatial(self, on: bool) -> None:
        self._check(self.lib.g6x_set_spatial(int(bool(on))), "g6x_set_spatial")

    def sync(self) -> None:
        self._check(self.lib.g6x_sync(), "g6x_sync")

    def timing(self, enable: bool) -> None:
        self.lib.g6x_timing(int(bool(enable)))

    def get_timing(self) -> dict:
        f, r, p, pairs = C.c_double(), C.c_double(), C.c_double(), C.c_double()
        n = C.c_longlong()
        self.lib.g6x_get_timing(C.byref(f), C.byref(r), C.byref(p), C.byref(n), C.byref(pairs))
        return {"force_ms": f.value, "reduce_ms": r.value, "predict_ms": p.value,
                "force_launches": n.value, "pairs": pairs.value}

    def autotune(self, nj: int, ni_max: int, eps2: float) -> int:
        return self._check(self.lib.g6x_autotune(int(nj), int(ni_max), float(eps2)), "g6x_autotune")

    def save_tuning(self, path: str) -> int:
        return self._check(self.lib.g6x_save_tuning(path.encode()), "g6x_save_tuning")

    def load_tuning(self, path: str) -> int:
        return self._check(self.lib.g6x_load_tuning(path.encode()), "g6x_load_tuning")
