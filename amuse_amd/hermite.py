"""Block-timestep Hermite-4 integrator (ph4's GPU code path) over a GRAPE-6
library, Python view of amuse_amd/lib/libhermite4.so.

The force library is bound at run time by path: amuse_amd/lib/libg6.so (the
MI355X library) by default; any library exporting the GRAPE-6 Fortran entry
points works.  Parameters follow ph4 (src/amuse_ph4/interface.py:506-628,
interface.cc:29-31): eta (timestep_parameter), eps2, and the initial-step
factors fac = 0.0625, limit = 0.03125, median limit = 8.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .g6 import LIBPATH as G6_LIBPATH

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "lib", "libhermite4.so")

_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)
_L = C.POINTER(C.c_longlong)


def _lib():
    if not os.path.exists(LIBPATH):
        raise RuntimeError(f"{LIBPATH} is not built; run amuse_amd.build")
    from .g6 import _share_hip_runtime
    _share_hip_runtime()  # h4_new dlopens the GRAPE-6 library
    L = C.CDLL(LIBPATH)
    L.h4_new.restype = C.c_void_p
    L.h4_new.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int]
    L.h4_commit.argtypes = [C.c_void_p, C.c_int, _I, _D, _D, _D, C.c_double, C.c_double,
                            C.c_double, C.c_double, C.c_double]
    L.h4_evolve.argtypes = [C.c_void_p, C.c_double, C.c_longlong]
    L.h4_steps.argtypes = [C.c_void_p, C.c_longlong]
    L.h4_energy.argtypes = [C.c_void_p, _D, _D]
    L.h4_get_state.argtypes = [C.c_void_p, _D, _D]
    L.h4_get_steps.argtypes = [C.c_void_p, _D, _D]
    L.h4_set_precision.argtypes = [C.c_void_p, C.c_int]
    L.h4_set_ph4_initial_load.argtypes = [C.c_void_p, C.c_int]
    L.h4_time.argtypes = [C.c_void_p]
    L.h4_time.restype = C.c_double
    L.h4_counters.argtypes = [C.c_void_p, _L, _L, _L]
    L.h4_free.argtypes = [C.c_void_p]
    L.h4_free.restype = None
    return L


class Hermite4:
    def __init__(self, g6_library: str = G6_LIBPATH, npipes: int = 0, precision: int = 0,
                 compat: bool = False, cluster: int = 0, ph4_initial_load: bool = True):
        self.L = _lib()
        self.h = self.L.h4_new(g6_library.encode(), int(cluster), int(npipes), int(precision),
                               int(bool(compat)))
        if not self.h:
            raise RuntimeError(f"could not open GRAPE-6 library {g6_library}")
        # ph4 leaves the j-set with acc = jerk = 0 after the initial force
        # pass (idata.cc:66-83); False reloads it with the initial values
        self.L.h4_set_ph4_initial_load(self.h, int(bool(ph4_initial_load)))
        self.n = 0

    def commit(self, m, x, v, eps2, eta=0.14, ids=None, dt_fac=0.0625, dt_limit=0.03125,
               dt_limitm=8.0):
        m = np.ascontiguousarray(m, dtype=np.float64)
        n = m.size
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(n, 3)
        v = np.ascontiguousarray(v, dtype=np.float64).reshape(n, 3)
        ids = np.arange(1, n + 1, dtype=np.int32) if ids is None else np.ascontiguousarray(
            ids, dtype=np.int32)
        rc = self.L.h4_commit(self.h, n, ids.ctypes.data_as(_I), m.ctypes.data_as(_D),
                              x.ctypes.data_as(_D), v.ctypes.data_as(_D), float(eps2), float(eta),
                              float(dt_fac), float(dt_limit), float(dt_limitm))
        if rc < 0:
            raise RuntimeError(f"h4_commit failed ({rc})")
        self.n = n
        self.m = m

    def evolve(self, t_end: float, max_blocks: int = 0) -> bool:
        """Block steps while the system time is below t_end, as ph4's
        evolve_model (src/amuse_ph4/interface.cc:614-674; the time ends past
        t_end when t_end is not commensurate with the steps); False if
        max_blocks (> 0) ran out first."""
        rc = self.L.h4_evolve(self.h, float(t_end), int(max_blocks))
        if rc < 0:
            raise RuntimeError(f"h4_evolve failed ({rc})")
        return rc == 0

    def steps(self, nblocks: int) -> None:
        rc = self.L.h4_steps(self.h, int(nblocks))
        if rc < 0:
            raise RuntimeError(f"h4_steps failed ({rc})")

    def energy(self):
        k, u = C.c_double(), C.c_double()
        rc = self.L.h4_energy(self.h, C.byref(k), C.byref(u))
        if rc < 0:
            raise RuntimeError(f"h4_energy failed ({rc})")
        return k.value, u.value

    def state(self):
        x = np.zeros((self.n, 3))
        v = np.zeros((self.n, 3))
        self.L.h4_get_state(self.h, x.ctypes.data_as(_D), v.ctypes.data_as(_D))
        return x, v

    def timesteps(self):
        t = np.zeros(self.n)
        dt = np.zeros(self.n)
        self.L.h4_get_steps(self.h, t.ctypes.data_as(_D), dt.ctypes.data_as(_D))
        return t, dt

    def set_precision(self, precision: int) -> None:
        if self.L.h4_set_precision(self.h, int(precision)) < 0:
            raise RuntimeError("force library has no precision control")

    @property
    def time(self) -> float:
        return self.L.h4_time(self.h)

    def counters(self):
        b, p, c = C.c_longlong(), C.c_longlong(), C.c_longlong()
        self.L.h4_counters(self.h, C.byref(b), C.byref(p), C.byref(c))
        return {"block_steps": b.value, "particle_steps": p.value, "g6_calls": c.value}

    def close(self):
        if self.h:
            self.L.h4_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
