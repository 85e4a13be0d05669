"""Device-resident block-step Hermite integrator (include/g6x.h g6d_*).

The integration loop of ph4's GPU path — next block time and block
selection (src/amuse_ph4/src/scheduler.cc:66-140), i-particle predictor
(idata.cc:341-360), force call, corrector with the Aarseth step
(idata.cc:431-512) — runs on the GPU; the host only replays captured steps
and reads a control block back.  Same interface as amuse_amd.hermite.Hermite4
(the host-side driver over the GRAPE-6 ABI), so the two are interchangeable.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .g6 import G6Error, load_library

_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)
_LL = C.POINTER(C.c_longlong)


def _bind(lib):
    sig = {
        "g6d_create": (C.c_void_p, [C.c_int, _I, _D, _D, _D, C.c_double, C.c_double, C.c_double,
                                    C.c_double, C.c_double, C.c_int, C.c_int, C.c_int]),
        "g6d_run": (C.c_longlong, [C.c_void_p, C.c_double, C.c_longlong, C.c_int, _I]),
        "g6d_energy": (C.c_int, [C.c_void_p, _D, _D]),
        "g6d_state": (C.c_int, [C.c_void_p, _D, _D]),
        "g6d_state_raw": (C.c_int, [C.c_void_p, _D, _D]),
        "g6d_steps": (C.c_int, [C.c_void_p, _D, _D]),
        "g6d_counters": (C.c_int, [C.c_void_p, _LL, _LL, _D]),
        "g6d_set_precision": (C.c_int, [C.c_void_p, C.c_int]),
        "g6d_synchronize": (C.c_int, [C.c_void_p]),
        "g6d_last_error": (C.c_char_p, [C.c_void_p]),
        "g6d_free": (None, [C.c_void_p]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


class DeviceHermite4:
    """ph4-style block-step integration with the whole step on the GPU."""

    def __init__(self, precision: int = 0, steps_per_graph: int = 64, device: int = 0,
                 ph4_initial_load: bool = True):
        self.lib = _bind(load_library())
        self.precision = int(precision)
        self.steps_per_graph = int(steps_per_graph)
        self.device = int(device)
        self.ph4_initial_load = bool(ph4_initial_load)
        self.h = None
        self.n = 0

    def commit(self, m, x, v, eps2, eta=0.14, ids=None, dt_fac=0.0625, dt_limit=0.03125,
               dt_limitm=8.0):
        if self.h:
            self.lib.g6d_free(self.h)
            self.h = None
        m = np.ascontiguousarray(m, dtype=np.float64)
        n = m.size
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(n, 3)
        v = np.ascontiguousarray(v, dtype=np.float64).reshape(n, 3)
        ids = np.arange(1, n + 1, dtype=np.int32) if ids is None else np.ascontiguousarray(
            ids, dtype=np.int32)
        self.h = self.lib.g6d_create(n, ids.ctypes.data_as(_I), m.ctypes.data_as(_D),
                                     x.ctypes.data_as(_D), v.ctypes.data_as(_D), float(eps2),
                                     float(eta), float(dt_fac), float(dt_limit), float(dt_limitm),
                                     self.precision, int(self.ph4_initial_load), self.device)
        if not self.h:
            raise G6Error("g6d_create failed (no GPU, bad sizes or a HIP error)")
        self.n = n
        self.m = m

    def _check(self, rc, what):
        if rc < 0:
            raise G6Error(f"{what} failed ({rc}): {self.lib.g6d_last_error(self.h).decode()}")
        return rc

    def evolve(self, t_end: float, max_blocks: int = 0) -> bool:
        """Block steps while the system time is below t_end, as ph4's
        evolve_model (src/amuse_ph4/interface.cc:614-674; the time ends past
        t_end when t_end is not commensurate with the steps); False if
        max_blocks (> 0) ran out first."""
        reached = C.c_int(0)
        self._check(self.lib.g6d_run(self.h, float(t_end), int(max_blocks),
                                     self.steps_per_graph, C.byref(reached)), "g6d_run")
        return bool(reached.value)

    def steps(self, nblocks: int) -> None:
        self._check(self.lib.g6d_run(self.h, float("inf"), int(nblocks), self.steps_per_graph,
                                     None), "g6d_run")

    def synchronize(self) -> None:
        """Every particle behind the system time steps to it (ph4's
        synchronize_all)."""
        self._check(self.lib.g6d_synchronize(self.h), "g6d_synchronize")

    def energy(self):
        k, u = C.c_double(), C.c_double()
        self._check(self.lib.g6d_energy(self.h, C.byref(k), C.byref(u)), "g6d_energy")
        return k.value, u.value

    def state(self):
        x = np.zeros((self.n, 3))
        v = np.zeros((self.n, 3))
        self._check(self.lib.g6d_state(self.h, x.ctypes.data_as(_D), v.ctypes.data_as(_D)),
                    "g6d_state")
        return x, v

    def state_raw(self):
        """x and v as last corrected, each at its particle's own time (ph4's
        get_state, src/amuse_ph4/interface.cc:485-502)."""
        x = np.zeros((self.n, 3))
        v = np.zeros((self.n, 3))
        self._check(self.lib.g6d_state_raw(self.h, x.ctypes.data_as(_D), v.ctypes.data_as(_D)),
                    "g6d_state_raw")
        return x, v

    def timesteps(self):
        t = np.zeros(self.n)
        dt = np.zeros(self.n)
        self._check(self.lib.g6d_steps(self.h, t.ctypes.data_as(_D), dt.ctypes.data_as(_D)),
                    "g6d_steps")
        return t, dt

    def set_precision(self, precision: int) -> None:
        self.precision = int(precision)
        if self.h:
            self._check(self.lib.g6d_set_precision(self.h, self.precision), "g6d_set_precision")

    @property
    def time(self) -> float:
        return self.counters()["system_time"]

    def counters(self):
        b, p, t = C.c_longlong(), C.c_longlong(), C.c_double()
        self._check(self.lib.g6d_counters(self.h, C.byref(b), C.byref(p), C.byref(t)),
                    "g6d_counters")
        return {"block_steps": b.value, "particle_steps": p.value, "system_time": t.value}

    def close(self):
        if self.h:
            self.lib.g6d_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
