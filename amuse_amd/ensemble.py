"""Batched ensembles of independent clusters on one GPU (include/g6x.h g6c_*).

Each cluster is integrated with the shared-step 4th-order Hermite scheme of
src/amuse_hermite/interface.cc (predictor :266-281, collision-time estimate
:496-531, corrector :536-556, step = dt_param x collision time :618-637)
with the step rounded down to a power of two dividing the cluster time; all
clusters advance together, a step round at a time: the step kernels
launched one by one (run(use_graph=False), the fastest here: queued kernels
run back to back), replayed from a hipGraph (run(use_graph=True)), or the
whole run in one persistent launch with grid barriers between the phases
(run(persistent=True)); the three give the same bits.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .g6 import G6Error, load_library

_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)


def _bind(lib):
    sig = {
        "g6c_create": (C.c_void_p, [C.c_int, _I, _D, _D, _D, C.c_double, C.c_double, C.c_int]),
        "g6c_start": (C.c_int, [C.c_void_p, C.c_double]),
        "g6c_run": (C.c_longlong, [C.c_void_p, C.c_int, C.c_int, C.c_longlong]),
        "g6c_energy": (C.c_int, [C.c_void_p, _D]),
        "g6c_state": (C.c_int, [C.c_void_p, _D, _D, _D, _I]),
        "g6c_forces": (C.c_int, [C.c_void_p, C.c_int, _D, _D, _D]),
        "g6c_set_tiles": (C.c_int, [C.c_void_p, _I]),
        "g6c_get_tiles": (C.c_int, [C.c_void_p, _I]),
        "g6c_set_jtiles": (C.c_int, [C.c_void_p, _I]),
        "g6c_get_jtiles": (C.c_int, [C.c_void_p, _I]),
        "g6c_autotune": (C.c_int, [C.c_void_p]),
        "g6c_set_precision": (C.c_int, [C.c_void_p, C.c_int]),
        "g6c_elapsed_ms": (C.c_double, [C.c_void_p]),
        "g6c_last_error": (C.c_char_p, [C.c_void_p]),
        "g6c_free": (None, [C.c_void_p]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


class Ensemble:
    def __init__(self, counts, m, x, v, eps2=1e-6, dt_param=0.03, precision=1):
        self.lib = _bind(load_library())
        self.counts = np.ascontiguousarray(counts, dtype=np.int32)
        self.m = np.ascontiguousarray(m, dtype=np.float64)
        n = self.m.size
        if self.counts.ndim != 1 or (self.counts <= 0).any() or self.counts.sum() != n:
            raise ValueError("counts must be positive and sum to the number of masses")
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(n, 3)
        v = np.ascontiguousarray(v, dtype=np.float64).reshape(n, 3)
        self.h = self.lib.g6c_create(self.counts.size, self.counts.ctypes.data_as(_I),
                                     self.m.ctypes.data_as(_D), x.ctypes.data_as(_D),
                                     v.ctypes.data_as(_D), float(eps2), float(dt_param),
                                     int(precision))
        if not self.h:
            raise G6Error("g6c_create failed (no GPU, bad sizes, or a coordinate not finite or beyond 1e15)")
        self.n = n
        self.nc = self.counts.size

    def _check(self, rc, what):
        if rc < 0:
            raise G6Error(f"{what} failed ({rc}): {self.lib.g6c_last_error(self.h).decode()}")
        return rc

    def start(self, t_end: float) -> None:
        self._check(self.lib.g6c_start(self.h, float(t_end)), "g6c_start")

    def run(self, steps_per_graph: int = 8, use_graph: bool = True, max_rounds: int = 0,
            persistent: bool = False) -> int:
        """Step rounds until every cluster reached t_end (or max_rounds > 0
        rounds).  persistent: the whole run in one cooperative launch (the
        four phases of a round separated by grid barriers, the work list
        rebuilt on the device); otherwise steps_per_graph steps per hipGraph
        replay (use_graph) or per eager batch."""
        mode = 2 if persistent else int(bool(use_graph))
        return self._check(self.lib.g6c_run(self.h, int(steps_per_graph), mode,
                                            int(max_rounds)), "g6c_run")

    def energy(self) -> np.ndarray:
        e = np.zeros(self.nc)
        self._check(self.lib.g6c_energy(self.h, e.ctypes.data_as(_D)), "g6c_energy")
        return e

    def state(self):
        x = np.zeros((self.n, 3))
        v = np.zeros((self.n, 3))
        t = np.zeros(self.nc)
        ns = np.zeros(self.nc, dtype=np.int32)
        self._check(self.lib.g6c_state(self.h, x.ctypes.data_as(_D), v.ctypes.data_as(_D),
                                       t.ctypes.data_as(_D), ns.ctypes.data_as(_I)), "g6c_state")
        return x, v, t, ns

    def forces(self, precision: int):
        a = np.zeros((self.n, 3))
        j = np.zeros((self.n, 3))
        p = np.zeros(self.n)
        self._check(self.lib.g6c_forces(self.h, int(precision), a.ctypes.data_as(_D),
                                        j.ctypes.data_as(_D), p.ctypes.data_as(_D)), "g6c_forces")
        return a, j, p

    def set_tiles(self, njs_per_class) -> None:
        arr = np.ascontiguousarray(njs_per_class, dtype=np.int32)
        self._check(self.lib.g6c_set_tiles(self.h, arr.ctypes.data_as(_I)), "g6c_set_tiles")

    def tiles(self) -> list:
        arr = np.zeros(8, dtype=np.int32)
        self.lib.g6c_get_tiles(self.h, arr.ctypes.data_as(_I))
        return arr.tolist()

    def set_jtiles(self, tj_per_class) -> None:
        arr = np.ascontiguousarray(tj_per_class, dtype=np.int32)
        self._check(self.lib.g6c_set_jtiles(self.h, arr.ctypes.data_as(_I)), "g6c_set_jtiles")

    def jtiles(self) -> list:
        arr = np.zeros(8, dtype=np.int32)
        self.lib.g6c_get_jtiles(self.h, arr.ctypes.data_as(_I))
        return arr.tolist()

    def autotune(self) -> int:
        return self._check(self.lib.g6c_autotune(self.h), "g6c_autotune")

    def set_precision(self, precision: int) -> None:
        self._check(self.lib.g6c_set_precision(self.h, int(precision)), "g6c_set_precision")

    def elapsed_ms(self) -> float:
        return self.lib.g6c_elapsed_ms(self.h)

    def close(self) -> None:
        if self.h:
            self.lib.g6c_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def make_ensemble(nclusters: int, seed: int, nmin=16, nmax=2048, kroupa=(0.08, 20.0)):
    """Clusters with N log-uniform in [nmin, nmax], Kroupa masses normalised to
    M = 1 per cluster, Plummer positions and velocities in standard units."""
    from . import ic
    rng = np.random.Generator(np.random.PCG64(seed))
    counts = np.rint(np.exp(rng.uniform(np.log(nmin), np.log(nmax), nclusters))).astype(np.int32)
    ms, xs, vs = [], [], []
    for c, n in enumerate(counts):
        masses = ic.kroupa(int(n), rng, kroupa[0], kroupa[1])
        m = masses / masses.sum()
        r2 = np.random.Generator(np.random.PCG64(rng.integers(1 << 62)))
        _, x, v = ic.plummer_raw(int(n), r2)
        x, v = ic.to_nbody_units(x, v)
        x, v = ic.center(m, x, v)
        ms.append(m); xs.append(x); vs.append(v)
    return counts, np.concatenate(ms), np.concatenate(xs), np.concatenate(vs)
