"""CPU oracle for the GRAPE-6 force path — TEST INFRASTRUCTURE ONLY.

Loaded only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg, always as the checker or as the timed CPU baseline, never as the thing
measured or shipped.  The product library never imports this package.

    liboracle.so       own C restatement of lib/g6lib/g6lib.c (g6_oracle.c)
    libcpubest.so      ph4's CPU pair loop (idata.cc:195-236) vectorised and
                       threaded: the best-effort CPU row of the benchmark
    _ref/libg6ref.so   the reference lib/g6lib/g6lib.c compiled from its
                       source by oracle/Makefile where the reference checkout
                       exists; elsewhere only the built file is used, never
                       the source

Arrays follow the ABI: [n][3] float64 vectors, int32 ids.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libg6ref.so")
CPUBEST_SO = os.path.join(HERE, "libcpubest.so")

_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)


def build(reference: str | None = None) -> None:
    """Compiles liboracle.so and, when the reference is readable, _ref."""
    cmd = ["make", "-s", "-C", HERE]
    if reference:
        cmd.append(f"REFERENCE={reference}")
    subprocess.run(cmd, check=True)


def _d(a):
    return None if a is None else a.ctypes.data_as(_D)


def _i(a):
    return None if a is None else a.ctypes.data_as(_I)


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a.reshape(shape) if shape is not None else a


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            build()
        L = C.CDLL(ORACLE_SO)
        L.orc_predict.argtypes = [C.c_int, C.c_double, _D, _D, _D, _D, _D, _D, C.c_int, _D, _D]
        L.orc_predict.restype = None
        L.orc_force.argtypes = [C.c_int, _I, _D, _D, C.c_double, C.c_int, _I, _D, _D, _D, _D, _D,
                                _D, _I, _D]
        L.orc_force.restype = None
        L.orc_force_ld.argtypes = [C.c_int, _I, _D, _D, C.c_double, C.c_int, _I, _D, _D, _D, _D,
                                   _D, _D, _D, _D, _D]
        L.orc_force_ld.restype = None
        L.orc_ngb.argtypes = [C.c_int, _I, _D, _D, C.c_int, _I, _D, C.c_int, _I, _I]
        L.orc_ngb.restype = None
        L.orc_baseline_run.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
                                       C.c_double, _I, _D, _D, _D, _D, _D, _D, _D, _D, _D, _D, _I]
        L.orc_baseline_run.restype = C.c_double
        _lib = L
    return _lib


def predict(ti, tj, x, v, a2, j6, k18=None, compat=False):
    """Predicted (xp, vp) of every j (g6lib.c:64-92)."""
    n = len(tj)
    tj = _f64(tj)
    x, v, a2, j6 = (_f64(q, (n, 3)) for q in (x, v, a2, j6))
    k18 = np.zeros((n, 3)) if k18 is None else _f64(k18, (n, 3))
    xp = np.zeros((n, 3))
    vp = np.zeros((n, 3))
    lib().orc_predict(n, float(ti), _d(tj), _d(x), _d(v), _d(a2), _d(j6), _d(k18),
                      int(bool(compat)), _d(xp), _d(vp))
    return xp, vp


def force(id_i, xi, vi, eps2, id_j, m, xp, vp):
    """g6lib pair loop in FP64 with pow() (g6lib.c:123-186).

    Returns acc (ni,3), jerk (ni,3), pot (ni,), nn ids (ni,), nn r^2 (ni,)."""
    id_i = np.ascontiguousarray(id_i, dtype=np.int32)
    id_j = np.ascontiguousarray(id_j, dtype=np.int32)
    ni, nj = id_i.size, id_j.size
    xi, vi = _f64(xi, (ni, 3)), _f64(vi, (ni, 3))
    xp, vp, m = _f64(xp, (nj, 3)), _f64(vp, (nj, 3)), _f64(m)
    acc = np.zeros((ni, 3)); jerk = np.zeros((ni, 3)); pot = np.zeros(ni)
    nn = np.zeros(ni, dtype=np.int32); r2 = np.zeros(ni)
    lib().orc_force(ni, _i(id_i), _d(xi), _d(vi), float(eps2), nj, _i(id_j), _d(m), _d(xp),
                    _d(vp), _d(acc), _d(jerk), _d(pot), _i(nn), _d(r2))
    return acc, jerk, pot, nn, r2


def force_ld(id_i, xi, vi, eps2, id_j, m, xp, vp):
    """Same sums in 80-bit long double plus per-i sums of |term|.

    Returns acc, jerk, pot, sacc, sjerk, spot."""
    id_i = np.ascontiguousarray(id_i, dtype=np.int32)
    id_j = np.ascontiguousarray(id_j, dtype=np.int32)
    ni, nj = id_i.size, id_j.size
    xi, vi = _f64(xi, (ni, 3)), _f64(vi, (ni, 3))
    xp, vp, m = _f64(xp, (nj, 3)), _f64(vp, (nj, 3)), _f64(m)
    acc = np.zeros((ni, 3)); jerk = np.zeros((ni, 3)); pot = np.zeros(ni)
    sa = np.zeros(ni); sj = np.zeros(ni); sp = np.zeros(ni)
    lib().orc_force_ld(ni, _i(id_i), _d(xi), _d(vi), float(eps2), nj, _i(id_j), _d(m), _d(xp),
                       _d(vp), _d(acc), _d(jerk), _d(pot), _d(sa), _d(sj), _d(sp))
    return acc, jerk, pot, sa, sj, sp


def ngb(id_i, xi, h2, id_j, xp, cap=256):
    """Neighbour ids (r^2 <= h2, self and id -1 excluded), ascending slot."""
    id_i = np.ascontiguousarray(id_i, dtype=np.int32)
    id_j = np.ascontiguousarray(id_j, dtype=np.int32)
    ni, nj = id_i.size, id_j.size
    xi, xp, h2 = _f64(xi, (ni, 3)), _f64(xp, (nj, 3)), _f64(h2)
    counts = np.zeros(ni, dtype=np.int32)
    lists = np.zeros((ni, cap), dtype=np.int32)
    lib().orc_ngb(ni, _i(id_i), _d(xi), _d(h2), nj, _i(id_j), _d(xp), cap, _i(counts), _i(lists))
    return counts, lists


def baseline_run(nthreads, npipes, nj, ni, ti, eps2, ids, tj, m, x, v, a2, j6, k18=None,
                 compat=False):
    """Timed CPU baseline: nthreads independent g6lib-style units over a
    disjoint split of i-slots [0, ni).  Returns (seconds, acc, jerk, pot, nn)."""
    ids = np.ascontiguousarray(ids, dtype=np.int32)
    n = ids.size
    tj, m = _f64(tj), _f64(m)
    x, v, a2, j6 = (_f64(q, (n, 3)) for q in (x, v, a2, j6))
    k18 = np.zeros((n, 3)) if k18 is None else _f64(k18, (n, 3))
    acc = np.zeros((ni, 3)); jerk = np.zeros((ni, 3)); pot = np.zeros(ni)
    nn = np.zeros(ni, dtype=np.int32)
    sec = lib().orc_baseline_run(int(nthreads), int(npipes), int(bool(compat)), int(nj), int(ni),
                                 float(ti), float(eps2), _i(ids), _d(tj), _d(m), _d(x), _d(v),
                                 _d(a2), _d(j6), _d(k18), _d(acc), _d(jerk), _d(pot), _i(nn))
    return sec, acc, jerk, pot, nn


def cpu_best_run(nthreads, ix, iv, eps2, jm, jx, jv):
    """Best-effort CPU forces (ph4's idata.cc:195-236 loop, OpenMP over i,
    SIMD over j).  Returns (seconds, acc, jerk, pot, nn slot)."""
    if not os.path.exists(CPUBEST_SO):
        build()
    L = C.CDLL(CPUBEST_SO)
    L.cpu_best_run.restype = C.c_double
    L.cpu_best_run.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, _D, _D, _D, _D, _D, _D,
                               _D, _D, _I]
    ix, iv = _f64(ix).reshape(-1, 3), _f64(iv).reshape(-1, 3)
    jx, jv, jm = _f64(jx).reshape(-1, 3), _f64(jv).reshape(-1, 3), _f64(jm)
    ni, nj = ix.shape[0], jm.size
    acc = np.zeros((ni, 3)); jerk = np.zeros((ni, 3)); pot = np.zeros(ni)
    nn = np.zeros(ni, dtype=np.int32)
    sec = L.cpu_best_run(int(nthreads), ni, nj, float(eps2), _d(ix), _d(iv), _d(jm), _d(jx),
                         _d(jv), _d(acc), _d(jerk), _d(pot), _i(nn))
    return sec, acc, jerk, pot, nn


class RefG6:
    """The reference lib/g6lib/g6lib.c, compiled by oracle/Makefile, driven
    through its own ABI (g6_open_ ... g6calc_lasthalf2_).  Capacity is its
    MAX_NUMBER_OF_PARTICLES (100000, g6lib.c:7)."""

    CAP = 100000

    def __init__(self, path: str = REF_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.L = C.CDLL(path)
        for name in ("g6_open_", "g6_close_", "g6_reset_"):
            getattr(self.L, name).argtypes = [_I]
            getattr(self.L, name).restype = C.c_int
        self.L.g6_set_ti_.argtypes = [_I, _D]
        self.L.g6_set_j_particle_.argtypes = [_I, _I, _I, _D, _D, _D, _D, _D, _D, _D, _D]
        self.L.g6calc_firsthalf_.argtypes = [_I, _I, _I, _I, _D, _D, _D, _D, _D, _D, _D]
        self.L.g6calc_firsthalf_.restype = None
        self.L.g6calc_lasthalf_.argtypes = [_I, _I, _I, _I, _D, _D, _D, _D, _D, _D, _D]
        self.L.g6calc_lasthalf2_.argtypes = [_I, _I, _I, _I, _D, _D, _D, _D, _D, _D, _D, _I]
        self._cid = C.c_int(0)

    def open(self):
        return self.L.g6_open_(C.byref(self._cid))

    def close(self):
        return self.L.g6_close_(C.byref(self._cid))

    def set_ti(self, ti):
        t = C.c_double(ti)
        return self.L.g6_set_ti_(C.byref(self._cid), C.byref(t))

    def set_j_particle(self, address, index, tj, dtj, mass, k18, j6, a2, v, x):
        a, i = C.c_int(int(address)), C.c_int(int(index))
        t, dt, mm = C.c_double(tj), C.c_double(dtj), C.c_double(mass)
        arrs = [np.ascontiguousarray(np.asarray(q, dtype=np.float64).reshape(3))
                for q in (k18, j6, a2, v, x)]
        return self.L.g6_set_j_particle_(C.byref(self._cid), C.byref(a), C.byref(i), C.byref(t),
                                         C.byref(dt), C.byref(mm), *[_d(q) for q in arrs])

    def calc(self, nj, index, xi, vi, eps2, nn=True, h2=None):
        idx = np.ascontiguousarray(index, dtype=np.int32).copy()
        ni = idx.size
        xi, vi = _f64(xi, (ni, 3)), _f64(vi, (ni, 3))
        h2 = np.zeros(ni) if h2 is None else _f64(h2)
        dummy = np.zeros((max(ni, 1), 3))
        e = C.c_double(eps2)
        cnj, cni = C.c_int(nj), C.c_int(ni)
        self.L.g6calc_firsthalf_(C.byref(self._cid), C.byref(cnj), C.byref(cni), _i(idx),
                                 _d(xi), _d(vi), _d(dummy), _d(dummy), _d(dummy), C.byref(e),
                                 _d(h2))
        acc = np.zeros((ni, 3)); jerk = np.zeros((ni, 3)); pot = np.zeros(ni)
        if nn:
            nnb = np.zeros(ni, dtype=np.int32)
            self.L.g6calc_lasthalf2_(C.byref(self._cid), C.byref(cnj), C.byref(cni), _i(idx),
                                     _d(xi), _d(vi), C.byref(e), _d(h2), _d(acc), _d(jerk),
                                     _d(pot), _i(nnb))
            return acc, jerk, pot, nnb, idx
        self.L.g6calc_lasthalf_(C.byref(self._cid), C.byref(cnj), C.byref(cni), _i(idx), _d(xi),
                                _d(vi), C.byref(e), _d(h2), _d(acc), _d(jerk), _d(pot))
        return acc, jerk, pot, idx
