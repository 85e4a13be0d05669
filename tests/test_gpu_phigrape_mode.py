"""phiGRAPE's use of the GRAPE-6 ABI, through the Fortran by-pointer entry
points its objects call (phiGRAPE is Fortran; no compiler for it here, so
the call sequence is restated):

* initgrape.F: g6_open(clusterid), g6_npipes(), g6_set_tunit(48) and
  g6_set_xunit(48) with INTEGER arguments;
* sendbodies2grape.F / update_grape.F: g6_set_j_particle with address
  j-1-nbh, index = the body's 1-based global index, a2by18 = 0,
  a1by6 = adot/6, aby2 = a/2;
* gravity.F:54-111: g6_set_ti(tnow), then per chunk of npipe active bodies a
  first firsthalf/lasthalf on a rough guess (ifirst = 1) and a second
  firsthalf/lasthalf2 whose nearest-neighbour output is used as a global
  index (gravity.F:117-121);
* gravity.F:300-371 (gravity_at_point): field probes with index -1, zero
  velocity, the same two calls per chunk;
* interface.F: g6_close / g6_open around each evolve.

acc, jerk and pot must match the 80-bit oracle on the same predicted j-set to
the plain path's bar, nn (a global index) the g6lib restatement exactly, and
the probes the direct sum with nothing skipped.
"""
import ctypes as C

import numpy as np
import pytest

import oracle
from amuse_amd import ic
from amuse_amd.g6 import LIBPATH
from conftest import print_case
from test_gpu_parity import COND

pytestmark = pytest.mark.gpu

_I = C.POINTER(C.c_int)
_D = C.POINTER(C.c_double)


def fortran_lib():
    L = C.CDLL(LIBPATH)
    for name, res, args in (
            ("g6_open_", C.c_int, [_I]), ("g6_close_", C.c_int, [_I]),
            ("g6_npipes_", C.c_int, []), ("g6_set_tunit_", C.c_int, [_I]),
            ("g6_set_xunit_", C.c_int, [_I]), ("g6_set_ti_", C.c_int, [_I, _D]),
            ("g6_set_j_particle_", C.c_int, [_I, _I, _I, _D, _D, _D, _D, _D, _D, _D, _D]),
            ("g6calc_firsthalf_", None, [_I, _I, _I, _I, _D, _D, _D, _D, _D, _D, _D]),
            ("g6calc_lasthalf_", C.c_int, [_I, _I, _I, _I, _D, _D, _D, _D, _D, _D, _D]),
            ("g6calc_lasthalf2_", C.c_int, [_I, _I, _I, _I, _D, _D, _D, _D, _D, _D, _D, _I]),
            ("g6x_set_npipes", C.c_int, [C.c_int]), ("g6x_set_precision", C.c_int, [C.c_int])):
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    return L


def ref(v, ctype):
    return C.byref(ctype(v))


def arr(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def bodies(n=2500, seed=41):
    m, x, v = ic.plummer(n, seed)
    rng = np.random.Generator(np.random.PCG64(seed))
    a = rng.normal(0, 1.0, (n, 3))
    adot = rng.normal(0, 5.0, (n, 3))
    tnow = 0.5
    dt = 2.0 ** -rng.integers(3, 11, n)
    t = tnow - rng.uniform(0, 1, n) * dt
    return m, x, v, a, adot, t, dt, tnow


def two_calls(L, cl, nj, idx, xi, vi, eps2, guess_a, guess_j, guess_p, nn_out):
    """One chunk as gravity.F does it: a guess pass, then the real one."""
    ni = idx.size
    h2 = np.full(ni, eps2)
    a, jk, p = arr(guess_a), arr(guess_j), arr(guess_p)
    L.g6calc_firsthalf_(cl, ref(nj, C.c_int), ref(ni, C.c_int), idx.ctypes.data_as(_I),
                        xi.ctypes.data_as(_D), vi.ctypes.data_as(_D), a.ctypes.data_as(_D),
                        jk.ctypes.data_as(_D), p.ctypes.data_as(_D), ref(eps2, C.c_double),
                        h2.ctypes.data_as(_D))
    L.g6calc_lasthalf_(cl, ref(nj, C.c_int), ref(ni, C.c_int), idx.ctypes.data_as(_I),
                       xi.ctypes.data_as(_D), vi.ctypes.data_as(_D), ref(eps2, C.c_double),
                       h2.ctypes.data_as(_D), a.ctypes.data_as(_D), jk.ctypes.data_as(_D),
                       p.ctypes.data_as(_D))
    jk[jk == 0.0] = 1e-5  # gravity.F:103-107
    L.g6calc_firsthalf_(cl, ref(nj, C.c_int), ref(ni, C.c_int), idx.ctypes.data_as(_I),
                        xi.ctypes.data_as(_D), vi.ctypes.data_as(_D), a.ctypes.data_as(_D),
                        jk.ctypes.data_as(_D), p.ctypes.data_as(_D), ref(eps2, C.c_double),
                        h2.ctypes.data_as(_D))
    if nn_out is None:
        rc = L.g6calc_lasthalf_(cl, ref(nj, C.c_int), ref(ni, C.c_int), idx.ctypes.data_as(_I),
                                xi.ctypes.data_as(_D), vi.ctypes.data_as(_D),
                                ref(eps2, C.c_double), h2.ctypes.data_as(_D),
                                a.ctypes.data_as(_D), jk.ctypes.data_as(_D), p.ctypes.data_as(_D))
    else:
        rc = L.g6calc_lasthalf2_(cl, ref(nj, C.c_int), ref(ni, C.c_int), idx.ctypes.data_as(_I),
                                 xi.ctypes.data_as(_D), vi.ctypes.data_as(_D),
                                 ref(eps2, C.c_double), h2.ctypes.data_as(_D),
                                 a.ctypes.data_as(_D), jk.ctypes.data_as(_D),
                                 p.ctypes.data_as(_D), nn_out.ctypes.data_as(_I))
    assert rc == 0
    return a, jk, p


@pytest.mark.parametrize("npipe", [48, 16384])
def test_phigrape_gravity_and_probes(npipe):
    m, x, v, a, adot, t, dt, tnow = bodies()
    n = m.size
    eps2 = 1e-4
    gidx = np.arange(1, n + 1, dtype=np.int32)  # 1-based global indices
    L = fortran_lib()
    cl = ref(0, C.c_int)
    print_case(f"phiGRAPE mode n={n} npipe={npipe}")
    for cycle in range(2):  # interface.F reopens the unit per evolve
        assert L.g6_open_(cl) == 0
        L.g6x_set_precision(0)
        L.g6x_set_npipes(npipe)
        try:
            assert L.g6_npipes_() == npipe
            assert L.g6_set_tunit_(ref(48, C.c_int)) == 0
            assert L.g6_set_xunit_(ref(48, C.c_int)) == 0
            z = np.zeros(3)
            for j in range(n):
                L.g6_set_j_particle_(cl, ref(j, C.c_int), ref(int(gidx[j]), C.c_int),
                                     ref(float(t[j]), C.c_double), ref(float(dt[j]), C.c_double),
                                     ref(float(m[j]), C.c_double), z.ctypes.data_as(_D),
                                     arr(adot[j] / 6).ctypes.data_as(_D),
                                     arr(a[j] / 2).ctypes.data_as(_D),
                                     arr(v[j]).ctypes.data_as(_D), arr(x[j]).ctypes.data_as(_D))
            assert L.g6_set_ti_(cl, ref(tnow, C.c_double)) == 0
            xp, vp = oracle.predict(tnow, t, x, v, a / 2, adot / 6)

            # active bodies: every third one, in a shuffled order
            act = np.random.Generator(np.random.PCG64(5 + cycle)).permutation(n)[: n // 3]
            A = np.zeros((act.size, 3)); J = np.zeros((act.size, 3)); P = np.zeros(act.size)
            NN = np.zeros(act.size, dtype=np.int32)
            for i0 in range(0, act.size, npipe):
                sel = act[i0:i0 + npipe]
                k = sel.size
                idx = np.ascontiguousarray(gidx[sel])
                nn = np.zeros(k, dtype=np.int32)
                A[i0:i0 + k], J[i0:i0 + k], P[i0:i0 + k] = two_calls(
                    L, cl, n, idx, arr(xp[sel]), arr(vp[sel]), eps2, np.ones((k, 3)),
                    np.full((k, 3), 10.0), np.full(k, -1.0), nn)
                NN[i0:i0 + k] = nn
            ea, ej, ep, sa, sj, sp = oracle.force_ld(gidx[act], xp[act], vp[act], eps2, gidx, m,
                                                     xp, vp)
            _, _, _, enn, _ = oracle.force(gidx[act], xp[act], vp[act], eps2, gidx, m, xp, vp)
            assert (np.abs(A - ea).max(1) <= COND["acc"] * sa).all()
            assert (np.abs(J - ej).max(1) <= COND["jerk"] * sj).all()
            assert (np.abs(P - ep) <= COND["pot"] * sp).all()
            assert np.array_equal(NN, enn)
            assert (NN >= 1).all() and (NN <= n).all() and (NN != gidx[act]).all()

            # gravity_at_point: probes with index -1, nothing skipped
            pts = np.random.Generator(np.random.PCG64(9)).normal(0, 1.0, (300, 3))
            PA = np.zeros((300, 3)); PP = np.zeros(300)
            for i0 in range(0, 300, npipe):
                k = min(npipe, 300 - i0)
                idx = np.full(k, -1, dtype=np.int32)
                PA[i0:i0 + k], _, PP[i0:i0 + k] = two_calls(
                    L, cl, n, idx, arr(pts[i0:i0 + k]), np.zeros((k, 3)), eps2,
                    np.ones((k, 3)), np.full((k, 3), 10.0), np.full(k, -1.0), None)
            pa, _, pp, psa, _, psp = oracle.force_ld(np.full(300, -1), pts, np.zeros((300, 3)),
                                                     eps2, gidx, m, xp, vp)
            assert (np.abs(PA - pa).max(1) <= COND["acc"] * psa).all()
            assert (np.abs(PP - pp) <= COND["pot"] * psp).all()
        finally:
            L.g6x_set_npipes(16384)
            assert L.g6_close_(cl) == 0
