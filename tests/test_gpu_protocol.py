"""The GRAPE-6 call protocol beyond set_ti -> firsthalf -> lasthalf.

* ph4's jdata::sync_gpu (src/amuse_ph4/src/gpu.cc:234-296, called from
  two_body.cc:480,1056) latches g6_npipes() rows of one repeated id through
  firsthalf and never calls lasthalf; set_j_particle updates follow, then
  an ordinary firsthalf + lasthalf2 that must give the oracle's answer for
  the new block and the updated j-set (a later firsthalf supersedes the
  latched one).  Run with the force launched lazily at lasthalf (default)
  and eagerly at firsthalf (G6_EAGER=1, as sapporo launches in firsthalf,
  lib/sapporo_light/sapporo.cpp:114-152); over a short j-range (the force
  kernel predicts its j-slots) and a long one (predictor launch).
* The neighbour query of a small block (inputs read in place from pinned
  host memory, j-slots predicted by the force kernel) with j-slots that
  have their own times, so prediction matters; set_ti moved between the
  calc and read_neighbour_list must not change the lists: they belong to
  the calc (sapporo builds them inside the force kernel,
  lib/sapporo_light/dev_evaluate_gravity.cu:60-67).
"""
import os

import numpy as np
import pytest

import oracle
from amuse_amd import ic
from conftest import print_case

pytestmark = pytest.mark.gpu


def jset(n, seed, t0):
    m, x, v = ic.plummer(n, seed)
    rng = np.random.Generator(np.random.PCG64(seed + 1))
    a2 = rng.normal(0, 0.3, (n, 3))
    j6 = rng.normal(0, 0.3, (n, 3))
    tj = t0 - rng.uniform(0, 1 / 32, n)
    ids = np.arange(1, n + 1, dtype=np.int32)
    return dict(m=m, x=x, v=v, a2=a2, j6=j6, tj=tj, ids=ids, k18=np.zeros((n, 3)))


def upload(g6, J, slots=None):
    s = np.arange(J["ids"].size) if slots is None else slots
    g6.set_j_particles(s, J["ids"][s], J["tj"][s], 0.0, J["m"][s], J["k18"][s], J["j6"][s],
                       J["a2"][s], J["v"][s], J["x"][s])


def check_fp64(out, ids_i, xi, vi, eps2, J, ti):
    acc, jerk, pot, nn, idx = out
    xp, vp = oracle.predict(ti, J["tj"], J["x"], J["v"], J["a2"], J["j6"], J["k18"])
    _, _, _, onn, _ = oracle.force(ids_i, xi, vi, eps2, J["ids"], J["m"], xp, vp)
    ea, ej, ep, sa, sj, sp = oracle.force_ld(ids_i, xi, vi, eps2, J["ids"], J["m"], xp, vp)
    assert np.array_equal(idx, ids_i)
    assert np.array_equal(nn, onn)
    for val, ex, s, tol in ((acc, ea, sa, 1e-15), (jerk, ej, sj, 4e-15), (pot, ep, sp, 1e-15)):
        d = np.abs(val - ex).reshape(len(ids_i), -1).max(1)
        assert (d <= tol * s).all()


def check_mixed(out, ids_i, xi, vi, eps2, J, ti):
    acc, jerk, pot, nn, idx = out
    xp, vp = oracle.predict(ti, J["tj"], J["x"], J["v"], J["a2"], J["j6"], J["k18"])
    ea, ej, ep, sa, sj, sp = oracle.force_ld(ids_i, xi, vi, eps2, J["ids"], J["m"], xp, vp)
    assert np.array_equal(idx, ids_i)
    assert (np.linalg.norm(acc - ea, axis=1) <= 2e-6 * sa).all()
    assert (np.linalg.norm(jerk - ej, axis=1) <= 2e-6 * sj).all()
    assert (np.abs(pot - ep) <= 2e-6 * sp).all()


@pytest.mark.parametrize("eager", [0, 1])
@pytest.mark.parametrize("n", [1500, 20000])
@pytest.mark.parametrize("prec", [0, 1])
def test_sync_gpu_firsthalf_without_lasthalf(g6lib, eager, n, prec):
    print_case(f"sync_gpu protocol eager={eager} n={n} prec={prec}")
    old = os.environ.get("G6_EAGER")
    os.environ["G6_EAGER"] = str(eager)
    try:
        g6lib.open()
    finally:
        if old is None:
            os.environ.pop("G6_EAGER", None)
        else:
            os.environ["G6_EAGER"] = old
    g6 = g6lib
    try:
        g6.set_precision(prec)
        g6.set_predictor_compat(False)
        g6.set_tiles(0, 0)
        eps2 = 1e-4
        t1, t2 = 0.5, 0.5 + 1 / 64
        J = jset(n, 90 + n, t1)
        upload(g6, J)
        g6.set_ti(t1)
        # jdata::sync_gpu: g6_npipes() copies of particle 0, h2 = eps2, no lasthalf
        npipes = g6.npipes()
        reps = np.full(npipes, J["ids"][0], dtype=np.int32)
        g6.calc_firsthalf(n, reps, np.repeat(J["x"][:1], npipes, 0),
                          np.repeat(J["v"][:1], npipes, 0), eps2,
                          h2=np.full(npipes, eps2))
        # the corrector's j updates for a block, then the next block step
        rng = np.random.Generator(np.random.PCG64(7))
        upd = np.sort(rng.choice(n, 200, replace=False))
        J["x"][upd] += rng.normal(0, 1e-3, (upd.size, 3))
        J["v"][upd] += rng.normal(0, 1e-3, (upd.size, 3))
        J["tj"][upd] = t2
        upload(g6, J, upd)
        g6.set_ti(t2)
        sel = np.arange(3, n, max(1, n // 300))[:300]
        xp, vp = oracle.predict(t2, J["tj"], J["x"], J["v"], J["a2"], J["j6"], J["k18"])
        out = g6.calc(n, J["ids"][sel], xp[sel], vp[sel], eps2)
        (check_fp64 if prec == 0 else check_mixed)(out, J["ids"][sel], xp[sel], vp[sel], eps2,
                                                   J, t2)
        # a second sync_gpu latch and a calc with lasthalf (no nn) after it
        g6.calc_firsthalf(n, reps[:7], np.repeat(xp[:1], 7, 0), np.repeat(vp[:1], 7, 0), eps2)
        g6.calc_firsthalf(n, J["ids"][sel], xp[sel], vp[sel], eps2)
        acc, jerk, pot, idx = g6.calc_lasthalf(n, sel.size, eps2)
        assert np.array_equal(idx, J["ids"][sel])
        assert np.array_equal(acc, out[0]) and np.array_equal(jerk, out[1])
        assert np.array_equal(pot, out[2])
    finally:
        g6.set_tiles(0, 0)
        g6.set_precision(0)
        g6.close()


@pytest.mark.parametrize("prec", [0, 1])
def test_neighbour_query_small_block_in_place(g6, prec):
    print_case(f"neighbour query ni=32 nj=1024 prec={prec}")
    n, eps2 = 1024, 1e-4
    t1, t2 = 1.0, 1.25
    J = jset(n, 55, t1)
    J["v"] *= 4.0            # lists change visibly between t1 and t2
    upload(g6, J)
    g6.set_precision(prec)
    g6.set_ti(t1)
    sel = np.arange(0, n, 32)
    xp1, vp1 = oracle.predict(t1, J["tj"], J["x"], J["v"], J["a2"], J["j6"], J["k18"])
    xp2, _ = oracle.predict(t2, J["tj"], J["x"], J["v"], J["a2"], J["j6"], J["k18"])
    h2 = np.full(sel.size, 0.01)
    h2[:4] = 0.3 ** 2
    out = g6.calc(n, J["ids"][sel], xp1[sel], vp1[sel], eps2, h2=h2)
    if prec == 0:
        check_fp64(out, J["ids"][sel], xp1[sel], vp1[sel], eps2, J, t1)
    c1, l1 = oracle.ngb(J["ids"][sel], xp1[sel], h2, J["ids"], xp1, cap=n)
    c2, _ = oracle.ngb(J["ids"][sel], xp1[sel], h2, J["ids"], xp2, cap=n)
    assert (c1 != c2).any(), "the two times must give different lists"
    assert c1.max() <= 256
    for move_ti in (False, True):
        if move_ti:
            g6.set_ti(t2)
        flag = g6.read_neighbour_list()
        assert flag == 0
        for q in range(sel.size):
            rc, cnt, got = g6.get_neighbour_list(q, 256)
            assert rc == 0 and cnt == c1[q], (move_ti, q)
            assert np.array_equal(got, np.sort(l1[q, :c1[q]])), (move_ti, q)
    # the next calc at t2 gives the lists at t2
    out2 = g6.calc(n, J["ids"][sel], xp1[sel], vp1[sel], eps2, h2=h2)
    assert g6.read_neighbour_list() == 0
    for q in range(sel.size):
        _, cnt, _ = g6.get_neighbour_list(q, 256)
        assert cnt == c2[q]
    assert out2[0].shape == (sel.size, 3)
