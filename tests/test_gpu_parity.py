"""Parity of the HIP library with the oracle, through the GRAPE-6 ABI.

Bars (plain FP64 path):
  * nearest-neighbour ids bit-exact, ties included (the predictor and r^2
    round exactly like lib/g6lib/g6lib.c and ties go to the lowest slot);
  * acc, jerk, pot per particle within 1e-12 of the 80-bit oracle relative to
    |value|, for particles whose sum is not cancelling by more than 1e3
    (|value| >= 1e-3 sum|term|), and within 1e-15 * sum|term| for all;
  * index[] written back unchanged (g6lib.c:397,427).
Mixed path: per-particle error against the 80-bit oracle stated and bounded
in test_mixed_precision_error_bounds.
"""
import numpy as np
import pytest

import oracle
from conftest import load_golden, print_case

pytestmark = pytest.mark.gpu

REL = 1e-12          # per-particle relative bound, plain path
# bound relative to sum|term|, plain path; a jerk term cancels inside itself
# (v/s^1.5 - 3 (r.v) r/s^2.5), so its rounding is larger relative to |term|
COND = {"acc": 1e-15, "jerk": 4e-15, "pot": 1e-15}


def upload(g6, x, v, a2, j6, k18, m, tj, ids, skip=None):
    n = len(ids)
    sel = np.arange(n) if skip is None else np.flatnonzero(~skip)
    g6.set_j_particles(sel, ids[sel], tj[sel], 0.0, m[sel], k18[sel], j6[sel], a2[sel], v[sel],
                       x[sel])


def errors(val, exact, sums):
    """Per-particle |d| / |exact| and |d| / sum|term| (vectors by norm)."""
    if val.ndim == 2:
        d = np.linalg.norm(val - exact, axis=1)
        mag = np.linalg.norm(exact, axis=1)
    else:
        d = np.abs(val - exact)
        mag = np.abs(exact)
    with np.errstate(divide="ignore", invalid="ignore"):
        rel = np.where(mag > 0, d / mag, d)
        cond = np.where(sums > 0, d / sums, d)
    return rel, cond, mag / np.where(sums > 0, sums, 1)


def assert_plain_parity(acc, jerk, pot, exact, cond_scale=1.0):
    ea, ej, ep, sa, sj, sp = exact
    for val, ex, s, name in ((acc, ea, sa, "acc"), (jerk, ej, sj, "jerk"), (pot, ep, sp, "pot")):
        rel, cond, ratio = errors(val, ex, s)
        assert cond.max() <= cond_scale * COND[name], (name, cond.max())
        ok = ratio >= 1e-3
        if ok.any():
            assert rel[ok].max() <= REL, (name, rel[ok].max())


def test_golden_reference_vectors(g6):
    """The compiled reference's outputs, reproduced with the predictor-compat
    polynomial (the reference's own)."""
    g6.set_predictor_compat(True)
    for name in ["g6ref_dt0", "g6ref_dt", "g6ref_lattice", "g6ref_special", "g6ref_probe"]:
        print_case(f"golden {name}")
        d = load_golden(name)
        g6.reset()
        never = d["id_j"] == -2
        upload(g6, d["x"], d["v"], d["a2"], d["j6"], d["k18"], d["m"], d["tj"], d["id_j"],
               skip=never)
        g6.set_ti(float(d["ti"]))
        nj = int(d["nj"])
        acc, jerk, pot, nn, idx = g6.calc(nj, d["id_i"], d["xi"], d["vi"], float(d["eps2"]))
        assert np.array_equal(nn, d["nn"]), name
        assert np.array_equal(idx, d["id_i"])
        ids = d["id_j"][:nj].copy()
        ids[ids == -2] = -1
        xp, vp = oracle.predict(float(d["ti"]), d["tj"][:nj], d["x"][:nj], d["v"][:nj],
                                d["a2"][:nj], d["j6"][:nj], d["k18"][:nj], compat=True)
        exact = oracle.force_ld(d["id_i"], d["xi"], d["vi"], float(d["eps2"]), ids,
                                d["m"][:nj], xp, vp)
        assert_plain_parity(acc, jerk, pot, exact)
        # the reference's own doubles (plain sequential sum of pow() terms)
        # are within 10x the same bars of the exact sum
        assert_plain_parity(d["acc"], d["jerk"], d["pot"], exact, cond_scale=10.0)
        # lasthalf without nn gives bit-identical forces
        acc2, jerk2, pot2, _ = g6.calc(nj, d["id_i"], d["xi"], d["vi"], float(d["eps2"]), nn=False)
        np.testing.assert_array_equal(acc2, acc)
        np.testing.assert_array_equal(jerk2, jerk)
        np.testing.assert_array_equal(pot2, pot)
        for prec in (1,):
            g6.set_precision(prec)
            r1 = g6.calc(nj, d["id_i"], d["xi"], d["vi"], float(d["eps2"]), nn=True)
            r2 = g6.calc(nj, d["id_i"], d["xi"], d["vi"], float(d["eps2"]), nn=False)
            for p_, q_ in zip(r1[:3], r2[:3]):
                np.testing.assert_array_equal(p_, q_)
            g6.set_precision(0)


def plummer_state(n, seed, dt_nonzero=True):
    from amuse_amd import ic
    rng = np.random.Generator(np.random.PCG64(seed))
    # exact virial scaling (an O(N^2) potential in numpy) only where it is cheap
    m, x, v = ic.plummer(n, seed, scale=n <= 4096)
    a2 = rng.normal(0, 0.3, (n, 3))
    j6 = rng.normal(0, 0.3, (n, 3))
    k18 = np.zeros((n, 3))
    ti = 1.0
    tj = ti - (rng.uniform(0, 1, n) * 2.0 ** -rng.integers(3, 11, n) if dt_nonzero else 0)
    tj = np.broadcast_to(tj, (n,)).astype(float).copy()
    ids = np.arange(1, n + 1, dtype=np.int32)
    return dict(m=m, x=x, v=v, a2=a2, j6=j6, k18=k18, tj=tj, ti=ti, ids=ids)


@pytest.mark.parametrize("compat", [False, True])
def test_plummer_4096_block_sizes(g6, compat):
    s = plummer_state(4096, 21)
    g6.set_predictor_compat(compat)
    upload(g6, s["x"], s["v"], s["a2"], s["j6"], s["k18"], s["m"], s["tj"], s["ids"])
    g6.set_ti(s["ti"])
    xp, vp = oracle.predict(s["ti"], s["tj"], s["x"], s["v"], s["a2"], s["j6"], s["k18"],
                            compat=compat)
    n = 4096
    for ni in (1, 2, 7, 64, 333, 1024, 4096):
        print_case(f"plummer4096 compat={compat} ni={ni}")
        sel = np.linspace(0, n - 1, ni).astype(int)
        id_i, xi, vi = s["ids"][sel], xp[sel], vp[sel]
        acc, jerk, pot, nn, _ = g6.calc(n, id_i, xi, vi, 1e-4)
        _, _, _, onn, _ = oracle.force(id_i, xi, vi, 1e-4, s["ids"], s["m"], xp, vp)
        assert np.array_equal(nn, onn), ni
        exact = oracle.force_ld(id_i, xi, vi, 1e-4, s["ids"], s["m"], xp, vp)
        assert_plain_parity(acc, jerk, pot, exact)


def test_mixed_nn_ties_on_lattice(g6):
    """On the integer lattice every r^2 is exact in FP32, so distance ties
    are exact in the mixed path too: its nearest neighbour must be the
    reference's (first slot wins) in every tile shape, per-pair compare
    (scalar-load kernels) and running minimum with tile rescan (LDS) alike."""
    d = load_golden("g6ref_lattice")
    g6.reset()
    upload(g6, d["x"], d["v"], d["a2"], d["j6"], d["k18"], d["m"], d["tj"], d["id_j"])
    g6.set_ti(float(d["ti"]))
    g6.set_precision(1)
    nj = int(d["nj"])
    try:
        for njs, ipl, nchunk, smem in [(1, 2, 0, 1), (1, 1, 0, 0), (1, 2, 0, 0), (1, 4, 0, 0),
                                       (4, 1, 0, 0), (16, 1, 0, 0), (64, 1, 0, 0),
                                       (1, 2, 1, 0), (8, 2, 1, 0)]:
            g6.set_tiles(njs, ipl, nchunk, smem)
            nn = g6.calc(nj, d["id_i"], d["xi"], d["vi"], float(d["eps2"]))[3]
            assert np.array_equal(nn, d["nn"]), (njs, ipl, nchunk, smem)
    finally:
        g6.set_tiles(0, 0)
        g6.set_precision(0)


@pytest.mark.parametrize("prec", [1, 0])
def test_nn_ties_long_items(g6, prec):
    """Items whose lanes see >= 16384 j keep the nearest neighbour as a
    running minimum per block and search the slot once at the end (the wave
    takes one lane's block at a time); shorter items compare per pair.  On an
    integer lattice (every r^2 exact in FP32 and FP64, six equidistant
    neighbours per point) in shuffled slot order, both give the reference's
    first slot, in both precisions, through the ABI and through full passes
    over the slots (g6x_calc_self, whose items away from the diagonal skip
    the id compare)."""
    g = np.stack(np.meshgrid(np.arange(32), np.arange(32), np.arange(40), indexing="ij"),
                 -1).reshape(-1, 3).astype(np.float64)
    nj = g.shape[0]
    rng = np.random.Generator(np.random.PCG64(77))
    x = g[rng.permutation(nj)]
    ids = np.arange(1, nj + 1, dtype=np.int32)
    m = rng.uniform(0.5, 1.5, nj) / nj
    v = rng.normal(0, 0.1, (nj, 3))
    z = np.zeros((nj, 3))
    g6.reset()
    g6.set_j_particles(np.arange(nj, dtype=np.int32), ids, 0.0, 0.0, m, z, z, z, v, x)
    g6.set_ti(0.0)
    sel = rng.choice(nj, 256, replace=False)
    id_i, xi, vi = ids[sel], x[sel], v[sel]
    _, _, _, onn, _ = oracle.force(id_i, xi, vi, 0.0, ids, m, x, v)
    g6.set_precision(prec)
    try:
        # (njs, ipl, nchunk): lanes' j per item 40960, 40960, 20480 (long);
        # 10240, 10240, 2560 (short)
        for njs, ipl, nchunk in [(1, 2, 1), (1, 1, 1), (1, 2, 2), (1, 2, 4), (4, 1, 1),
                                 (16, 1, 1)]:
            print_case(f"nn lattice prec={prec} nj={nj} njs={njs} ipl={ipl} nchunk={nchunk}")
            g6.set_tiles(njs, ipl, nchunk, 0)
            nn = g6.calc(nj, id_i, xi, vi, 0.0)[3]
            assert np.array_equal(nn, onn), (njs, ipl, nchunk)
            g6.calc_self(nj, 0, nj, 0.0, nn=True)
            assert np.array_equal(g6.fetch(nj)[3][sel], onn), ("self", njs, ipl, nchunk)
    finally:
        g6.set_tiles(0, 0)
        g6.set_precision(0)


def test_spatial_full_pass(g6):
    """Mixed full passes of 2^17 slots take the spatial order by default
    (slots sorted along a Morton curve, far i-block/j-tile pairs with one
    FP32 subtraction relative to the i-block centre).  For the shapes that
    have a spatial kernel, with long (running-minimum) and short
    nearest-neighbour items, the results are within the mixed path's bound
    of the 80-bit oracle on sampled i, every nearest neighbour is at the
    minimal r^2 up to FP32 rounding, sum m a = 0 over all slots, and the
    slot-order pass (g6x_set_spatial(0)) agrees to the same bound; a shape
    without a spatial kernel runs in slot order, bit for bit."""
    n, eps2 = 1 << 17, 1e-6
    st = plummer_state(n, 41, dt_nonzero=False)
    ids, m, x, v = st["ids"], st["m"], st["x"], st["v"]
    g6.set_precision(1)
    upload(g6, x, v, st["a2"], st["j6"], st["k18"], m, st["tj"], ids)
    g6.set_ti(st["ti"])
    s = np.linspace(0, n - 1, 192).astype(int)
    exact = oracle.force_ld(ids[s], x[s], v[s], eps2, ids, m, x, v)
    _, _, _, onn, or2 = oracle.force(ids[s], x[s], v[s], eps2, ids, m, x, v)
    ea, ej, ep, sa, sj, sp = exact
    try:
        # (njs, ipl, nchunk): lanes' j per item 131072, 32768 (long), 8192
        # (short); 1 i per lane; (4, 1) has no spatial kernel
        for shape in [(1, 2, 1), (1, 2, 4), (1, 2, 16), (1, 1, 2), (0, 0, 0), (4, 1, 8)]:
            print_case(f"spatial full pass n={n} shape={shape}")
            g6.set_tiles(*shape)
            res = {}
            for sp_on in (True, False):
                g6.set_spatial(sp_on)
                g6.calc_self(n, 0, n, eps2, nn=True)
                res[sp_on] = g6.fetch(n)
            g6.set_spatial(True)
            if shape[0] == 4:
                for p_, q_ in zip(res[True], res[False]):
                    assert np.array_equal(p_, q_)
            for a, j, p, nn in res.values():
                for val, e, sm, name in ((a[s], ea, sa, "acc"), (j[s], ej, sj, "jerk"),
                                         (p[s], ep, sp, "pot")):
                    d = np.abs(val - e).reshape(len(s), -1)
                    d = np.sqrt((d * d).sum(1))
                    assert (d <= 2e-6 * sm).all(), (shape, name, (d / sm).max())
                dd = x[nn[s] - 1] - x[s]
                r2 = (dd * dd).sum(1)
                assert (r2 <= or2 * (1 + 1e-5)).all(), (shape, (r2 / or2).max())
                assert ((nn >= 1) & (nn <= n) & (nn != ids)).all()
                ma = np.linalg.norm((m[:, None] * a).sum(0))
                assert ma <= 1e-5 * (m * np.linalg.norm(a, axis=1)).sum()
            da = np.linalg.norm(res[True][0] - res[False][0], axis=1)
            assert (da <= 4e-6 * np.linalg.norm(res[False][0], axis=1) + 1e-300).mean() > 0.999
    finally:
        g6.set_tiles(0, 0)
        g6.set_spatial(True)
        g6.set_precision(0)


@pytest.mark.parametrize("prec", [0, 1])
def test_self_pass_id_compare_forms(g6, prec):
    """g6x_calc_self leaves the per-pair id compare out of the work items
    whose j-range holds none of their i-slots, but only while every slot
    below nj holds an id and no id is held twice.  Its results equal the ABI
    path's on the same i-block (which compares ids in every pair) bit for
    bit: with distinct ids (most items without the compare; long and short
    items), with one id held by two slots, and with an empty slot (both: the
    compare everywhere); the latter two also against the oracle, which skips
    equal ids and empty slots (lib/g6lib/g6lib.c:135-140)."""
    n = 40000
    st = plummer_state(n, 31, dt_nonzero=False)
    ids = st["ids"].copy()
    g6.set_precision(prec)
    try:
        for case in ("distinct", "repeat", "empty"):
            idj = ids.copy()
            skip = np.zeros(n, bool)
            if case == "repeat":
                idj[30001] = idj[17]
            if case == "empty":
                skip[2222] = True
            g6.reset()
            upload(g6, st["x"], st["v"], st["a2"], st["j6"], st["k18"], st["m"], st["tj"], idj,
                   skip=skip)
            g6.set_ti(st["ti"])
            for njs, ipl, nchunk in [(1, 2, 2), (1, 2, 8), (4, 1, 4)]:
                print_case(f"self pass prec={prec} {case} njs={njs} ipl={ipl} nchunk={nchunk}")
                g6.set_tiles(njs, ipl, nchunk, 0)
                g6.calc_self(n, 0, n, 1e-4, nn=True)
                a2_, j2_, p2_, nn2 = g6.fetch(n)
                lo, hi = 1000, 5096
                idi = idj[lo:hi].copy()
                idi[skip[lo:hi]] = -1
                a, j, p, nn, _ = g6.calc(n, idi, st["x"][lo:hi], st["v"][lo:hi], 1e-4)
                live = ~skip[lo:hi]
                np.testing.assert_array_equal(a[live], a2_[lo:hi][live])
                np.testing.assert_array_equal(j[live], j2_[lo:hi][live])
                np.testing.assert_array_equal(p[live], p2_[lo:hi][live])
                assert np.array_equal(nn[live], nn2[lo:hi][live])
            if case != "distinct" and prec == 0:
                s = np.array([17, 2222, 30001, 100, 39999])
                mj = st["m"] * ~skip
                jid = np.where(skip, -1, idj).astype(np.int32)
                _, _, _, onn, _ = oracle.force(jid[s], st["x"][s], st["v"][s], 1e-4, jid, mj,
                                               st["x"], st["v"])
                keep = ~skip[s]
                assert np.array_equal(nn2[s][keep], onn[keep])
                ea, ej, ep, sa, sj, sp = oracle.force_ld(jid[s], st["x"][s], st["v"][s], 1e-4,
                                                         jid, mj, st["x"], st["v"])
                d = np.linalg.norm(a2_[s] - ea, axis=1)
                assert (d[keep] <= 1e-15 * sa[keep]).all()
    finally:
        g6.set_tiles(0, 0)
        g6.set_precision(0)


def test_tile_shapes_agree(g6):
    s = plummer_state(3000, 22)
    upload(g6, s["x"], s["v"], s["a2"], s["j6"], s["k18"], s["m"], s["tj"], s["ids"])
    g6.set_ti(s["ti"])
    xp, vp = oracle.predict(s["ti"], s["tj"], s["x"], s["v"], s["a2"], s["j6"], s["k18"])
    sel = np.arange(0, 3000, 7)
    id_i, xi, vi = s["ids"][sel], xp[sel], vp[sel]
    exact = oracle.force_ld(id_i, xi, vi, 1e-4, s["ids"], s["m"], xp, vp)
    _, _, _, onn, _ = oracle.force(id_i, xi, vi, 1e-4, s["ids"], s["m"], xp, vp)
    shapes = [(1, 2, 1), (1, 1, 1), (1, 2, 0), (1, 1, 0), (2, 2, 0), (2, 1, 0), (4, 2, 0),
              (4, 1, 0), (8, 2, 0), (8, 1, 0), (16, 1, 0), (32, 1, 0), (64, 1, 0), (128, 1, 0),
              (256, 1, 0), (1, 4, 0), (2, 4, 0)]
    for prec in (0, 1):
        g6.set_precision(prec)
        first = None
        for njs, ipl, smem in shapes:
            for nchunk in (1, 3, 0):
                print_case(f"tiles prec={prec} njs={njs} ipl={ipl} smem={smem} nchunk={nchunk}")
                g6.set_tiles(njs, ipl, nchunk, smem)
                acc, jerk, pot, nn, _ = g6.calc(3000, id_i, xi, vi, 1e-4)
                if prec == 0:
                    assert np.array_equal(nn, onn), (njs, ipl, nchunk, smem)
                    assert_plain_parity(acc, jerk, pot, exact)
                else:
                    # mixed: every shape within FP32 rounding of the first; the
                    # FP32 r^2 of a pair does not depend on the shape, so the
                    # nearest neighbour (per-pair compare in the scalar-load
                    # kernels, running minimum + tile rescan in the LDS ones)
                    # is the same slot in every shape
                    if first is None:
                        first = (acc, pot, nn)
                    scale = np.abs(first[0]).max()
                    assert np.abs(acc - first[0]).max() < 1e-5 * scale
                    assert np.array_equal(nn, first[2]), (njs, ipl, nchunk, smem)
    g6.set_tiles(0, 0)


def test_autotune_and_tuning_file(g6, tmp_path):
    """g6x_autotune times the shapes for ni = 1, 2, ..., ni_max against the
    loaded j-set and keeps one row per ni bucket, which the automatic plan
    then follows: results stay within the plain path's bars (nn exact), and
    a table saved and loaded again gives the same bits."""
    s = plummer_state(3000, 23)
    upload(g6, s["x"], s["v"], s["a2"], s["j6"], s["k18"], s["m"], s["tj"], s["ids"])
    g6.set_ti(s["ti"])
    xp, vp = oracle.predict(s["ti"], s["tj"], s["x"], s["v"], s["a2"], s["j6"], s["k18"])
    empty = str(tmp_path / "empty.txt")
    open(empty, "w").close()
    try:
        print_case("autotune nj=3000 ni_max=64")
        assert g6.autotune(3000, 64, 1e-4) > 0
        for ni in (1, 5, 64):
            sel = np.arange(0, 3000, 3000 // ni)[:ni]
            id_i, xi, vi = s["ids"][sel], xp[sel], vp[sel]
            exact = oracle.force_ld(id_i, xi, vi, 1e-4, s["ids"], s["m"], xp, vp)
            _, _, _, onn, _ = oracle.force(id_i, xi, vi, 1e-4, s["ids"], s["m"], xp, vp)
            print_case(f"tuned plan ni={ni}")
            got = g6.calc(3000, id_i, xi, vi, 1e-4)
            assert np.array_equal(got[3], onn), ni
            assert_plain_parity(got[0], got[1], got[2], exact)
            path = str(tmp_path / "tune.txt")
            assert g6.save_tuning(path) == 7      # buckets ni <= 1, 2, 4, ..., 64
            assert g6.load_tuning(empty) == 0
            assert g6.load_tuning(path) == 7
            again = g6.calc(3000, id_i, xi, vi, 1e-4)
            for p, q in zip(got[:4], again[:4]):
                assert np.array_equal(p, q), ni
    finally:
        g6.load_tuning(empty)


def test_host_io_boundaries(g6):
    """i-blocks and j-stages on both sides of the 4096-row limit below which
    the kernels read inputs and write results in pinned host memory
    themselves (above it the rows are copied), and a launch whose j-split
    would read each i-row too often in place (its rows are copied)."""
    n = 6000
    s = plummer_state(n, 27)
    # the first 4096 slots go in wrong (a 6000-record stage, copied), then
    # right (a 4096-record stage, read in place by the scatter kernel)
    wrong = s["x"].copy()
    wrong[:4096] += 0.25
    upload(g6, wrong, s["v"], s["a2"], s["j6"], s["k18"], s["m"], s["tj"], s["ids"])
    g6.set_ti(s["ti"])
    g6.calc(n, s["ids"][:1], s["x"][:1], s["v"][:1], 1e-4)  # predicted at ti
    sel = np.arange(4096)
    g6.set_j_particles(sel, s["ids"][sel], s["tj"][sel], 0.0, s["m"][sel], s["k18"][sel],
                       s["j6"][sel], s["a2"][sel], s["v"][sel], s["x"][sel])
    xp, vp = oracle.predict(s["ti"], s["tj"], s["x"], s["v"], s["a2"], s["j6"], s["k18"])
    for nj, ni in ((1024, 4096), (1024, 4097), (n, 64), (n, 2000), (n, 4097)):
        print_case(f"host i/o nj={nj} ni={ni}")
        isel = np.linspace(0, n - 1, ni).astype(int)
        id_i, xi, vi = s["ids"][isel], xp[isel], vp[isel]
        acc, jerk, pot, nn, _ = g6.calc(nj, id_i, xi, vi, 1e-4)
        jid, jm, jx, jv = s["ids"][:nj], s["m"][:nj], xp[:nj], vp[:nj]
        _, _, _, onn, _ = oracle.force(id_i, xi, vi, 1e-4, jid, jm, jx, jv)
        assert np.array_equal(nn, onn), (nj, ni)
        assert_plain_parity(acc, jerk, pot, oracle.force_ld(id_i, xi, vi, 1e-4, jid, jm, jx, jv))


def test_many_chunk_reduction(g6):
    """Small i-blocks against many j split over 16..79 chunks: the
    one-wave-per-i reduction keeps FP64 parity and the exact nn."""
    s = plummer_state(20000, 26)
    upload(g6, s["x"], s["v"], s["a2"], s["j6"], s["k18"], s["m"], s["tj"], s["ids"])
    g6.set_ti(s["ti"])
    xp, vp = oracle.predict(s["ti"], s["tj"], s["x"], s["v"], s["a2"], s["j6"], s["k18"])
    for ni in (1, 5, 64):
        sel = np.arange(0, 20000, 20000 // ni)[:ni]
        id_i, xi, vi = s["ids"][sel], xp[sel], vp[sel]
        exact = oracle.force_ld(id_i, xi, vi, 1e-4, s["ids"], s["m"], xp, vp)
        _, _, _, onn, _ = oracle.force(id_i, xi, vi, 1e-4, s["ids"], s["m"], xp, vp)
        for prec in (0, 1):
            g6.set_precision(prec)
            for nchunk in (0, 16, 40, 79):
                print_case(f"many chunks ni={ni} prec={prec} nchunk={nchunk}")
                if nchunk:
                    g6.set_tiles(256 if ni == 1 else 64, 1, nchunk, 0)
                acc, jerk, pot, nn, _ = g6.calc(20000, id_i, xi, vi, 1e-4)
                g6.set_tiles(0, 0)
                if prec == 0:
                    assert np.array_equal(nn, onn), (ni, nchunk)
                    assert_plain_parity(acc, jerk, pot, exact)
                else:
                    d = np.abs(acc - exact[0]).max(1)
                    assert (d <= 2e-6 * exact[3]).all(), (ni, nchunk)
    g6.set_precision(0)


def test_row_form_reduction(g6):
    """Small i-blocks against 40000 j: 157 one-tile chunks, folded by the
    reduce kernel's row form (one workgroup per row; from 128 chunks up to
    128 rows; 300 rows take the wave form); 100 forced chunks take the wave
    form.  FP64 parity and the exact nn either way; mixed within its
    bound."""
    n = 40000
    s = plummer_state(n, 29)
    upload(g6, s["x"], s["v"], s["a2"], s["j6"], s["k18"], s["m"], s["tj"], s["ids"])
    g6.set_ti(s["ti"])
    xp, vp = oracle.predict(s["ti"], s["tj"], s["x"], s["v"], s["a2"], s["j6"], s["k18"])
    for ni in (1, 7, 64, 300):
        sel = np.arange(0, n, n // ni)[:ni]
        id_i, xi, vi = s["ids"][sel], xp[sel], vp[sel]
        exact = oracle.force_ld(id_i, xi, vi, 1e-4, s["ids"], s["m"], xp, vp)
        _, _, _, onn, or2 = oracle.force(id_i, xi, vi, 1e-4, s["ids"], s["m"], xp, vp)
        for prec in (0, 1):
            g6.set_precision(prec)
            for nchunk in (0, 100):
                print_case(f"row form ni={ni} prec={prec} nchunk={nchunk}")
                if nchunk:
                    g6.set_tiles(256 if ni == 1 else 64, 1, nchunk, 0)
                acc, jerk, pot, nn, _ = g6.calc(n, id_i, xi, vi, 1e-4)
                # the fold has a fixed order: a second run gives the same bits
                again = g6.calc(n, id_i, xi, vi, 1e-4)
                g6.set_tiles(0, 0)
                for a, b in zip((acc, jerk, pot, nn), again[:4]):
                    assert np.array_equal(a, b), (ni, prec, nchunk)
                if prec == 0:
                    assert np.array_equal(nn, onn), (ni, nchunk)
                    assert_plain_parity(acc, jerk, pot, exact)
                else:
                    for val, ex, sm, name in ((acc, exact[0], exact[3], "acc"),
                                              (jerk, exact[1], exact[4], "jerk"),
                                              (pot, exact[2], exact[5], "pot")):
                        _, cond, _ = errors(val, ex, sm)
                        assert cond.max() < 2e-6, (name, ni, nchunk, cond.max())
                    assert_nn_near(nn, or2, xi, xp, s["ids"])
    g6.set_precision(0)


def assert_nn_near(nn, or2, xi, xp, ids, rtol=1e-5):
    """Mixed-path nearest neighbours: each chosen j lies at the minimal FP64
    r^2 up to FP32 rounding (near-ties may pick another j of that distance);
    a wrong choice (not the minimum, self, an empty slot) is far outside."""
    order = np.argsort(ids)
    slot = order[np.searchsorted(ids[order], nn)]
    assert np.array_equal(ids[slot], nn)
    d = xp[slot] - xi
    r2 = (d * d).sum(1)
    assert (r2 <= or2 * (1 + rtol)).all(), (r2 / or2).max()


def test_self_block_matches_abi_path(g6):
    s = plummer_state(2048, 23)
    upload(g6, s["x"], s["v"], s["a2"], s["j6"], s["k18"], s["m"], s["tj"], s["ids"])
    g6.set_ti(s["ti"])
    xp, vp = oracle.predict(s["ti"], s["tj"], s["x"], s["v"], s["a2"], s["j6"], s["k18"])
    acc, jerk, pot, nn, _ = g6.calc(2048, s["ids"][100:612], xp[100:612], vp[100:612], 1e-4)
    g6.calc_self(2048, 100, 512, 1e-4, nn=True)
    acc2, jerk2, pot2, nn2 = g6.fetch(512)
    assert np.array_equal(nn, nn2)
    np.testing.assert_array_equal(acc, acc2)
    np.testing.assert_array_equal(jerk, jerk2)
    np.testing.assert_array_equal(pot, pot2)


@pytest.mark.parametrize("prec,compat", [(0, False), (0, True), (1, False), (1, True)])
def test_predicted_on_load_matches_predictor(g6, prec, compat):
    """Short j-ranges: the ABI call's force kernel predicts the j-slots as it
    loads them; the same i-block through g6x_calc_self (the predictor
    kernel, then the same force shape) gives the same bits, nn included."""
    nj = 777  # a ragged last tile
    s = plummer_state(nj, 28)
    g6.set_precision(prec)
    g6.set_predictor_compat(compat)
    upload(g6, s["x"], s["v"], s["a2"], s["j6"], s["k18"], s["m"], s["tj"], s["ids"])
    g6.set_ti(s["ti"])
    xp, vp = oracle.predict(s["ti"], s["tj"], s["x"], s["v"], s["a2"], s["j6"], s["k18"],
                            compat=compat)
    lo, ni = 100, 300
    print_case(f"predicted on load prec={prec} compat={compat}")
    acc, jerk, pot, nn, _ = g6.calc(nj, s["ids"][lo:lo + ni], xp[lo:lo + ni], vp[lo:lo + ni], 1e-4)
    g6.calc_self(nj, lo, ni, 1e-4, nn=True)
    acc2, jerk2, pot2, nn2 = g6.fetch(ni)
    g6.set_predictor_compat(False)
    g6.set_precision(0)
    assert np.array_equal(nn, nn2)
    np.testing.assert_array_equal(acc, acc2)
    np.testing.assert_array_equal(jerk, jerk2)
    np.testing.assert_array_equal(pot, pot2)
    if prec == 0:
        _, _, _, onn, _ = oracle.force(s["ids"][lo:lo + ni], xp[lo:lo + ni], vp[lo:lo + ni],
                                       1e-4, s["ids"], s["m"], xp, vp)
        assert np.array_equal(nn, onn)


def test_deterministic_repeat(g6):
    s = plummer_state(5000, 24)
    upload(g6, s["x"], s["v"], s["a2"], s["j6"], s["k18"], s["m"], s["tj"], s["ids"])
    g6.set_ti(s["ti"])
    for prec in (0, 1):
        g6.set_precision(prec)
        g6.calc_self(5000, 0, 5000, 1e-4)
        r1 = g6.fetch(5000)
        g6.calc_self(5000, 0, 5000, 1e-4)
        r2 = g6.fetch(5000)
        for a, b in zip(r1, r2):
            assert np.array_equal(a, b)


def test_edge_cases(g6):
    s = plummer_state(64, 25)
    upload(g6, s["x"], s["v"], s["a2"], s["j6"], s["k18"], s["m"], s["tj"], s["ids"])
    g6.set_ti(s["ti"])
    xi = np.zeros((3, 3))
    vi = np.zeros((3, 3))
    # nj = 0: nothing interacts, nn = -1 (g6lib.c:112-121)
    acc, jerk, pot, nn, _ = g6.calc(0, [1, 2, 3], xi, vi, 1e-4)
    assert not acc.any() and not jerk.any() and not pot.any()
    assert list(nn) == [-1, -1, -1]
    # ni = 0 is a no-op
    out = g6.calc(64, np.zeros(0, np.int32), np.zeros((0, 3)), np.zeros((0, 3)), 1e-4)
    assert out[0].shape == (0, 3)
    # a slot far beyond the initial capacity grows the store
    far = 300000
    g6.set_j_particle(far, 777, s["ti"], 0.0, 0.5, [0, 0, 0], [0, 0, 0], [0, 0, 0], [0, 0, 0],
                      [10.0, 0, 0])
    acc, _, pot, nn, _ = g6.calc(far + 1, [-1], [[10.0, 0.0, 1.0]], [[0, 0, 0]], 0.0)
    assert nn[0] == 777
    xp, vp = oracle.predict(s["ti"], s["tj"], s["x"], s["v"], s["a2"], s["j6"], s["k18"])
    p_far = -0.5 / 1.0
    p_rest = -(s["m"] / np.linalg.norm(xp - [10.0, 0, 1.0], axis=1)).sum()
    assert pot[0] == pytest.approx(p_far + p_rest, rel=1e-13)
    # reset empties every slot
    g6.reset()
    acc, jerk, pot, nn, _ = g6.calc(far + 1, [-1], [[10.0, 0.0, 1.0]], [[0, 0, 0]], 0.0)
    assert pot[0] == 0.0 and nn[0] == -1


def test_error_paths(g6):
    import ctypes as C
    from amuse_amd.g6 import G6Error
    lib = g6.lib
    with pytest.raises(G6Error):
        g6.set_j_particle(-1, 1, 0, 0, 1, [0] * 3, [0] * 3, [0] * 3, [0] * 3, [0] * 3)
    # lasthalf without firsthalf (calc_self clears the latch)
    g6.calc_self(0, 0, 0, 0.0)
    idx = np.zeros(1, np.int32)
    buf = np.zeros((1, 3))
    D = C.POINTER(C.c_double)
    rc = lib.g6calc_lasthalf(0, 1, 1, idx.ctypes.data_as(C.POINTER(C.c_int)), buf.ctypes.data_as(D),
                             buf.ctypes.data_as(D), 0.0, buf.ctypes.data_as(D),
                             buf.ctypes.data_as(D), buf.ctypes.data_as(D), buf.ctypes.data_as(D))
    assert rc == -4
    g6.close()
    assert lib.g6_set_ti(0, 1.0) == -2
    # bad device id
    assert lib.g6_open(99) < 0
    g6.open()


def test_open_close_cycles_reset_state(g6):
    s = plummer_state(128, 26)
    for cycle in range(3):
        upload(g6, s["x"], s["v"], s["a2"], s["j6"], s["k18"], s["m"], s["tj"], s["ids"])
        g6.set_ti(s["ti"])
        acc, jerk, pot, nn, _ = g6.calc(128, s["ids"][:4], s["x"][:4], s["v"][:4], 1e-4)
        assert (nn > 0).all()
        g6.close()
        g6.open()
        acc, jerk, pot, nn, _ = g6.calc(128, s["ids"][:4], s["x"][:4], s["v"][:4], 1e-4)
        assert not pot.any() and (nn == -1).all()


def test_mixed_precision_error_bounds(g6):
    """Mixed path vs the 80-bit oracle: per-particle error relative to
    sum|term| below 2e-6, median below 2e-7; nn agrees except near-ties."""
    s = plummer_state(8192, 27)
    upload(g6, s["x"], s["v"], s["a2"], s["j6"], s["k18"], s["m"], s["tj"], s["ids"])
    g6.set_ti(s["ti"])
    g6.set_precision(1)
    xp, vp = oracle.predict(s["ti"], s["tj"], s["x"], s["v"], s["a2"], s["j6"], s["k18"])
    sel = np.arange(0, 8192, 5)
    id_i, xi, vi = s["ids"][sel], xp[sel], vp[sel]
    acc, jerk, pot, nn, _ = g6.calc(8192, id_i, xi, vi, 1e-4)
    ea, ej, ep, sa, sj, sp = oracle.force_ld(id_i, xi, vi, 1e-4, s["ids"], s["m"], xp, vp)
    _, _, _, onn, or2 = oracle.force(id_i, xi, vi, 1e-4, s["ids"], s["m"], xp, vp)
    stats = {}
    for val, ex, sm, name in ((acc, ea, sa, "acc"), (jerk, ej, sj, "jerk"), (pot, ep, sp, "pot")):
        rel, cond, _ = errors(val, ex, sm)
        stats[name] = (cond.max(), np.median(cond), np.median(rel))
        assert cond.max() < 2e-6, (name, cond.max())
        assert np.median(cond) < 2e-7, (name, np.median(cond))
    print("mixed error (max cond, median cond, median rel):", stats)
    assert (nn == onn).mean() > 0.99
