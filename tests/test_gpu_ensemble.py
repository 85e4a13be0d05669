"""Batched ensembles on the GPU against the oracle (oracle/ensemble.py).

* forces of every particle of every cluster: FP64 within 1e-15 sum|term| of
  the 80-bit oracle, mixed within 2e-6 sum|term|;
* FP64 integration of clusters in several size classes (so several lanes-
  per-i shapes) follows the oracle: identical step counts, states within
  1e-10 after 1/16 time unit; also with shapes whose LDS tiles are mostly
  zero-mass padding (64 lanes per i over 16- or 64-entry tiles);
* a hipGraph replay equals eager launches bit for bit, replay after replay,
  and so does the whole run in one persistent cooperative launch;
* every tile shape (lanes per i, j per LDS tile) gives the same FP64 forces
  to rounding, and mixed forces within the mixed bound of the oracle;
* mixed-precision energy error stays within a small factor of FP64's;
* clusters of 1, 2, 3 and 17 particles: FP64 follows the oracle, a lone
  particle moves in a straight line.
"""
import numpy as np
import pytest

import oracle
from oracle import ensemble as oens
from amuse_amd.ensemble import Ensemble, make_ensemble
from conftest import print_case

pytestmark = pytest.mark.gpu

EPS2 = 1e-4
DTP = 0.03       # hermite dt_param (src/amuse_hermite/interface.cc:86)
MAX_ROUNDS = 200000


def small_ensemble():
    # sizes across classes <=16, <=32, <=128, <=512 and one above 2048
    counts = np.array([12, 16, 31, 100, 257, 2100], dtype=np.int32)
    from amuse_amd import ic
    rng = np.random.Generator(np.random.PCG64(9))
    ms, xs, vs = [], [], []
    for n in counts:
        mm = ic.kroupa(int(n), rng, 0.08, 20.0)
        mm /= mm.sum()
        _, x, v = ic.plummer_raw(int(n), rng)
        x, v = ic.to_nbody_units(x, v)
        x, v = ic.center(mm, x, v)
        ms.append(mm); xs.append(x); vs.append(v)
    return counts, np.concatenate(ms), np.concatenate(xs), np.concatenate(vs)


def per_cluster(counts):
    off = np.concatenate([[0], np.cumsum(counts)])
    return [slice(off[c], off[c + 1]) for c in range(len(counts))]


def test_create_refuses_far_or_nonfinite_coordinates():
    """The force kernel pads short j-tiles with zero-mass entries at 1e17,
    so creation refuses coordinates beyond 1e15 (and NaN/inf); 1e15 itself
    and a valid ensemble are accepted."""
    from amuse_amd.g6 import G6Error
    counts, m, x, v = small_ensemble()
    for bad in (2e15, -1e16, np.nan, np.inf):
        xb = x.copy()
        xb[40, 1] = bad
        with pytest.raises(G6Error):
            Ensemble(counts, m, xb, v, eps2=EPS2, dt_param=DTP, precision=1)
    xb = x.copy()
    xb[40, 1] = 1e15
    Ensemble(counts, m, xb, v, eps2=EPS2, dt_param=DTP, precision=1).close()


def test_ensemble_forces():
    counts, m, x, v = small_ensemble()
    ens = Ensemble(counts, m, x, v, eps2=EPS2, dt_param=DTP, precision=0)
    a64, j64, p64 = ens.forces(0)
    amx, jmx, pmx = ens.forces(1)
    for sl in per_cluster(counts):
        ids = np.arange(sl.stop - sl.start, dtype=np.int32)
        ea, ej, ep, sa, sj, sp = oracle.force_ld(ids, x[sl], v[sl], EPS2, ids, m[sl], x[sl], v[sl])
        n = sl.stop - sl.start
        for val, ex, s, tol, name in ((a64[sl], ea, sa, 1e-15, "acc64"), (j64[sl], ej, sj, 4e-15, "jerk64"),
                                      (p64[sl], ep, sp, 1e-15, "pot64"), (amx[sl], ea, sa, 2e-6, "accmx"),
                                      (jmx[sl], ej, sj, 1e-5, "jerkmx"), (pmx[sl], ep, sp, 2e-6, "potmx")):
            d = np.abs(val - ex).reshape(len(s), -1).max(1)
            print(f"cluster n={n} {name}: max |d|/sum|term| {(d / s).max():.3e}", flush=True)
            assert (d <= tol * s).all(), (n, name, (d / s).max())
        # Newton's third law within each cluster: sum m a = sum m jerk = 0
        mm = m[sl][:, None]
        # (measured: FP64 <= 1.1e-16, mixed <= 4.6e-8 of the sums of magnitudes)
        for val, tol, name in ((a64[sl], 1e-14, "acc64"), (j64[sl], 1e-14, "jerk64"),
                               (amx[sl], 1e-6, "accmx"), (jmx[sl], 1e-6, "jerkmx")):
            r = np.linalg.norm((mm * val).sum(0)) / (mm[:, 0] * np.linalg.norm(val, axis=1)).sum()
            print(f"cluster n={n} {name}: |sum m q|/sum|m q| {r:.2e}", flush=True)
            assert r <= tol, (n, name, r)
    ens.close()


@pytest.mark.parametrize("shape", [None, (64, 16), (64, 64), (8, 64)])
def test_ensemble_fp64_integration_matches_oracle(shape):
    """Also with lanes per i / j-tile shapes whose tiles are mostly padding
    (every lane runs the same number of j; the entries past the tile's last
    j are zero-mass padding that must change no sum and no step)."""
    counts, m, x, v = small_ensemble()
    counts, m, x, v = counts[:5], m[:counts[:5].sum()], x[:counts[:5].sum()], v[:counts[:5].sum()]
    ens = Ensemble(counts, m, x, v, eps2=EPS2, dt_param=DTP, precision=0)
    if shape is not None:
        ens.set_tiles([shape[0]] * 8)
        ens.set_jtiles([shape[1]] * 8)
    ens.start(0.0625)
    ens.run(steps_per_graph=4, use_graph=True, max_rounds=MAX_ROUNDS)
    xg, vg, tg, ng = ens.state()
    ens.close()
    for c, sl in enumerate(per_cluster(counts)):
        print_case(f"ensemble cluster {c} n={counts[c]} shape {shape}")
        xo, vo, to, so = _oracle_run(c, m[sl], x[sl], v[sl])
        assert tg[c] == to == 0.0625
        assert ng[c] == so, (c, ng[c], so)
        assert np.abs(xg[sl] - xo).max() < 1e-10
        assert np.abs(vg[sl] - vo).max() < 1e-9


_ORACLE_RUNS = {}


def _oracle_run(c, m, x, v):
    if c not in _ORACLE_RUNS:
        _ORACLE_RUNS[c] = oens.integrate_cluster(m, x, v, EPS2, DTP, 0.0625)
    return _ORACLE_RUNS[c]


def test_graph_replay_equals_eager():
    counts, m, x, v = small_ensemble()
    print_case("ensemble graph vs eager, mixed, t=1/8")
    res = []
    for use_graph in (True, False):
        ens = Ensemble(counts, m, x, v, eps2=EPS2, dt_param=DTP, precision=1)
        ens.start(0.125)
        rounds = ens.run(steps_per_graph=3, use_graph=use_graph, max_rounds=MAX_ROUNDS)
        res.append((ens.state(), rounds))
        ens.close()
    (sa, ra), (sb, rb) = res
    assert ra == rb
    for p, q in zip(sa, sb):
        assert np.array_equal(p, q)


@pytest.mark.parametrize("prec", [1, 0])
def test_persistent_run_equals_eager(prec):
    """The whole run in one persistent cooperative launch (phases separated
    by grid barriers, work list rebuilt on the device as clusters retire)
    equals eager launches that check for a rebuild after every round, bit
    for bit: the same lanes per i for every cluster at every round, so the
    same forces, steps and states; also when the run is split into
    launches of a few rounds each."""
    counts, m, x, v = small_ensemble()
    res = {}
    for mode in ("eager", "persistent", "persistent-chunks"):
        print_case(f"ensemble {mode} prec={prec} t=1/8")
        ens = Ensemble(counts, m, x, v, eps2=EPS2, dt_param=DTP, precision=prec)
        ens.start(0.125)
        if mode == "eager":
            rounds = ens.run(steps_per_graph=1, use_graph=False, max_rounds=MAX_ROUNDS)
        elif mode == "persistent":
            rounds = ens.run(persistent=True, max_rounds=MAX_ROUNDS)
        else:
            rounds = 0
            while True:
                r = ens.run(persistent=True, max_rounds=7)
                rounds += r
                if r < 7:
                    break
                assert rounds < MAX_ROUNDS
        res[mode] = (ens.state(), rounds)
        ens.close()
    (se, re_), (sp, rp), (sc, rc) = res["eager"], res["persistent"], res["persistent-chunks"]
    assert re_ == rp == rc, (re_, rp, rc)
    assert (se[2] == 0.125).all()
    for p, q, r in zip(se, sp, sc):
        assert np.array_equal(p, q)
        assert np.array_equal(p, r)


def test_persistent_run_many_clusters():
    """256 clusters of 16..512 particles in mixed precision to t = 1/16: the
    persistent launch equals eager launches bit for bit with autotuned lanes
    per i and j-tile lengths, and every cluster finishes."""
    counts, m, x, v = make_ensemble(256, 17, nmin=16, nmax=512)
    res = []
    tiles = None
    for mode in ("eager", "persistent"):
        print_case(f"ensemble 256 clusters {mode}")
        ens = Ensemble(counts, m, x, v, eps2=1e-6, dt_param=DTP, precision=1)
        if tiles is None:
            assert ens.autotune() > 0
            tiles = (ens.tiles(), ens.jtiles())
        else:
            ens.set_tiles(tiles[0])
            ens.set_jtiles(tiles[1])
        ens.start(0.0625)
        if mode == "eager":
            rounds = ens.run(steps_per_graph=1, use_graph=False, max_rounds=MAX_ROUNDS)
        else:
            rounds = ens.run(persistent=True, max_rounds=MAX_ROUNDS)
        res.append((ens.state(), rounds))
        ens.close()
    print("tiles", tiles, "rounds", res[0][1], flush=True)
    assert res[0][1] == res[1][1]
    assert (res[1][0][2] == 0.0625).all()
    for p, q in zip(res[0][0], res[1][0]):
        assert np.array_equal(p, q)


def test_tile_shapes_agree():
    counts, m, x, v = small_ensemble()
    ens = Ensemble(counts, m, x, v, eps2=EPS2, dt_param=DTP, precision=0)
    ref = ens.forces(0)
    for njs in (1, 2, 4, 8, 16, 32, 64):
        ens.set_tiles([njs] * 8)
        got = ens.forces(0)
        for g, r in zip(got, ref):
            assert np.abs(g - r).max() <= 1e-13 * np.abs(r).max()
    for tj in (64, 128):
        ens.set_tiles([4] * 8)
        ens.set_jtiles([tj] * 8)
        got = ens.forces(0)
        for g, r in zip(got, ref):
            assert np.abs(g - r).max() <= 1e-13 * np.abs(r).max()
    ens.set_jtiles([256] * 8)
    assert ens.autotune() > 0
    tuned = ens.forces(0)
    assert np.abs(tuned[0] - ref[0]).max() <= 1e-13 * np.abs(ref[0]).max()
    ens.close()


def test_tile_shapes_mixed_within_bound():
    """Mixed forces of every lanes-per-i / j-tile shape within the mixed
    bound of the 80-bit oracle (tiles whose last entries are padding
    included: clusters of 12..2100 against 1..64 lanes per i)."""
    counts, m, x, v = small_ensemble()
    ens = Ensemble(counts, m, x, v, eps2=EPS2, dt_param=DTP, precision=1)
    exact = []
    for sl in per_cluster(counts):
        ids = np.arange(sl.stop - sl.start, dtype=np.int32)
        exact.append(oracle.force_ld(ids, x[sl], v[sl], EPS2, ids, m[sl], x[sl], v[sl]))
    for njs, tj in [(n, 256) for n in (1, 2, 4, 8, 16, 32, 64)] + [(64, 16), (16, 64), (2, 128)]:
        ens.set_tiles([njs] * 8)
        ens.set_jtiles([tj] * 8)
        amx, jmx, pmx = ens.forces(1)
        for sl, (ea, ej, ep, sa, sj, sp) in zip(per_cluster(counts), exact):
            for val, ex, s, tol in ((amx[sl], ea, sa, 2e-6), (jmx[sl], ej, sj, 1e-5), (pmx[sl], ep, sp, 2e-6)):
                d = np.abs(val - ex).reshape(len(s), -1).max(1)
                assert (d <= tol * s).all(), (njs, tj, sl, (d / s).max())
    ens.close()


def test_mixed_energy_error_vs_fp64():
    counts, m, x, v = make_ensemble(64, 5, nmin=16, nmax=256)
    de = {}
    for prec in (0, 1):
        ens = Ensemble(counts, m, x, v, eps2=1e-4, dt_param=DTP, precision=prec)
        e0 = ens.energy()
        ens.start(0.25)
        ens.run(steps_per_graph=8, max_rounds=MAX_ROUNDS)
        _, _, tc, _ = ens.state()
        assert (tc == 0.25).all()
        e1 = ens.energy()
        de[prec] = np.abs((e1 - e0) / e0)
        ens.close()
    print("median |dE/E| fp64", np.median(de[0]), "mixed", np.median(de[1]))
    assert np.median(de[0]) < 1e-7
    assert np.median(de[1]) < 10 * np.median(de[0]) + 1e-8


def test_tiny_clusters():
    """Clusters of 1, 2, 3 and 17 particles (single tiles that are mostly
    padding, a lone particle with no pair at all): FP64 integration follows
    the oracle (same step counts, states within 1e-10 after 1/16), the lone
    particle moves in a straight line at dt_limit steps, and the mixed run
    finishes with finite states."""
    from amuse_amd import ic
    rng = np.random.Generator(np.random.PCG64(21))
    counts = np.array([1, 2, 3, 17], dtype=np.int32)
    ms, xs, vs = [], [], []
    for n in counts:
        mm = rng.uniform(0.5, 1.5, int(n))
        mm /= mm.sum()
        if n == 1:
            x, v = np.array([[0.3, -0.2, 0.1]]), np.array([[0.5, 0.25, -0.125]])
        else:
            _, x, v = ic.plummer_raw(int(n), rng)
            x, v = ic.to_nbody_units(x, v)
            x, v = ic.center(mm, x, v)
        ms.append(mm); xs.append(x); vs.append(v)
    m, x, v = np.concatenate(ms), np.concatenate(xs), np.concatenate(vs)
    t_end = 0.0625
    for prec in (0, 1):
        print_case(f"tiny clusters prec={prec}")
        ens = Ensemble(counts, m, x, v, eps2=EPS2, dt_param=DTP, precision=prec)
        ens.start(t_end)
        ens.run(steps_per_graph=4, use_graph=False, max_rounds=MAX_ROUNDS)
        xg, vg, tg, ng = ens.state()
        ens.close()
        assert (tg == t_end).all()
        assert np.isfinite(xg).all() and np.isfinite(vg).all()
        # the lone particle: no force, steps of dt_limit = 1/32
        assert ng[0] == 2
        assert np.abs(xg[0] - (x[0] + v[0] * t_end)).max() < 1e-15
        assert np.array_equal(vg[0], v[0])
        if prec == 0:
            for c, sl in enumerate(per_cluster(counts)):
                xo, vo, to, so = oens.integrate_cluster(m[sl], x[sl], v[sl], EPS2, DTP, t_end)
                assert ng[c] == so, (c, ng[c], so)
                assert np.abs(xg[sl] - xo).max() < 1e-10
                assert np.abs(vg[sl] - vo).max() < 1e-9
