"""The CPU oracle, pinned before it is trusted.

* against the golden vectors the compiled reference lib/g6lib/g6lib.c wrote
  (tests/golden/make_golden.py), bit for bit;
* against the compiled reference itself on fresh random inputs, when
  oracle/_ref/libg6ref.so is present;
* against closed forms and invariants that a plausible mistake (a dropped
  term, a wrong sign or power, a transposed operand) would break:
  two-body potential and field, the ph4 field-at-point potentials
  (src/amuse_ph4/tests/test_ph4.py:786-823), three bodies on a line
  (src/tests/codes_tests/gd_tests.py:144-177), Newton's third law
  (sum m a = sum m jerk = 0), translation and Galilean invariance, the
  derivative relation jerk = d(acc)/dt along the motion, and the exact
  Taylor polynomial of the predictor.
"""
import os

import numpy as np
import pytest

import oracle
from conftest import load_golden

CASES = ["g6ref_dt0", "g6ref_dt", "g6ref_lattice", "g6ref_special", "g6ref_probe"]


def golden_inputs(d):
    """Slot arrays as the reference saw them (never-written slots id -1)."""
    ids = d["id_j"].copy()
    ids[ids == -2] = -1
    nj = int(d["nj"])
    return ids[:nj], nj


def oracle_on_golden(d, compat=True):
    ids, nj = golden_inputs(d)
    xp, vp = oracle.predict(float(d["ti"]), d["tj"][:nj], d["x"][:nj], d["v"][:nj],
                            d["a2"][:nj], d["j6"][:nj], d["k18"][:nj], compat=compat)
    m = d["m"][:nj].copy()
    never = d["id_j"][:nj] == -2
    # never-written slots hold zeros in the reference (no predictor input)
    xp[never] = 0.0
    vp[never] = 0.0
    m[never] = 0.0
    return oracle.force(d["id_i"], d["xi"], d["vi"], float(d["eps2"]), ids, m, xp, vp)


@pytest.mark.parametrize("name", CASES)
def test_oracle_matches_reference_golden_bitwise(name):
    d = load_golden(name)
    acc, jerk, pot, nn, _ = oracle_on_golden(d, compat=True)
    assert np.array_equal(nn, d["nn"])
    assert np.array_equal(acc, d["acc"])
    assert np.array_equal(jerk, d["jerk"])
    assert np.array_equal(pot, d["pot"])
    assert np.array_equal(d["index_out"], d["id_i"])


def test_golden_probe_is_ph4_known_answer():
    d = load_golden("g6ref_probe")
    # test_ph4.py:803
    expect = np.array([-4.57110767688, -2.66662518639, -2.13331892165])
    np.testing.assert_allclose(d["pot"][:3], expect, rtol=1e-11)
    np.testing.assert_allclose(d["pot"][3:], expect, rtol=1e-11)
    # closed form -1/sqrt(x^2+eps2) - 1/sqrt((2-x)^2+eps2)
    x = d["xi"][:, 0]
    e = float(d["eps2"])
    np.testing.assert_allclose(d["pot"], -1 / np.sqrt(x * x + e) - 1 / np.sqrt((2 - x) ** 2 + e),
                               rtol=1e-14)


@pytest.mark.skipif(not os.path.exists(oracle.REF_SO), reason="reference g6lib not compiled here")
@pytest.mark.parametrize("seed", [3, 4])
def test_oracle_matches_compiled_reference_random(seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    n = 300
    ref = oracle.RefG6()
    ids = rng.permutation(np.arange(1, n + 1)).astype(np.int32)
    x, v = rng.normal(size=(n, 3)), rng.normal(size=(n, 3))
    a2, j6, k18 = rng.normal(size=(n, 3)), rng.normal(size=(n, 3)), rng.normal(size=(n, 3))
    m = rng.uniform(0.1, 1, n)
    tj = rng.uniform(0, 0.1, n)
    ti = 0.125
    ref.open()
    for j in range(n):
        ref.set_j_particle(j, ids[j], tj[j], 0.0, m[j], k18[j], j6[j], a2[j], v[j], x[j])
    ref.set_ti(ti)
    ni = 77
    id_i = ids[rng.integers(0, n, ni)]
    xi, vi = rng.normal(size=(ni, 3)), rng.normal(size=(ni, 3))
    racc, rjerk, rpot, rnn, _ = ref.calc(n, id_i, xi, vi, 0.01)
    ref.close()
    xp, vp = oracle.predict(ti, tj, x, v, a2, j6, k18, compat=True)
    acc, jerk, pot, nn, _ = oracle.force(id_i, xi, vi, 0.01, ids, m, xp, vp)
    assert np.array_equal(nn, rnn)
    assert np.array_equal(acc, racc) and np.array_equal(jerk, rjerk) and np.array_equal(pot, rpot)


def test_predictor_taylor_closed_form():
    rng = np.random.Generator(np.random.PCG64(5))
    n = 50
    x, v, a2, j6, k18 = (rng.normal(size=(n, 3)) for _ in range(5))
    tj = rng.uniform(-1, 0, n)
    ti = 0.3
    dt = (ti - tj)[:, None]
    xp, vp = oracle.predict(ti, tj, x, v, a2, j6, k18, compat=False)
    # x(t) = x + v dt + a/2 dt^2 + j/6 dt^3 + s/24 dt^4 with a2=a/2, j6=j/6, k18=s/18
    acc, jerk, snap = 2 * a2, 6 * j6, 18 * k18
    np.testing.assert_allclose(xp, x + v * dt + acc / 2 * dt ** 2 + jerk / 6 * dt ** 3
                               + snap / 24 * dt ** 4, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(vp, v + acc * dt + jerk / 2 * dt ** 2 + snap / 6 * dt ** 3,
                               rtol=1e-13, atol=1e-13)
    # compat polynomial (g6lib.c:85-87): dt^3 and dt^4 in the velocity
    xc, vc = oracle.predict(ti, tj, x, v, a2, j6, k18, compat=True)
    assert np.array_equal(xc, xp)
    np.testing.assert_allclose(vc, v + acc * dt + jerk / 2 * dt ** 3 + snap / 3 * dt ** 4,
                               rtol=1e-13, atol=1e-13)


def _random_system(n, seed, eps2=1e-3):
    rng = np.random.Generator(np.random.PCG64(seed))
    x = rng.normal(size=(n, 3))
    v = rng.normal(size=(n, 3))
    m = rng.uniform(0.5, 2.0, n) / n
    ids = np.arange(n, dtype=np.int32)
    return ids, x, v, m, eps2


def test_two_body_closed_form():
    ids = np.array([1, 2], np.int32)
    x = np.array([[0.0, 0, 0], [2.0, 0, 0]])
    v = np.array([[0.0, 0, 0], [10.0, 0, 0]])
    m = np.array([10.0, 10.0])
    eps2 = 0.01
    acc, jerk, pot, nn, r2 = oracle.force(ids, x, v, eps2, ids, m, x, v)
    s = 4.0 + eps2
    # test_ph4.py:34-52: potential -10/sqrt(2^2 + 0.1^2)
    np.testing.assert_allclose(pot, [-10 / np.sqrt(s)] * 2, rtol=1e-15)
    np.testing.assert_allclose(acc[0], [10 * 2 / s ** 1.5, 0, 0], rtol=1e-15)
    np.testing.assert_allclose(acc[1], [-10 * 2 / s ** 1.5, 0, 0], rtol=1e-15)
    # jerk = m (v/s^1.5 - 3 (r.v) r / s^2.5), r = (2,0,0), v = (10,0,0)
    jx = 10 * (10 / s ** 1.5 - 3 * 20 * 2 / s ** 2.5)
    np.testing.assert_allclose(jerk[0], [jx, 0, 0], rtol=1e-14)
    assert list(nn) == [2, 1]
    np.testing.assert_allclose(r2, [4.0, 4.0])


def test_three_body_line_energy():
    # gd_tests.py:144-177 / test_phigrape.py:96-126: unit masses at x=-1,0,1,
    # eps2=0, vx = 0, 1, 0 -> Ep = -2.5, Ek = 0.5
    ids = np.array([1, 2, 3], np.int32)
    x = np.array([[-1.0, 0, 0], [0.0, 0, 0], [1.0, 0, 0]])
    v = np.array([[0.0, 0, 0], [1.0, 0, 0], [0.0, 0, 0]])
    m = np.ones(3)
    _, _, pot, _, _ = oracle.force(ids, x, v, 0.0, ids, m, x, v)
    assert 0.5 * (m * pot).sum() == pytest.approx(-2.5, rel=1e-15)
    assert 0.5 * (m * (v * v).sum(1)).sum() == 0.5
    # delete the middle particle (id -1 slot): -0.5
    idj = ids.copy()
    idj[1] = -1
    _, _, pot, _, _ = oracle.force(ids[[0, 2]], x[[0, 2]], v[[0, 2]], 0.0, idj, m, x, v)
    assert 0.5 * pot.sum() == pytest.approx(-0.5, rel=1e-15)


def test_newton_third_law_and_energy_pairs():
    ids, x, v, m, eps2 = _random_system(200, 6)
    acc, jerk, pot, _, _ = oracle.force(ids, x, v, eps2, ids, m, x, v)
    scale = (np.abs(m[:, None] * acc)).sum()
    assert np.abs((m[:, None] * acc).sum(0)).max() < 1e-13 * scale
    jscale = (np.abs(m[:, None] * jerk)).sum()
    assert np.abs((m[:, None] * jerk).sum(0)).max() < 1e-13 * jscale
    # 1/2 sum m phi equals the explicit pair sum
    d = x[:, None, :] - x[None, :, :]
    r = np.sqrt((d * d).sum(-1) + eps2)
    iu = np.triu_indices(200, 1)
    U = -(m[:, None] * m[None, :] / r)[iu].sum()
    assert 0.5 * (m * pot).sum() == pytest.approx(U, rel=1e-13)


def test_translation_and_boost_invariance():
    ids, x, v, m, eps2 = _random_system(128, 7)
    base = oracle.force_ld(ids, x, v, eps2, ids, m, x, v)
    shift, boost = np.array([3.0, -2.0, 0.5]), np.array([0.7, 0.1, -4.0])
    moved = oracle.force_ld(ids, x + shift, v + boost, eps2, ids, m, x + shift, v + boost)
    for a, b in zip(base[:3], moved[:3]):
        np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-12 * np.abs(a).max())


def test_jerk_is_time_derivative_of_acc():
    # free straight-line motion: acc(t) along x + v t has d/dt equal to jerk
    ids, x, v, m, eps2 = _random_system(64, 8, eps2=0.05)
    h = 1e-5
    ap = oracle.force_ld(ids, x + v * h, v, eps2, ids, m, x + v * h, v)[0]
    am = oracle.force_ld(ids, x - v * h, v, eps2, ids, m, x - v * h, v)[0]
    jerk = oracle.force_ld(ids, x, v, eps2, ids, m, x, v)[1]
    np.testing.assert_allclose((ap - am) / (2 * h), jerk, rtol=1e-6, atol=1e-6 * np.abs(jerk).max())


def test_long_double_agrees_with_pow_path():
    ids, x, v, m, eps2 = _random_system(300, 9)
    a, j, p, _, _ = oracle.force(ids, x, v, eps2, ids, m, x, v)
    al, jl, pl, sa, sj, sp = oracle.force_ld(ids, x, v, eps2, ids, m, x, v)
    assert (np.linalg.norm(a - al, axis=1) / sa).max() < 1e-14
    assert (np.linalg.norm(j - jl, axis=1) / sj).max() < 1e-14
    assert (np.abs(p - pl) / sp).max() < 1e-14


def test_brute_force_fractions_tiny():
    # exact rational arithmetic where sqrt is exact: eps2 chosen so that
    # r^2 + eps2 are perfect squares
    from fractions import Fraction as F
    ids = np.array([1, 2, 3], np.int32)
    x = np.array([[0.0, 0, 0], [3.0, 0, 0], [0.0, 4.0, 0]])
    v = np.zeros((3, 3))
    m = np.array([1.0, 2.0, 3.0])
    eps2 = 16.0     # s = 25 (d=3), 32? -> use pairs 1-2 (9+16=25) and 1-3 (16+16=32)
    acc, _, pot, nn, _ = oracle.force(ids[:1], x[:1], v[:1], eps2, ids, m, x, v)
    # particle 1: from 2 at r^2=9 -> s=25, from 3 at r^2=16 -> s=32
    expect_pot = -(F(2) / 5) - 3 / np.sqrt(32.0)
    assert pot[0] == pytest.approx(float(expect_pot), rel=1e-15)
    assert acc[0, 0] == pytest.approx(2 * 3 / 125, rel=1e-15)
    assert nn[0] == 2
