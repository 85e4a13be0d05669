"""Pins the ensemble oracle (oracle/ensemble.py) to what the mathematics
fixes, independently of its own formulas:

* collision time: for massless bodies in linear motion it is the crossing
  time r/|v|; for bodies at rest it is the free-fall scale sqrt(r^3/M)
  (src/amuse_hermite/interface.cc:496-524 builds it from exactly these two);
  the smaller of the two wins;
* step rule: a power of two, no larger than dt_param x collision time, and a
  divisor of the current time;
* integrator: a circular binary (period 2 pi) stays on its analytic orbit
  and the error falls as dt^4 (the scheme is 4th order);
* energy of two bodies: the closed form.
"""
import math

import numpy as np
import pytest

from oracle import ensemble as oens


def test_collision_time_linear_motion():
    m = np.zeros(2)
    x = np.array([[0.0, 0.0, 0.0], [0.3, -1.2, 0.4]])
    v = np.array([[0.0, 0.0, 0.0], [2.0, 1.0, -2.0]])
    r = math.sqrt(0.09 + 1.44 + 0.16)
    q = oens.collision_qmax(m, x, v, 0.0)
    assert q ** -0.25 == pytest.approx(r / 3.0, rel=1e-6)


def test_collision_time_free_fall():
    m = np.array([0.25, 0.5])
    x = np.array([[1.0, 0.0, 0.0], [1.0, 2.0, 0.0]])
    v = np.zeros((2, 3))
    q = oens.collision_qmax(m, x, v, 0.0)
    assert q ** -0.25 == pytest.approx(math.sqrt(2.0 ** 3 / 0.75), rel=1e-6)
    # softening lengthens it: s = r^2 + eps2 replaces r^2 in the scale
    q2 = oens.collision_qmax(m, x, v, 1.0)
    assert q2 < q


def test_collision_time_takes_the_shorter():
    m = np.array([0.5, 0.5])
    x = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0]])
    for speed in (0.1, 10.0):
        v = np.array([[0.0, 0.0, 0.0], [0.0, speed, 0.0]])
        q = oens.collision_qmax(m, x, v, 0.0)
        assert q ** -0.25 == pytest.approx(min(1.0 / speed, 1.0), rel=1e-6)


@pytest.mark.parametrize("t", [0.0, 0.5, 0.375, 0.0625 * 3])
def test_step_rule(t):
    for q in (1e-3, 1.0, 37.0, 1e6):
        s = oens.next_step(t, q, 0.03)
        assert math.frexp(s)[0] == 0.5  # a power of two
        assert s <= 0.03 * q ** -0.25 or s == 2.0 ** -40
        assert s <= 0.03125
        assert math.fmod(t, s) == 0.0
        if s < 0.03125 and math.fmod(t, 2 * s) == 0.0:
            assert 2 * s > 0.03 * q ** -0.25  # largest admissible


def circular_binary():
    m = np.array([0.5, 0.5])
    x = np.array([[-0.5, 0.0, 0.0], [0.5, 0.0, 0.0]])
    v = np.array([[0.0, -0.5, 0.0], [0.0, 0.5, 0.0]])
    return m, x, v


def test_circular_binary_fourth_order():
    m, x, v = circular_binary()
    t_end = 1.0
    exact = np.array([0.5 * math.cos(t_end), 0.5 * math.sin(t_end), 0.0])
    errs = []
    for dtp in (0.03, 0.015):
        xf, vf, tf, steps = oens.integrate_cluster(m, x, v, 0.0, dtp, t_end)
        assert tf == t_end
        errs.append(np.abs(xf[1] - exact).max())
    ratio = errs[0] / errs[1]
    assert errs[0] < 1e-8
    assert 10.0 < ratio < 24.0, (errs, ratio)


def test_two_body_energy():
    m, x, v = circular_binary()
    e = oens.energy(m, x, v, 0.01)
    expect = 0.5 * 0.5 * 0.25 * 2 - 0.25 / math.sqrt(1.0 + 0.01)
    assert e == pytest.approx(expect, rel=1e-14)
