"""CPU oracle of the batched-ensemble integrator — TEST INFRASTRUCTURE ONLY.

Numpy restatement, cluster by cluster, of the scheme in
amuse_amd/csrc/g6_clusters.hip, i.e. the shared-step Hermite code
src/amuse_hermite/interface.cc:
  predictor                 :266-281
  collision-time estimate   :496-531 (min over pairs of (s/v^2)^2 and
                            s^4/(m_ij^2 r^2), s = r^2 + eps2), evaluated at
                            the predicted state of each step
  time-symmetric corrector  :536-556
  step = dt_param x collision time, dt_param = 0.03   :618-637, :86
with the step rounded down to a power of two dividing the cluster time, and
forces from the oracle's 80-bit pair sum (g6_oracle.c orc_force_ld).
"""
from __future__ import annotations

import math

import numpy as np

from . import force_ld


def cluster_forces(m, x, v, eps2):
    n = m.size
    ids = np.arange(n, dtype=np.int32)
    a, j, p, *_ = force_ld(ids, x, v, eps2, ids, m, x, v)
    return a, j, p


def collision_qmax(m, x, v, eps2):
    """max over pairs of 1/(collision time)^4, rounded to float32 as the
    device keeps it."""
    r = x[None, :, :] - x[:, None, :]
    w = v[None, :, :] - v[:, None, :]
    r2 = (r * r).sum(-1)
    s = r2 + eps2
    v2 = (w * w).sum(-1)
    mij = m[:, None] + m[None, :]
    with np.errstate(divide="ignore", invalid="ignore"):  # self pairs at eps2 = 0
        q1 = (v2 / s) ** 2
        q2 = mij ** 2 * r2 / s ** 4
    q = np.maximum(q1, q2)
    np.fill_diagonal(q, 0.0)
    return float(np.float32(q.max()))


def next_step(t, qmax, dt_param=0.03, dt_limit=0.03125, dt_floor=2.0 ** -40):
    # fourth root as two correctly rounded square roots, as the device does
    cand = dt_param / math.sqrt(math.sqrt(qmax)) if qmax > 0 else dt_limit
    s = dt_limit
    while s > cand and s > dt_floor:
        s *= 0.5
    while math.fmod(t, s) != 0.0 and s > dt_floor:
        s *= 0.5
    return max(s, dt_floor)


def integrate_cluster(m, x, v, eps2, dt_param, t_end):
    x, v = x.copy(), v.copy()
    a, j, _ = cluster_forces(m, x, v, eps2)
    dt = next_step(0.0, collision_qmax(m, x, v, eps2), dt_param)
    t, steps = 0.0, 0
    while True:
        xp = x + v * dt + a * dt * dt / 2 + j * dt * dt * dt / 6
        vp = v + a * dt + j * dt * dt / 2
        a1, j1, _ = cluster_forces(m, xp, vp, eps2)
        q = collision_qmax(m, xp, vp, eps2)
        v1 = v + (a + a1) * dt / 2 + (j - j1) * dt * dt / 12
        x = x + (v + v1) * dt / 2 + (a - a1) * dt * dt / 12
        v = v1
        a, j = a1, j1
        t += dt
        steps += 1
        if t >= t_end:
            break
        dt = next_step(t, q, dt_param)
    return x, v, t, steps


def energy(m, x, v, eps2):
    _, _, p = cluster_forces(m, x, v, eps2)
    return float((m * (0.5 * (v * v).sum(1) + 0.5 * p)).sum())
