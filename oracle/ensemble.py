"""CPU oracle of the batched-ensemble integrator — TEST INFRASTRUCTURE ONLY.

Numpy restatement, cluster by cluster, of the scheme in
amuse_amd/csrc/g6_clusters.hip:
  predictor   src/amuse_hermite/interface.cc:266-281
  corrector + Aarseth step   src/amuse_ph4/src/idata.cc:431-462
  first step  src/amuse_ph4/src/jdata.cc:514-544 (minimised over the cluster)
  step update (halve once, or double when commensurate)  idata.cc:464-510
with forces from the oracle's 80-bit pair sum (g6_oracle.c orc_force_ld).
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


def pow2_floor(x):
    mant, e = math.frexp(x)
    return x / (2 * mant)


def integrate_cluster(m, x, v, eps2, eta, t_end, dt_fac=0.0625, dt_limit=0.03125):
    x, v = x.copy(), v.copy()
    a, j, _ = cluster_forces(m, x, v, eps2)
    a2 = (a * a).sum(1)
    j2 = (j * j).sum(1)
    first = np.where((a2 == 0) | (j2 == 0), dt_fac * eta, dt_fac * eta * np.sqrt(a2 / np.where(j2 == 0, 1, j2)))
    dt = pow2_floor(float(first.min()))
    while dt > dt_limit:
        dt /= 2
    t, steps = 0.0, 0
    while t < t_end:
        xp = x + v * dt + a * dt * dt / 2 + j * dt * dt * dt / 6
        vp = v + a * dt + j * dt * dt / 2
        a1, j1, _ = cluster_forces(m, xp, vp, eps2)
        alpha = -3 * (a - a1) - dt * (2 * j + j1)
        beta = 2 * (a - a1) + dt * (j + j1)
        x = xp + (alpha / 12 + beta / 20) * dt * dt
        v = vp + (alpha / 3 + beta / 4) * dt
        A2 = (a1 * a1).sum(1)
        J2 = ((dt * j1) ** 2).sum(1)
        K2 = ((2 * alpha) ** 2).sum(1)
        L2 = ((6 * beta) ** 2).sum(1)
        cand = float((eta * dt * np.sqrt((np.sqrt(A2 * K2) + J2) / (np.sqrt(J2 * L2) + K2))).min())
        a, j = a1, j1
        t += dt
        steps += 1
        old2 = dt
        while math.fmod(t, old2) != 0:
            old2 /= 2
        if cand < old2:
            dt = old2 / 2
        else:
            t2 = 2 * old2
            dt = t2 if (cand >= t2 and math.fmod(t, t2) == 0 and t2 <= dt_limit) else old2
    return x, v, t, steps


def energy(m, x, v, eps2):
    _, _, p = cluster_forces(m, x, v, eps2)
    return float((m * (0.5 * (v * v).sum(1) + 0.5 * p)).sum())
