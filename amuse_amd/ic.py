"""Seeded initial conditions in N-body units (G = 1).

Own restatements of the distributions AMUSE's generators draw from; the
random streams are our own (numpy PCG64), so draws match in distribution,
not value by value.

    plummer        src/amuse/ic/plummer.py:49-135 (inverse-CDF radius with a
                   mass cut, von Neumann rejection on q^2 (1-q^2)^3.5 for the
                   speed, isotropic directions)
    scale_to_standard   src/amuse/datamodel/particle_attributes.py:48-110
                   (M = 1, U = -1/2 by a length scaling, K = Q |U|)
    salpeter       src/amuse/ic/salpeter.py:33-36 (inverse CDF of m^alpha)
    kroupa         src/amuse/ic/brokenimf.py:148-177,246-270 (continuous
                   broken power law, segment chosen by its mass fraction)
"""
from __future__ import annotations

import math

import numpy as np


def plummer_raw(n: int, rng: np.random.Generator, rcut: float = 10.0):
    """Plummer sphere in Plummer scale units, equal masses 1/n."""
    mcut = rcut ** 3 / (1.0 + rcut ** 2) ** 1.5
    X = rng.uniform(0.0, mcut, n)
    # X > 0 with probability 1; guard the log-singular draw anyway
    X = np.maximum(X, np.finfo(float).tiny)
    r = 1.0 / np.sqrt(X ** (-2.0 / 3.0) - 1.0)
    pos = r[:, None] * _isotropic(n, rng)
    q = np.empty(0)
    while q.size < n:
        k = n - q.size
        x = rng.uniform(0.0, 1.0, k)
        y = rng.uniform(0.0, 0.1, k)
        q = np.concatenate([q, x[y <= x * x * (1.0 - x * x) ** 3.5]])
    speed = q[:n] * math.sqrt(2.0) * (1.0 + r * r) ** -0.25
    vel = speed[:, None] * _isotropic(n, rng)
    m = np.full(n, 1.0 / n)
    return m, pos, vel


def _isotropic(n: int, rng: np.random.Generator) -> np.ndarray:
    cth = rng.uniform(-1.0, 1.0, n)
    phi = rng.uniform(0.0, 2.0 * math.pi, n)
    sth = np.sqrt(1.0 - cth * cth)
    return np.stack([sth * np.cos(phi), sth * np.sin(phi), cth], axis=1)


def to_nbody_units(pos: np.ndarray, vel: np.ndarray):
    """Plummer scale units -> N-body units (virial radius 1)."""
    s = 3.0 * math.pi / 16.0
    return pos * s, vel / math.sqrt(s)


def center(m, pos, vel):
    M = m.sum()
    pos = pos - (m[:, None] * pos).sum(0) / M
    vel = vel - (m[:, None] * vel).sum(0) / M
    return pos, vel


def potential_energy(m: np.ndarray, pos: np.ndarray, eps2: float = 0.0,
                     block: int = 2048) -> float:
    """Pairwise sum -sum_{i<j} m_i m_j / sqrt(r_ij^2 + eps2), O(N^2) numpy."""
    n = m.size
    U = 0.0
    for a in range(0, n, block):
        xa = pos[a:a + block]
        d = xa[:, None, :] - pos[None, :, :]
        r2 = np.einsum("ijk,ijk->ij", d, d) + eps2
        idx = np.arange(a, min(a + block, n))
        r2[np.arange(idx.size), idx] = np.inf
        U -= 0.5 * float((m[idx][:, None] * m[None, :] / np.sqrt(r2)).sum())
    return U


def kinetic_energy(m: np.ndarray, vel: np.ndarray) -> float:
    return 0.5 * float((m * (vel * vel).sum(1)).sum())


def scale_to_standard(m, pos, vel, U: float | None = None, virial_ratio: float = 0.5):
    """Mass to 1, positions so that U = -1/2, velocities so K = Q/2.

    U may be supplied (e.g. computed on the GPU for large N) for the masses
    after normalisation."""
    m = m / m.sum()
    if U is None:
        U = potential_energy(m, pos)
    pos = pos * (U / -0.5)
    K = kinetic_energy(m, vel)
    vel = vel * math.sqrt(virial_ratio * 0.5 / K)
    return m, pos, vel


def plummer(n: int, seed: int, rcut: float = 10.0, scale: bool = True, U=None):
    rng = np.random.Generator(np.random.PCG64(seed))
    m, pos, vel = plummer_raw(n, rng, rcut)
    pos, vel = to_nbody_units(pos, vel)
    pos, vel = center(m, pos, vel)
    if scale:
        m, pos, vel = scale_to_standard(m, pos, vel, U=U)
        pos, vel = center(m, pos, vel)
    return m, pos, vel


def salpeter(n: int, rng: np.random.Generator, mmin=0.1, mmax=100.0, alpha=-2.35):
    a1 = alpha + 1.0
    u = rng.random(n)
    return mmin * (1.0 + ((mmax / mmin) ** a1 - 1.0) * u) ** (1.0 / a1)


def broken_power_law(n: int, rng: np.random.Generator, bounds, alphas):
    """Continuous broken power law dN/dm ~ m^alpha_k on [b_k, b_{k+1}]."""
    bounds = np.asarray(bounds, float)
    alphas = np.asarray(alphas, float)
    # continuity coefficients c_k with c_k b_{k+1}^a_k = c_{k+1} b_{k+1}^a_{k+1}
    c = np.ones(alphas.size)
    for k in range(1, alphas.size):
        c[k] = c[k - 1] * bounds[k] ** (alphas[k - 1] - alphas[k])
    w = np.empty(alphas.size)
    for k, a in enumerate(alphas):
        lo, hi = bounds[k], bounds[k + 1]
        w[k] = c[k] * (math.log(hi / lo) if a == -1 else (hi ** (a + 1) - lo ** (a + 1)) / (a + 1))
    cum = np.concatenate([[0.0], np.cumsum(w / w.sum())])
    u = rng.random(n)
    k = np.clip(np.searchsorted(cum, u, side="right") - 1, 0, alphas.size - 1)
    s = (u - cum[k]) / (cum[k + 1] - cum[k])
    lo, hi, a = bounds[k], bounds[k + 1], alphas[k]
    out = np.empty(n)
    one = a == -1
    out[one] = lo[one] * (hi[one] / lo[one]) ** s[one]
    a1 = a[~one] + 1
    out[~one] = lo[~one] * (1.0 + ((hi[~one] / lo[~one]) ** a1 - 1.0) * s[~one]) ** (1.0 / a1)
    return out


def kroupa(n: int, rng: np.random.Generator, mmin=0.08, mmax=20.0):
    """Kroupa (2001) segments 0.08-0.5 (-1.3), 0.5-mmax (-2.3)."""
    return broken_power_law(n, rng, [mmin, 0.5, mmax], [-1.3, -2.3])


def plummer_with_masses(n: int, seed: int, masses: np.ndarray, eps2_U: float = 0.0,
                        U=None):
    """Plummer positions/velocities with the given mass spectrum, scaled to
    standard units."""
    rng = np.random.Generator(np.random.PCG64(seed))
    _, pos, vel = plummer_raw(n, rng)
    pos, vel = to_nbody_units(pos, vel)
    m = masses / masses.sum()
    pos, vel = center(m, pos, vel)
    m, pos, vel = scale_to_standard(m, pos, vel, U=U)
    pos, vel = center(m, pos, vel)
    return m, pos, vel
