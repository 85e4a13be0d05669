"""Writes the golden fixtures of tests/golden/ from the reference CPU GRAPE-6
emulator lib/g6lib/g6lib.c, compiled from its source by oracle/Makefile into
oracle/_ref/libg6ref.so (this container only).

    python tests/golden/make_golden.py

Each fixture holds inputs and the reference's outputs, obtained through the
reference's own ABI (g6_open_, g6_set_j_particle_, g6_set_ti_,
g6calc_firsthalf_, g6calc_lasthalf2_, g6lib.c:190-441).  Note that the
reference predicts velocities with its own polynomial (g6lib.c:85-87), so
cases with t_i != t_j pin the library's predictor-compat mode.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from amuse_amd import ic  # noqa: E402


def run_ref(ref, case):
    ref.open()
    for j in range(case["nj"]):
        if case["id_j"][j] == -2:   # slot left untouched (stays id -1 from reset)
            continue
        ref.set_j_particle(j, case["id_j"][j], case["tj"][j], 0.0, case["m"][j], case["k18"][j],
                           case["j6"][j], case["a2"][j], case["v"][j], case["x"][j])
    ref.set_ti(case["ti"])
    acc, jerk, pot, nn, idx = ref.calc(case["nj"], case["id_i"], case["xi"], case["vi"],
                                      case["eps2"], nn=True)
    ref.close()
    out = dict(case)
    out.update(acc=acc, jerk=jerk, pot=pot, nn=nn, index_out=idx)
    return out


def base_case(n, seed, dt_nonzero, k18_scale=0.0):
    rng = np.random.Generator(np.random.PCG64(seed))
    m, x, v = ic.plummer(n, seed)
    a2 = rng.normal(0, 0.5, (n, 3))
    j6 = rng.normal(0, 0.5, (n, 3))
    k18 = rng.normal(0, 1.0, (n, 3)) * k18_scale
    ti = 0.625
    tj = ti - (rng.uniform(0, 1, n) * 2.0 ** -rng.integers(3, 11, n) if dt_nonzero else 0.0)
    tj = np.broadcast_to(tj, (n,)).astype(float).copy()
    ids = np.arange(1, n + 1, dtype=np.int32)
    # i-set: the first half of the j-set predicted by the caller's own
    # polynomial (exact Taylor), plus its ids -> self-interaction is skipped
    ni = n // 2
    dt = ti - tj[:ni]
    xi = x[:ni] + v[:ni] * dt[:, None] + a2[:ni] * dt[:, None] ** 2 + j6[:ni] * dt[:, None] ** 3
    vi = v[:ni] + 2 * a2[:ni] * dt[:, None] + 3 * j6[:ni] * dt[:, None] ** 2
    return dict(nj=n, id_j=ids, tj=tj, m=m, x=x, v=v, a2=a2, j6=j6, k18=k18, ti=ti,
                id_i=ids[:ni].copy(), xi=xi, vi=vi, eps2=1e-4)


def lattice_case():
    # 6x6x6 unit lattice; i-points at lattice nodes and at cell centres, so
    # many j are at exactly equal distance (ties resolved by lowest slot)
    g = np.arange(6, dtype=float)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    x = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    n = x.shape[0]
    rng = np.random.Generator(np.random.PCG64(7))
    perm = rng.permutation(n)      # slots not in geometric order
    x = x[perm]
    ids = np.arange(10, 10 + n, dtype=np.int32)
    centres = np.array([[0.5, 0.5, 0.5], [2.5, 1.5, 3.5], [4.5, 4.5, 0.5], [1.0, 2.0, 2.5]])
    xi = np.concatenate([x[:40], centres])
    id_i = np.concatenate([ids[:40], np.full(len(centres), -1, np.int32)])
    z = np.zeros((n, 3))
    return dict(nj=n, id_j=ids, tj=np.zeros(n), m=np.full(n, 1.0 / n), x=x, v=z.copy(),
                a2=z.copy(), j6=z.copy(), k18=z.copy(), ti=0.0, id_i=id_i, xi=xi,
                vi=np.zeros((len(xi), 3)), eps2=0.01)


def special_case():
    # empty slots inside [0, nj), duplicate ids, an i whose id is not in the
    # j-set, field probes with id -1, eps2 = 0
    c = base_case(96, 11, dt_nonzero=True)
    idj = c["id_j"].copy()
    idj[[3, 17, 50]] = -1           # explicitly emptied slots
    idj[[60, 61]] = 5               # duplicates of particle 5
    idj[[70]] = -2                  # never written (reset state)
    c["id_j"] = idj
    c["id_i"] = c["id_i"].copy()
    c["id_i"][[0, 1]] = [5, 9999]
    c["id_i"][[2, 3]] = -1
    c["eps2"] = 0.0
    c["nj"] = 80                    # slots >= 80 ignored by the calc
    return c


def probe_case():
    # src/amuse_ph4/tests/test_ph4.py:786-823: unit masses at x=0 and x=2,
    # eps2 = 1e-5, potentials at x = 0.25, 0.5, 0.75 (and mirrors)
    x = np.array([[0.0, 0, 0], [2.0, 0, 0]])
    z = np.zeros((2, 3))
    xs = np.array([0.25, 0.5, 0.75, 1.75, 1.5, 1.25])
    xi = np.stack([xs, np.zeros(6), np.zeros(6)], 1)
    return dict(nj=2, id_j=np.array([1, 2], np.int32), tj=np.zeros(2), m=np.ones(2), x=x,
                v=z.copy(), a2=z.copy(), j6=z.copy(), k18=z.copy(), ti=0.0,
                id_i=np.full(6, -1, np.int32), xi=xi, vi=np.zeros((6, 3)), eps2=1e-5)


def main():
    ref = oracle.RefG6()
    cases = {
        "g6ref_dt0": base_case(512, 1, dt_nonzero=False),
        "g6ref_dt": base_case(512, 2, dt_nonzero=True, k18_scale=0.3),
        "g6ref_lattice": lattice_case(),
        "g6ref_special": special_case(),
        "g6ref_probe": probe_case(),
    }
    for name, case in cases.items():
        out = run_ref(ref, case)
        arrays = {k: np.asarray(v) for k, v in out.items()}
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **arrays)
        print("wrote", name, {k: v.shape for k, v in arrays.items() if v.ndim})


if __name__ == "__main__":
    main()
