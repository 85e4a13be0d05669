"""BHtree's use of the GRAPE-6 ABI (src/amuse_bhtree/src/BHtree.C:764-885):
an interaction list (particles and tree-node pseudo-particles) loaded as
j-particles with index = list position, zero velocity and derivatives and
t_j = 0; t_i = 0; the leaf particles, which sit in the list at first_leaf..,
as the i-set in chunks of g6_npipes() with zero velocities, through
firsthalf/lasthalf (acc and pot used); and a close/reopen of the unit after
a call budget.  Acc and pot must match the 80-bit oracle on the same list to
the plain path's bar, self-interaction excluded by the shared index.
"""
import numpy as np
import pytest

import oracle
from conftest import print_case
from test_gpu_parity import COND

pytestmark = pytest.mark.gpu


def interaction_list(seed, n_leaf=300, n_part=1500, n_nodes=400):
    rng = np.random.Generator(np.random.PCG64(seed))
    # ordinary particles, then node pseudo-particles (heavier, farther out)
    x = np.concatenate([rng.normal(0, 0.5, (n_part, 3)), rng.normal(0, 3.0, (n_nodes, 3))])
    m = np.concatenate([rng.uniform(0.5, 1.5, n_part) / n_part,
                        rng.uniform(5, 50, n_nodes) / n_part])
    first_leaf = 200
    return x, m, first_leaf, n_leaf


def bhtree_call(g6, x, m, first_leaf, ni, eps2, max_ni):
    n = m.size
    z = np.zeros(3)
    g6.set_ti(0.0)
    for i in range(n):
        g6.set_j_particle(i, i, 0.0, 0.0, float(m[i]), z, z, z, z, x[i])
    index = np.arange(first_leaf, first_leaf + ni, dtype=np.int32)
    pos = x[first_leaf:first_leaf + ni]
    acc = np.zeros((ni, 3))
    pot = np.zeros(ni)
    for p0 in range(0, ni, max_ni):
        k = min(max_ni, ni - p0)
        g6.calc_firsthalf(n, index[p0:p0 + k], pos[p0:p0 + k], np.zeros((k, 3)), eps2)
        a, _, ph, idx = g6.calc_lasthalf(n, k, eps2)
        assert np.array_equal(idx, index[p0:p0 + k])  # g6lib writes the ids back
        acc[p0:p0 + k] = a
        pot[p0:p0 + k] = ph
    return acc, pot


@pytest.mark.parametrize("npipes", [64, 16384])
def test_bhtree_interaction_list(g6, npipes):
    x, m, first_leaf, ni = interaction_list(31)
    eps2 = 1e-4
    print_case(f"bhtree list n={m.size} ni={ni} npipes={npipes}")
    g6.set_npipes(npipes)
    try:
        acc, pot = bhtree_call(g6, x, m, first_leaf, ni, eps2, g6.npipes())
    finally:
        g6.set_npipes(16384)
    ids = np.arange(m.size, dtype=np.int32)
    zero = np.zeros_like(x)
    ea, _, ep, sa, _, sp = oracle.force_ld(ids[first_leaf:first_leaf + ni],
                                           x[first_leaf:first_leaf + ni], zero[:ni], eps2, ids,
                                           m, x, zero)
    da = np.abs(acc - ea).max(1)
    assert (da <= COND["acc"] * sa).all(), (da / sa).max()
    assert (np.abs(pot - ep) <= COND["pot"] * sp).all()


def test_bhtree_reopen_cycle(g6):
    """The budget-triggered g6_close / g6_open of BHtree.C:879-884 between
    two calls leaves no state behind: the second list gives the same result
    as on a fresh unit."""
    x, m, first_leaf, ni = interaction_list(32, n_part=800, n_nodes=100)
    a1, p1 = bhtree_call(g6, x, m, first_leaf, ni, 1e-4, g6.npipes())
    g6.close()
    g6.open()
    x2, m2, fl2, ni2 = interaction_list(33, n_part=600, n_nodes=50)
    a2, p2 = bhtree_call(g6, x2, m2, fl2, ni2, 1e-4, g6.npipes())
    g6.close()
    g6.open()
    g6.reset()
    a3, p3 = bhtree_call(g6, x2, m2, fl2, ni2, 1e-4, g6.npipes())
    assert np.array_equal(a2, a3) and np.array_equal(p2, p3)
