"""g6x_calc_self_noalloc, the allocation-free form of g6x_calc_self: its
results (written to caller-owned device arrays) equal g6x_calc_self +
g6x_fetch on the same j-store bit for bit, in both precisions, with and
without nearest neighbours, on two calls in a row (the second smaller, over
another slot range, after set_j_particle updates that the call applies from
the staging buffers); sizes beyond its static scratch are refused."""
import numpy as np
import pytest

from conftest import print_case

# torch's HIP runtime is loaded before the library's (at collection), as in
# the other modules that use torch
torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

G6X_ERR_ARG = -3
MAX_NI = 1 << 19


def plummer_store(n, seed):
    from amuse_amd import ic
    rng = np.random.Generator(np.random.PCG64(seed))
    m, x, v = ic.plummer(n, seed, scale=False)
    a2 = rng.normal(0, 0.3, (n, 3))
    j6 = rng.normal(0, 0.3, (n, 3))
    tj = 1.0 - rng.uniform(0, 1, n) * 2.0 ** -rng.integers(3, 11, n)
    ids = np.arange(1, n + 1, dtype=np.int32)
    return m, x, v, a2, j6, tj, ids


def device_out(ni):
    dev = torch.device("cuda")
    return (torch.full((ni, 3), np.nan, dtype=torch.float64, device=dev),
            torch.full((ni, 3), np.nan, dtype=torch.float64, device=dev),
            torch.full((ni,), np.nan, dtype=torch.float64, device=dev),
            torch.full((ni,), -7, dtype=torch.int32, device=dev))


def noalloc(g6, nj, i0, ni, eps2, nn=True):
    acc, jerk, pot, nnb = device_out(ni)
    torch.cuda.synchronize()   # the fills are done before the library's stream writes
    g6.calc_self_noalloc(nj, i0, ni, eps2, acc.data_ptr(), jerk.data_ptr(), pot.data_ptr(),
                         nnb.data_ptr() if nn else None)
    g6.sync()
    return (acc.cpu().numpy(), jerk.cpu().numpy(), pot.cpu().numpy(),
            nnb.cpu().numpy() if nn else None)


def assert_same(a, b):
    for p, q in zip(a, b):
        if p is None:
            assert q is None
            continue
        np.testing.assert_array_equal(p, q)


@pytest.mark.parametrize("prec", [0, 1])
def test_noalloc_equals_calc_self(g6, prec):
    n, eps2 = 65536, 1e-4
    m, x, v, a2, j6, tj, ids = plummer_store(n, 5150)
    z = np.zeros((n, 3))
    g6.set_precision(prec)
    g6.set_j_particles(np.arange(n), ids, tj, 0.0, m, z, j6, a2, v, x)
    # set-up may allocate: the staging buffers grow to the first upload here
    assert g6.lib.g6_flush_jp_buffer(0) == 0
    g6.set_ti(1.0)
    # first call: every slot, with a chunking both paths take (the automatic
    # one of a full pass would give the allocation-free call fewer chunks)
    g6.set_tiles(1, 2, 4)
    print_case(f"noalloc prec={prec} full pass n={n}")
    got = noalloc(g6, n, 0, n, eps2)
    g6.calc_self(n, 0, n, eps2, nn=True)
    assert_same(got, g6.fetch(n))
    assert ((got[3] >= 1) & (got[3] <= n)).all()
    g6.set_tiles(0, 0)
    # updates staged by set_j_particle, applied by the next call
    rng = np.random.Generator(np.random.PCG64(3))
    upd = rng.choice(n, 500, replace=False)
    x2 = x.copy()
    x2[upd] += rng.normal(0, 1e-3, (500, 3))
    g6.set_j_particles(upd, ids[upd], tj[upd], 0.0, m[upd], z[upd], j6[upd], a2[upd], v[upd],
                       x2[upd])
    # second call: smaller, another slot range, automatic chunking
    i0, ni = 1234, 3000
    print_case(f"noalloc prec={prec} second call i0={i0} ni={ni}")
    got2 = noalloc(g6, n, i0, ni, eps2)
    g6.calc_self(n, i0, ni, eps2, nn=True)
    assert_same(got2, g6.fetch(ni))
    # without nearest neighbours
    got3 = noalloc(g6, n, i0, ni, eps2, nn=False)
    assert_same(got3[:3], got2[:3])
    g6.set_precision(0)


def test_noalloc_limits(g6):
    n = 4096
    m, x, v, a2, j6, tj, ids = plummer_store(n, 7)
    z = np.zeros((n, 3))
    g6.set_j_particles(np.arange(n), ids, tj, 0.0, m, z, j6, a2, v, x)
    g6.set_ti(1.0)
    lib = g6.lib
    assert lib.g6_flush_jp_buffer(0) == 0
    acc, jerk, pot, nnb = device_out(8)
    args = (acc.data_ptr(), jerk.data_ptr(), pot.data_ptr(), nnb.data_ptr())
    # more i than the static scratch holds
    assert lib.g6x_calc_self_noalloc(n, 0, MAX_NI + 1, 1e-4, *args) == G6X_ERR_ARG
    assert b"G6X_NOALLOC_MAX_NI" in lib.g6x_last_error()
    # slots beyond the j-store's capacity (the call cannot grow it)
    assert lib.g6x_calc_self_noalloc(1 << 27, 0, 8, 1e-4, *args) == G6X_ERR_ARG
    assert lib.g6x_calc_self_noalloc(n, -1, 8, 1e-4, *args) == G6X_ERR_ARG
    # a valid call still works after the refusals
    got = noalloc(g6, n, 0, 8, 1e-4)
    g6.calc_self(n, 0, 8, 1e-4, nn=True)
    assert_same(got, g6.fetch(8))
