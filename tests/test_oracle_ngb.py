"""The oracle's neighbour query, pinned before it is trusted.

The reference criterion (lib/sapporo_light/dev_evaluate_gravity.cu:60-67):
j is a neighbour of i when the unsoftened r^2 <= h2[i] and the ids differ;
j slots with id -1 are empty and skipped (lib/g6lib/g6lib.c:135-140). The
lists are reported in ascending slot order with the full count, at most
`cap` ids kept (lib/sapporo_light/sapporo.cpp:234-272).

What pins it, chosen so that a plausible slip fails one check:
* an integer lattice, where every r^2 is exact: an interior point has 6
  neighbours at h2 = 1 (the r^2 == h2 case: a strict < gives 0), 18 at
  h2 = 2, 26 at h2 = 3 and 32 at h2 = 4, and none just below 1; a corner
  point has 3 / 6 / 7; the face neighbours are the six slots at +-1 along
  each axis;
* self (r^2 = 0) is never listed, nor is any other slot carrying i's id or
  an id of -1; a probe with a fresh id at a lattice point lists the point;
* lists are in slot order (ids assigned in reverse so slot and id order
  differ), counts stay exact past the cap and the kept ids are the first
  `cap` in slot order;
* a numpy brute force on random sets with duplicate and -1 ids.
"""
import numpy as np

import oracle

L = 7          # lattice points per axis
S = 0.5        # spacing (a power of two: every coordinate and r^2 is exact)


def lattice():
    g = np.arange(L) * S
    x = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    return np.ascontiguousarray(x)


def slot(a, b, c):
    return (a * L + b) * L + c


def query(ids_j, xj, id_i, xi, h2, cap=4096):
    xi = np.atleast_2d(np.asarray(xi, dtype=np.float64))
    id_i = np.atleast_1d(np.asarray(id_i, dtype=np.int32))
    h2 = np.broadcast_to(np.asarray(h2, dtype=np.float64), (id_i.size,)).copy()
    return oracle.ngb(id_i, xi, h2, ids_j, xj, cap=cap)


def test_lattice_shell_counts():
    x = lattice()
    ids = np.arange(1, x.shape[0] + 1, dtype=np.int32)
    c = slot(3, 3, 3)
    for k, want in ((0.999, 0), (1, 6), (2, 18), (3, 26), (4, 32)):
        counts, _ = query(ids, x, ids[c], x[c], k * S * S)
        assert counts[0] == want, (k, counts[0])
    corner = slot(0, 0, 0)
    for k, want in ((1, 3), (2, 6), (3, 7)):
        counts, _ = query(ids, x, ids[corner], x[corner], k * S * S)
        assert counts[0] == want, (k, counts[0])


def test_lattice_face_neighbours_exact():
    x = lattice()
    ids = np.arange(1, x.shape[0] + 1, dtype=np.int32)
    a, b, c = 2, 4, 3
    counts, lists = query(ids, x, ids[slot(a, b, c)], x[slot(a, b, c)], S * S)
    faces = sorted(slot(a + d[0], b + d[1], c + d[2]) for d in
                   ((-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)))
    assert counts[0] == 6
    assert list(lists[0, :6]) == [ids[s] for s in faces]


def test_self_duplicate_and_empty_ids_excluded():
    x = lattice()
    ids = np.arange(1, x.shape[0] + 1, dtype=np.int32)
    c = slot(3, 3, 3)
    up, down, east = slot(4, 3, 3), slot(2, 3, 3), slot(3, 4, 3)
    ids2 = ids.copy()
    ids2[up] = -1            # empty slot at a neighbour position
    ids2[down] = ids[c]      # another slot carrying i's id
    counts, lists = query(ids2, x, ids[c], x[c], S * S)
    assert counts[0] == 4
    kept = set(lists[0, :4].tolist())
    assert ids[east] in kept and -1 not in kept and ids[c] not in kept
    # a probe with a fresh id at the lattice point lists the point itself
    counts, lists = query(ids, x, 10 ** 6, x[c], S * S)
    assert counts[0] == 7 and ids[c] in lists[0, :7]
    # a probe with id -1 (phiGRAPE's field-at-point calls) skips only -1 slots
    counts, _ = query(ids2, x, -1, x[c], S * S)
    assert counts[0] == 6


def test_slot_order_and_cap():
    x = lattice()
    n = x.shape[0]
    ids = np.arange(n, 0, -1, dtype=np.int32)        # ids descend with the slot
    c = slot(3, 3, 3)
    counts, lists = query(ids, x, ids[c], x[c], 4 * S * S, cap=n)
    assert counts[0] == 32
    full = lists[0, :32]
    assert np.all(np.diff(full) < 0)                  # slot order, not id order
    counts10, lists10 = query(ids, x, ids[c], x[c], 4 * S * S, cap=10)
    assert counts10[0] == 32
    assert np.array_equal(lists10[0], full[:10])


def brute(id_i, xi, h2, id_j, xj):
    d = xj[None, :, :] - xi[:, None, :]
    r2 = (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]
    ok = (r2 <= h2[:, None]) & (id_j[None, :] != id_i[:, None]) & (id_j[None, :] != -1)
    return ok


def test_random_sets_match_brute_force():
    rng = np.random.Generator(np.random.PCG64(77))
    for trial in range(4):
        nj, ni = 500, 120
        xj = rng.normal(0, 1, (nj, 3))
        id_j = rng.integers(1, 400, nj).astype(np.int32)     # duplicates
        id_j[rng.choice(nj, 25, replace=False)] = -1
        pick = rng.choice(nj, ni, replace=False)
        xi = xj[pick] + rng.normal(0, 1e-3, (ni, 3)) * (trial % 2)
        id_i = id_j[pick].copy()
        id_i[:10] = 10 ** 6
        h2 = rng.uniform(0, 1.0, ni)
        counts, lists = oracle.ngb(id_i, xi, h2, id_j, xj, cap=nj)
        ok = brute(id_i, xi, h2, id_j, xj)
        assert np.array_equal(counts, ok.sum(1))
        for i in range(ni):
            assert np.array_equal(lists[i, :counts[i]], id_j[np.nonzero(ok[i])[0]])
