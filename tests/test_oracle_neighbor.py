"""The CPU oracle's cell list and half Verlet list, pinned against brute force,
the reference's structural invariants and closed-form grid facts.

Reference: CellList (src/cpu/neighbor/cell_list.cpp:24-895),
tests/cpu/dynamics/test_full_neighbor_list.cpp:51-153.
"""
import numpy as np
import pytest

from conftest import brute_pairs, jittered_lattice

KB = 8.617333262e-5


def c0_geometry(orc):
    a = (4 / 0.8442) ** (1 / 3)
    x, y, z = orc.fcc_lattice(10, 10, 10, a)
    L = 10 * a
    return x, y, z, np.array([L, L, L])


def test_grid_and_stencil_closed_forms(orc):
    # 4000-atom FCC: L = 16.79596, rc+skin = 2.8 -> 11 bins of 1.52691, reach 2; every
    # offset of the 5^3 block has min cell distance^2 = 3*cs^2 = 6.99 < 7.84.
    a = (4 / 0.8442) ** (1 / 3)
    nc, cs, reach, nhalf, nfull = orc.cell_grid(np.full(3, 10 * a), 2.8)
    assert list(nc) == [11, 11, 11]
    assert np.allclose(cs, 10 * a / 11, rtol=0, atol=1e-15)
    assert list(reach) == [2, 2, 2]
    assert nfull == 125 and nhalf == 63
    # 2,048,000-atom FCC: 80 cells -> 95 bins per axis (floor(134.3677/1.4))
    nc, cs, reach, nhalf, nfull = orc.cell_grid(np.full(3, 80 * a), 2.8)
    assert list(nc) == [95, 95, 95] and list(reach) == [2, 2, 2]
    # jittered lattice fixture: L = 7.6, rc+skin = 3.0 -> 5 bins (the half path's minimum)
    nc, cs, reach, nhalf, nfull = orc.cell_grid(np.full(3, 7.6), 3.0)
    assert list(nc) == [5, 5, 5] and list(reach) == [2, 2, 2]
    # pruning: L = 3.6, rc+skin = 3.0 -> 2 bins of 1.8, reach 2; the 8 corner
    # offsets have min distance^2 = 3*1.8^2 = 9.72 > 9 and are dropped
    nc, cs, reach, nhalf, nfull = orc.cell_grid(np.full(3, 3.6), 3.0)
    assert list(nc) == [2, 2, 2] and list(reach) == [2, 2, 2]
    assert nfull == 117 and nhalf == 59


def test_counting_sort_is_stable(orc):
    rng = np.random.default_rng(5)
    L = 9.0
    x, y, z = (rng.uniform(-L, 2 * L, 300) for _ in range(3))  # also exercises the wrap
    atom_cell, count, offset, binned = orc.bin_atoms(x, y, z, np.full(3, L), 3.0)
    assert count.sum() == 300
    assert np.array_equal(offset, np.concatenate([[0], np.cumsum(count)[:-1]]))
    for c in np.nonzero(count)[0]:
        seg = binned[offset[c]:offset[c] + count[c]]
        assert np.all(np.diff(seg) > 0), "ascending i inside a cell"
        assert np.all(atom_cell[seg] == c)
    # 6 bins of 1.5 per axis; cell index = cx*ny*nz + cy*nz + cz of the wrapped coordinate
    nc = 6
    cs = L / nc
    w = lambda v: v - L * np.floor(v / L)
    b = lambda v: np.minimum(np.floor(w(v) / cs), nc - 1)
    cidx = (b(x) * nc * nc + b(y) * nc + b(z)).astype(int)
    assert np.array_equal(cidx, atom_cell)


@pytest.mark.parametrize("case", ["jittered", "fcc_c0", "fcc_c0_perturbed", "small_grid", "random_dense"])
def test_half_list_equals_brute_force(orc, case):
    rcs = 2.8
    if case == "jittered":
        x, y, z, box = jittered_lattice(4, 1.9)
        rcs = 3.0
    elif case.startswith("fcc_c0"):
        x, y, z, box = c0_geometry(orc)
        if case.endswith("perturbed"):
            rng = np.random.default_rng(11)
            x, y, z = (c + rng.normal(0, 0.12, c.size) for c in (x, y, z))
            x, y, z = (c - L * np.floor(c / L) for c, L in zip((x, y, z), box))
    elif case == "small_grid":
        # 4 bins per axis < 2*reach+1: the full-stencil min-image fallback
        rng = np.random.default_rng(3)
        L = 6.0
        x, y, z = (rng.uniform(0, L, 120) for _ in range(3))
        box = np.full(3, L)
    else:
        rng = np.random.default_rng(7)
        L = 12.0
        x, y, z = (rng.uniform(0, L, 1500) for _ in range(3))
        box = np.full(3, L)
    nn, first, nb = orc.build_half(x, y, z, box, rcs)
    got = orc.pair_keys(orc.half_pairs(nn, first, nb), x.size)
    assert np.all(np.diff(got) > 0), "every unordered pair appears once"
    want = brute_pairs(x, y, z, box, rcs)
    assert want.size > 0
    np.testing.assert_array_equal(got, want)


def test_threads_reproduce_serial_csr(orc):
    x, y, z, box = c0_geometry(orc)
    rng = np.random.default_rng(2)
    x = x + rng.normal(0, 0.1, x.size)
    a = orc.build_half(x, y, z, box, 2.8, nthreads=1)
    b = orc.build_half(x, y, z, box, 2.8, nthreads=4)
    for u, v in zip(a, b):
        np.testing.assert_array_equal(u, v)


def test_full_union_complete_and_bounded(orc):
    """test_full_neighbor_list.cpp:51-153 restated on the half list: the
    symmetrised list has no self-legs, holds every pair within the cutoff and
    no pair beyond cutoff + skin."""
    cutoff, skin = 2.6, 0.4
    x, y, z, box = jittered_lattice(4, 1.9)
    nn, first, nb = orc.build_half(x, y, z, box, cutoff + skin)
    pairs = orc.half_pairs(nn, first, nb)
    assert np.all(pairs[:, 0] != pairs[:, 1])
    listed = set(map(tuple, np.sort(pairs, axis=1).tolist()))
    n = x.size
    inside = 0
    for i in range(n):
        for j in range(i + 1, n):
            d = np.array([x[i] - x[j], y[i] - y[j], z[i] - z[j]])
            d -= box * np.rint(d / box)
            r = np.sqrt(d @ d)
            if r <= cutoff:
                assert (i, j) in listed
                inside += 1
            if r > cutoff + skin:
                assert (i, j) not in listed
    assert inside > 0


def test_needs_rebuild_threshold(orc):
    """CellList::needs_rebuild: rebuild iff 4*max|dr|^2 > skin^2, min-imaged."""
    L = 10.0
    box = np.full(3, L)
    xs = np.array([0.05, 5.0])
    ys = np.array([1.0, 1.0])
    zs = np.array([1.0, 1.0])
    skin = 0.3
    # atom 0 crosses the boundary by 0.14 (min-image 0.14 < 0.15): no rebuild
    x = np.array([L - 0.09, 5.0])
    r, m = orc.needs_rebuild(x, ys, zs, xs, ys, zs, box, skin)
    assert not r and abs(m - 0.14 ** 2) < 1e-12
    x = np.array([L - 0.11, 5.0])
    r, m = orc.needs_rebuild(x, ys, zs, xs, ys, zs, box, skin)
    assert r and abs(m - 0.16 ** 2) < 1e-12
