"""The oracle's geom_axis / ml_geom_axis splitters (oracle/bvh_oracle.c
orc_split_geom_axis / orc_split_ml_geom_axis) against the reference's own
known answers (tests/SplitTest.cpp) and against the properties the
algorithm fixes: a permutation; for geom_axis, one mean partition of the
identity followed by midpoint splits; for ml_geom_axis, at every node of the
recursion the range's own longest axis and mean separate the two sides, and
the two-pointer (libstdc++ std::partition) order is reproduced by a direct
simulation written from the algorithm's description."""
import numpy as np
import pytest

import oracle as O


def spheres(centers, r=1.0):
    c = np.asarray(centers, np.float64)
    return np.stack([O.kdop_from_sphere(ci, r) for ci in c]), c


def line8():
    c = np.zeros((8, 3))
    c[:, 0] = np.arange(8)
    return spheres(c)


def grid4x3():
    return spheres([[i, j, 0.0] for j in range(3) for i in range(4)])


# tests/SplitTest.cpp:168-239 ("recursive mean splitting 1D elements")
@pytest.mark.parametrize("depth,want", [(0, []), (1, [4]), (2, [2, 4, 6]), (3, [1, 2, 3, 4, 5, 6, 7])])
def test_geom_axis_1d_known_answers(depth, want):
    kd, c = line8()
    idx, sp = O.split_permutations(kd, c, depth)
    assert sp.tolist() == want
    assert idx.tolist() == list(range(8))


def test_geom_axis_1d_depth4():
    kd, c = line8()
    _, sp = O.split_permutations(kd, c, 4)
    assert len(sp) == 15
    assert all(0 <= int(sp[i + 1]) - int(sp[i]) <= 1 for i in range(14))


# tests/SplitTest.cpp:327-438 ("new recursive mean splitting 2D elements", 26-DOP snapshots)
def test_ml_geom_axis_2d_known_answers():
    kd, c = grid4x3()
    n = len(c)
    assert len(O.split_permutations(kd, c, 0, ml=True)[1]) == 0
    assert O.split_permutations(kd, c, 1, ml=True)[1].tolist() == [n // 2]
    sp = O.split_permutations(kd, c, 2, ml=True)[1].astype(np.int64)
    assert len(sp) == 3 and sp[1] == 6 and n - sp[1] == 6
    assert sp[0] >= 2 and sp[1] - sp[0] >= 2 and sp[2] - sp[1] >= 2 and n - sp[2] >= 2
    sp = O.split_permutations(kd, c, 3, ml=True)[1].astype(np.int64)
    assert len(sp) == 7 and np.all(np.diff(sp) >= 0) and np.all(np.diff(sp) <= 4)
    sp = O.split_permutations(kd, c, 4, ml=True)[1].astype(np.int64)
    assert len(sp) == 15 and np.all(np.diff(sp) >= 0) and np.all(np.diff(sp) <= 2)


def _hoare(idx, b, e, pred):
    """Two-pointer partition: scan from the left for an element that fails,
    from the right for one that holds, swap them, repeat until the scans
    meet (the libstdc++ std::partition loop for bidirectional iterators)."""
    first, last = b, e
    while True:
        while first != last and pred(idx[first]):
            first += 1
        if first == last:
            return first
        last -= 1
        while first != last and not pred(idx[last]):
            last -= 1
        if first == last:
            return first
        idx[first], idx[last] = idx[last], idx[first]
        first += 1


def _union_axis(kd, ids):
    lo = kd[ids, 0::2].min(axis=0)
    hi = kd[ids, 1::2].max(axis=0)
    return int(np.argmax(hi - lo))  # first maximum, as std::max_element


def _simulate(kd, c, depth, ml):
    n = len(c)
    idx = list(range(n))
    splits = []
    proj = lambda i, a: O.project(c[i], a)
    if not ml:
        axis = _union_axis(kd, np.arange(n))
        s = 0.0
        for i in range(n):
            s += proj(i, axis)
        m = s / n

    def rec(d, b, e):
        if ml:
            a = _union_axis(kd, np.array(idx[b:e])) if e > b else 0
            if e - b < 2:
                sp = b
            else:
                s = 0.0
                for i in idx[b:e]:
                    s += proj(i, a)
                mm = s / (e - b)
                sp = _hoare(idx, b, e, lambda i: proj(i, a) < mm)
                if sp in (b, e):
                    sp = b + (e - b) // 2
        else:
            sp = _hoare(idx, b, e, lambda i: proj(i, axis) < m)
            if sp in (b, e):
                sp = b + (e - b) // 2
        if d > 0:
            rec(d - 1, b, sp)
        splits.append(sp)
        if d > 0:
            rec(d - 1, sp, e)

    if depth > 0:
        rec(depth - 1, 0, n)
    return np.array(idx, np.uint32), np.array(splits, np.uint64)


@pytest.mark.parametrize("ml", [False, True])
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_splitters_match_simulation(ml, seed):
    rng = np.random.default_rng(seed)
    n = 300
    c = rng.normal(size=(n, 3)) * np.array([3.0, 1.0, 0.5])
    kd, c = spheres(c, 0.1)
    for depth in range(0, 6):
        idx, sp = O.split_permutations(kd, c, depth, ml=ml)
        widx, wsp = _simulate(kd, c, depth, ml)
        assert np.array_equal(idx, widx) and np.array_equal(sp, wsp), depth
        assert np.array_equal(np.sort(idx), np.arange(n, dtype=np.uint32))


def test_geom_axis_properties():
    """One mean partition, then midpoints: the first split separates
    projections below the mean from the rest, deeper splits halve ranges."""
    rng = np.random.default_rng(7)
    kd, c = spheres(rng.uniform(size=(257, 3)) * np.array([5.0, 1.0, 1.0]), 0.05)
    idx, sp = O.split_permutations(kd, c, 3)
    axis = O.longest_axis(np.concatenate([kd[:, 0::2].min(0)[:, None], kd[:, 1::2].max(0)[:, None]], 1).ravel())
    p = np.array([O.project(ci, axis) for ci in c])
    mean = np.sum(p) / len(p)
    s0 = int(sp[3])  # the root split is the middle entry of the in-order list
    assert np.all(p[idx[:s0]] < mean) and np.all(p[idx[s0:]] >= mean)
    assert int(sp[1]) == s0 // 2 and int(sp[5]) == s0 + (len(p) - s0) // 2


def test_object_patches_from_splitters():
    """orc_object_build with a splitter: patches are the split ranges of the
    permutation, od a power of two."""
    rng = np.random.default_rng(11)
    kd, c = spheres(rng.uniform(size=(500, 3)), 0.02)
    for split, ml in ((O.SPLIT_GEOM_AXIS, False), (O.SPLIT_ML_GEOM_AXIS, True)):
        o = O.Object(kd, c, od=8, split=split)
        idx, sp = O.split_permutations(kd, c, 3, ml=ml)
        assert np.array_equal(o.perm, idx) and np.array_equal(o.splits, sp)
        b = np.concatenate([[0], sp]).astype(np.int64)
        e = np.concatenate([sp, [len(c)]]).astype(np.int64)
        assert np.array_equal(o.pbegin.astype(np.int64), b) and np.array_equal(o.pend.astype(np.int64), e)
