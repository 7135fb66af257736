"""GPU parity: every stage of the HIP pipeline against the CPU oracle on the
same seeded inputs.  Integer / index outputs and fp64 slabs are compared bit
for bit (slabs by value, so +0 == -0); pair sets exactly."""
import numpy as np
import pytest

import oracle as O
from distbvh_amd import CollisionWorld
from distbvh_amd import workloads as W
from helpers import FORMS, from_cuda, gpu_object, oracle_elements, sort_pairs, sort_records, to_cuda

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU test needs a GPU"
    return torch


# ---- primitives -------------------------------------------------------------------
@pytest.mark.parametrize("n", [1, 2, 5, 4095, 4096, 4097, 100003, 1 << 20])
@pytest.mark.parametrize("bits", [64, 32])
def test_sort_matches_stable_argsort(torch_cuda, n, bits):
    rng = np.random.default_rng(n * 7 + bits)
    if bits == 64:
        keys = rng.integers(0, 1 << 63, n, dtype=np.uint64)
        keys[::3] = keys[0]  # duplicates exercise stability
        keys = keys & np.uint64(0x7fff00ff00ff00ff)
        dt = np.uint64
    else:
        keys = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
        keys[1::4] = 7
        dt = np.uint32
    vals = np.arange(n, dtype=np.uint32)
    with CollisionWorld(1) as w:
        tk = to_cuda(keys)
        tv = to_cuda(vals)
        w.sort(tk, tv)
        torch_cuda.cuda.synchronize()
        gk = from_cuda(tk, dt)
        gv = from_cuda(tv, np.uint32)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(gk, keys[order])
    assert np.array_equal(gv, order.astype(np.uint32))


def test_sort_known_answer(torch_cuda):
    # tests/sort_test.cpp:9-30
    keys = np.array([3, 7, 2, 1, 4, 9, 1, 3], np.uint32)
    vals = np.arange(8, dtype=np.uint32)
    with CollisionWorld(1) as w:
        tk, tv = to_cuda(keys), to_cuda(vals)
        w.sort(tk, tv)
        assert list(from_cuda(tk, np.uint32)) == [1, 1, 2, 3, 3, 4, 7, 9]
        assert list(from_cuda(tv, np.uint32)) == [3, 6, 2, 0, 7, 4, 1, 5]


@pytest.mark.parametrize("n", [1, 8, 4097, 1000001])
def test_exclusive_scan(torch_cuda, n):
    rng = np.random.default_rng(n)
    v = rng.integers(0, 1000, n, dtype=np.uint32)
    if n == 8:
        v = np.array([3, 7, 2, 1, 4, 9, 1, 3], np.uint32)  # tests/UtilTest.cpp:42-55
    with CollisionWorld(1) as w:
        t = to_cuda(v)
        w.exclusive_scan(t)
        got = from_cuda(t, np.uint32)
    exp = (np.concatenate([np.zeros(1, np.uint64), np.cumsum(v.astype(np.uint64))[:-1]])
           & np.uint64(0xffffffff)).astype(np.uint32)
    assert np.array_equal(got, exp)


# ---- object build -----------------------------------------------------------------
def _oracle_object(kind, data, eps, od, bits, leaf=1):
    kd, ce = oracle_elements(kind, data, eps)
    return O.Object(kd, ce, od=od, morton_bits=bits, leaf_size=leaf), kd, ce


@pytest.mark.parametrize("bits", [63, 30])
@pytest.mark.parametrize("od", [1, 2, 4, 32, 64])
@pytest.mark.parametrize("device", [False, True])
def test_build_stages_flat_sheets(torch_cuda, bits, od, device):
    a, b = W.flat_sheets(64)
    for data in (a, b):
        ref, kd, ce = _oracle_object("poly", data, 0.0, od, bits)
        with CollisionWorld(od, morton_bits=bits) as w:
            obj = gpu_object(w, "poly", data, 0.0, device=device)
            mn, inv = obj.bounds()
            rmn, rinv = O.compute_bounds(kd)
            assert np.array_equal(mn, rmn) and np.array_equal(inv, rinv, equal_nan=True)
            codes, perm = obj.sorted()
            assert np.array_equal(codes, ref.codes)
            assert np.array_equal(perm, ref.perm)
            assert np.array_equal(obj.splits(), ref.splits)
            assert np.array_equal(obj.sorted_kdops(), kd[perm])
            pats = obj.local_patches()
            for k, p in enumerate(pats):
                assert p["size"] == ref.pend[k] - ref.pbegin[k]
                if p["size"]:
                    assert np.array_equal(p["kdop"], ref.patch_kdops[k])
                    assert np.array_equal(p["centroid"], ref.patch_centroids[k])


@pytest.mark.parametrize("nq,od,bits", [(64, 16, 63), (64, 16, 30), (128, 64, 63), (96, 1, 63)])
def test_build_stages_wavy(torch_cuda, nq, od, bits):
    a, b = W.wavy_sheets(nq)
    ref, kd, ce = _oracle_object("poly", b, 0.0, od, bits)
    with CollisionWorld(od, morton_bits=bits) as w:
        obj = gpu_object(w, "poly", b, 0.0, device=True)
        codes, perm = obj.sorted()
        assert np.array_equal(codes, ref.codes)
        assert np.array_equal(perm, ref.perm)
        assert np.array_equal(obj.splits(), ref.splits)
        assert np.array_equal(obj.sorted_kdops(), kd[perm])


# ---- pair sets ----------------------------------------------------------------------
def _collide_both(kind, da, db, eps, od, bits, leaf, device=True):
    ra, ka, _ = _oracle_object(kind, da, eps, od, bits, leaf)
    rb, kb, _ = _oracle_object(kind, db, eps, od, bits, leaf)
    ref = ra.collide(rb)
    brute = O.brute_pairs(ka, kb)
    with CollisionWorld(od, morton_bits=bits, leaf_size=leaf) as w:
        oa = gpu_object(w, kind, da, eps, device=device)
        ob = gpu_object(w, kind, db, eps, device=device)
        oa.broadphase(ob)
        recs = sort_records(oa.results())
        pp = oa.patch_pairs()
    return recs, ref, brute, pp, ra, rb


@pytest.mark.parametrize("leaf", [1, 4])
def test_pairs_flat_sheets(torch_cuda, leaf):
    a, b = W.flat_sheets(64)
    # strict overlap with eps = 0: B is 0.02 above A, so the set is empty
    recs, ref, brute, _, _, _ = _collide_both("poly", a, b, 0.0, 16, 63, leaf)
    assert len(brute) == 0 and len(ref) == 0 and len(recs) == 0
    # from_vertices epsilon 0.01 (kdop.hpp:238-264): non-empty set
    recs, ref, brute, pp, ra, rb = _collide_both("poly", a, b, 0.01, 16, 63, leaf)
    assert len(brute) > 0
    assert np.array_equal(recs, sort_records(ref))
    assert np.array_equal(sort_pairs(recs[:, [1, 3]]), sort_pairs(brute))


@pytest.mark.parametrize("nq,od,bits,leaf", [(64, 16, 63, 1), (64, 16, 30, 4), (128, 32, 63, 4),
                                             (128, 1, 63, 2)])
def test_pairs_wavy_sheets(torch_cuda, nq, od, bits, leaf):
    a, b = W.wavy_sheets(nq)
    recs, ref, brute, pp, ra, rb = _collide_both("poly", a, b, 0.0, od, bits, leaf)
    assert len(brute) > 0
    assert np.array_equal(recs, sort_records(ref))
    assert np.array_equal(sort_pairs(recs[:, [1, 3]]), sort_pairs(brute))
    # broadphase patch pairs: every pair of non-empty patches with overlapping k-DOPs
    exp = []
    for p in range(od):
        if ra.pend[p] == ra.pbegin[p]:
            continue
        for q in range(od):
            if rb.pend[q] > rb.pbegin[q] and O.overlap(ra.patch_kdops[p], rb.patch_kdops[q]):
                exp.append((p, q))
    assert np.array_equal(pp, np.array(exp, np.uint64).reshape(-1, 2))


def test_pairs_spheres(torch_cuda):
    a = W.sphere_cloud(20000, radius=5e-3, ncomp=8, sigma=0.05, shift=(0, 0, 0), seed=11)
    b = W.sphere_cloud(20000, radius=5e-3, ncomp=8, sigma=0.05, shift=(0, 0, 0), seed=12)
    recs, ref, brute, _, _, _ = _collide_both("sphere", a, b, 0.0, 8, 63, 1)
    assert len(brute) > 0
    assert np.array_equal(recs, sort_records(ref))


def test_self_collision_spheres(torch_cuda):
    s = W.sphere_cloud(30000, radius=5e-3, ncomp=8, sigma=0.05, shift=(0, 0, 0), seed=13)
    kd, _ = O.elements_sphere(s)
    brute = O.brute_self(kd)
    with CollisionWorld(8, leaf_size=2) as w:
        o = gpu_object(w, "sphere", s, device=True)
        o.self_collision()
        got = sort_pairs(o.pairs())
    assert len(brute) > 0
    assert np.array_equal(got, sort_pairs(brute))


def test_collision_known_answers(torch_cuda):
    # tests/CollisionTest.cpp:43-81
    g = W.hex_grid(4, 4, 4)
    cube = W.hex_grid(1, 1, 1)
    with CollisionWorld(1) as w:
        oa = gpu_object(w, "poly", g)
        ob = gpu_object(w, "poly", cube)
        oa.broadphase(ob)
        assert oa.result_count() == 64
    with CollisionWorld(1) as w:
        o = gpu_object(w, "poly", W.hex_grid(2, 2, 1))
        o.self_collision()
        assert o.result_count() == 0
    with CollisionWorld(1) as w:
        both = np.concatenate([W.hex_grid(2, 2, 2), cube], axis=2)
        o = gpu_object(w, "poly", both)
        o.self_collision()
        assert o.result_count() == 8


def test_full_size_wavy_sheets(torch_cuda):
    """The benchmark workload at full size (2 x 2,097,152 triangles, 256
    patches, leaf size 4): sorted codes and permutation bit-exact, the
    broadphase patch pairs and the narrowphase records identical to the
    oracle's, plus invariants (codes sorted, permutation complete)."""
    import os
    a, b = W.wavy_sheets(1024)
    nt = min(len(os.sched_getaffinity(0)), 16)
    ka, ca = O.elements_poly(a)
    kb, cb = O.elements_poly(b)
    ra = O.Object(ka, ca, od=256, morton_bits=63, leaf_size=4, nthreads=nt)
    rb = O.Object(kb, cb, od=256, morton_bits=63, leaf_size=4, nthreads=nt)
    ref = sort_records(ra.collide(rb))
    with CollisionWorld(256, morton_bits=63, leaf_size=4) as w:
        oa = gpu_object(w, "poly", a, device=True)
        ob = gpu_object(w, "poly", b, device=True)
        codes, perm = oa.sorted()
        assert np.all(codes[1:] >= codes[:-1])
        assert np.array_equal(np.sort(perm), np.arange(len(perm), dtype=np.uint32))
        assert np.array_equal(codes, ra.codes) and np.array_equal(perm, ra.perm)
        assert np.array_equal(oa.splits(), ra.splits)
        oa.broadphase(ob)
        recs = sort_records(oa.results())
        pp = oa.patch_pairs()
    assert len(ref) > 0
    assert np.array_equal(recs, ref)
    # every narrowphase record lies in a broadphase patch pair
    pset = set(map(tuple, pp.tolist()))
    assert all((p, q) in pset for p, q in set(map(tuple, recs[:, [0, 2]].tolist())))


def _check_tree(tree, skd, n, od, splits):
    """Structural invariants of the radix tree: internal node k splits
    positions k | k+1, a node's range is the union of its children's, its
    slabs are the exact (min, max) union of theirs, leaf p holds the k-DOP of
    sorted element p, every node is reachable from the root once, and every
    non-empty patch is one subtree."""
    kd, ch, rg, root = tree["kdop"], tree["children"], tree["range"], tree["root"]
    assert len(kd) == 2 * n - 1
    leaf0 = n - 1
    assert np.array_equal(kd[leaf0:], skd)
    assert np.array_equal(rg[leaf0:, 0], np.arange(n)) and np.array_equal(rg[leaf0:, 1], np.arange(n))
    assert np.all(ch[leaf0:] == -1)
    if n == 1:
        assert root == 0
        return
    assert tuple(rg[root]) == (0, n - 1)
    k = np.arange(n - 1)
    left, right = ch[:n - 1, 0], ch[:n - 1, 1]
    assert np.all((left >= 0) & (left < 2 * n - 1) & (right >= 0) & (right < 2 * n - 1))
    assert np.array_equal(rg[left, 1], k.astype(np.uint64))
    assert np.array_equal(rg[right, 0], (k + 1).astype(np.uint64))
    assert np.array_equal(rg[:n - 1, 0], rg[left, 0]) and np.array_equal(rg[:n - 1, 1], rg[right, 1])
    lo = np.minimum(kd[left][:, 0::2], kd[right][:, 0::2])
    hi = np.maximum(kd[left][:, 1::2], kd[right][:, 1::2])
    assert np.array_equal(kd[:n - 1, 0::2], lo) and np.array_equal(kd[:n - 1, 1::2], hi)
    # reachability: every node has exactly one parent, except the root
    parents = np.zeros(2 * n - 1, np.int64)
    np.add.at(parents, left, 1)
    np.add.at(parents, right, 1)
    assert parents[root] == 0 and np.all(np.delete(parents, root) == 1)
    # patches are subtrees: a node spans exactly [begin, end) of each non-empty patch
    b = np.concatenate([[0], splits]).astype(np.uint64)
    e = np.concatenate([splits, [n]]).astype(np.uint64)
    spans = set(map(tuple, rg.tolist()))
    for p in range(od):
        if e[p] > b[p]:
            assert (int(b[p]), int(e[p]) - 1) in spans, p


@pytest.mark.parametrize("case", ["wavy64_od16", "wavy200_od64", "spheres_od8", "wavy1024_od256", "flat_od1",
                                  # around the build phases' chunk boundaries (128 leaves, x128 per phase)
                                  "spheres_n127", "spheres_n128", "spheres_n129", "spheres_n16383",
                                  "spheres_n16384", "spheres_n16385", "spheres_n2097153"])
def test_tree_structure(torch_cuda, case):
    if case.startswith("spheres_n"):
        n = int(case.split("_n")[1])
        data = W.sphere_cloud(n, radius=5e-3, ncomp=8, sigma=0.05, shift=(0, 0, 0), seed=n)
        kind, od = "sphere", 8
    elif case.startswith("wavy"):
        nq = int(case[4:].split("_")[0])
        od = int(case.split("od")[1])
        data, _ = W.wavy_sheets(nq)
        kind = "poly"
    elif case == "spheres_od8":
        data = W.sphere_cloud(70001, radius=5e-3, ncomp=8, sigma=0.05, shift=(0, 0, 0), seed=31)
        kind, od = "sphere", 8
    else:
        data, _ = W.flat_sheets(64)
        kind, od = "poly", 1
    with CollisionWorld(od, morton_bits=63, leaf_size=4, precision="fp64_records") as w:
        obj = gpu_object(w, kind, data, device=True)
        tree = obj.tree()
        skd = obj.sorted_kdops()
        splits = obj.splits()
        n = skd.shape[0]
    _check_tree(tree, skd, n, od, splits)


@pytest.mark.parametrize("precision", FORMS)
def test_refit_pairs(torch_cuda, precision):
    """bvh_refit keeps the topology and recomputes every k-DOP: after moving
    sheet B the pair set is brute force on the moved sheet, and refitting back
    gives the original set again."""
    a, b0 = W.wavy_sheets(64)
    rng = np.random.default_rng(7)
    b1 = b0 + np.array([0.003, -0.002, -0.001]).reshape(1, 3, 1) + rng.normal(0, 1e-4, b0.shape)
    ka, _ = O.elements_poly(a)
    want0 = sort_pairs(O.brute_pairs(ka, O.elements_poly(b0)[0]))
    want1 = sort_pairs(O.brute_pairs(ka, O.elements_poly(b1)[0]))
    assert len(want0) > 0 and len(want1) > 0 and not np.array_equal(want0, want1)
    with CollisionWorld(16, leaf_size=4, precision=precision) as w:
        oa = gpu_object(w, "poly", a, device=True)
        ob = gpu_object(w, "poly", b0, device=True)
        t1 = to_cuda(b1)
        ob.refit(t1)
        oa.broadphase(ob)
        got1 = sort_pairs(oa.results()[:, [1, 3]])
        ob.refit(to_cuda(b0))
        oa.broadphase(ob)
        got0 = sort_pairs(oa.results()[:, [1, 3]])
    assert np.array_equal(got1, want1)
    assert np.array_equal(got0, want0)
