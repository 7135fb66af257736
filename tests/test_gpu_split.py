"""geom_axis / ml_geom_axis splitters on the GPU (distbvh_amd/csrc/split.hip)
against the oracle (oracle/bvh_oracle.c orc_split_geom_axis /
orc_split_ml_geom_axis, pinned by tests/test_oracle_split.py): the reference
order (the patches' elements in order) and the split offsets bit for bit,
the patch k-DOPs and centroids bit for bit, and the pairs against brute force
(the pair set does not depend on the split algorithm)."""
import numpy as np
import pytest

import oracle as O
from distbvh_amd import CollisionWorld
from distbvh_amd import capi
from distbvh_amd import workloads as W
from helpers import FORMS, sort_pairs, to_cuda

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU test needs a GPU"
    return torch


SPLITS = [(capi.BVH_SPLIT_GEOM_AXIS, O.SPLIT_GEOM_AXIS, False),
          (capi.BVH_SPLIT_ML_GEOM_AXIS, O.SPLIT_ML_GEOM_AXIS, True)]
SPLIT_IDS = ["geom_axis", "ml_geom_axis"]


def _build(w, kind, data, split, eps=0.0):
    o = w.create_collision_object()
    if kind == "poly":
        o.set_entity_data_poly(to_cuda(data), epsilon=eps, split_algorithm=split)
    else:
        o.set_entity_data_sphere(to_cuda(data), split_algorithm=split)
    o.init_broadphase()
    return o


def _elements(kind, data, eps=0.0):
    return O.elements_poly(data, eps) if kind == "poly" else O.elements_sphere(data)


def _check(kind, data, od, split, osplit, ml, eps=0.0, precision="fp64"):
    kd, ce = _elements(kind, data, eps)
    depth = od.bit_length() - 1
    idx, sp = O.split_permutations(kd, ce, depth, ml=ml)
    ref = O.Object(kd, ce, od=od, split=osplit)
    with CollisionWorld(od, precision=precision) as w:
        o = _build(w, kind, data, split, eps)
        parts = [o.patch_elements(k) for k in range(od)]
        order = np.concatenate(parts).astype(np.uint32) if parts else np.zeros(0, np.uint32)
        assert np.array_equal(order, idx), "reference order"
        assert np.array_equal(np.cumsum([len(p) for p in parts])[:-1].astype(np.uint64), sp), "splits"
        pats = o.local_patches()
        for k, p in enumerate(pats):
            assert p["size"] == ref.pend[k] - ref.pbegin[k]
            if p["size"]:
                assert np.array_equal(p["kdop"], ref.patch_kdops[k]), k
                assert np.array_equal(p["centroid"], ref.patch_centroids[k]), k
        # the tree over these patches: self collision against brute force
        o.self_collision()
        assert np.array_equal(sort_pairs(o.pairs()), sort_pairs(O.brute_self(kd)))


def _line(n, r=1.0):
    xyzr = np.zeros((4, n))
    xyzr[0] = np.arange(n)
    xyzr[3] = r
    return xyzr


def _grid4x3():
    xyzr = np.zeros((4, 12))
    for j in range(3):
        for i in range(4):
            xyzr[:3, i + 4 * j] = (i, j, 0.0)
    xyzr[3] = 1.0
    return xyzr


@pytest.mark.parametrize("split,osplit,ml", SPLITS, ids=SPLIT_IDS)
@pytest.mark.parametrize("od", [1, 2, 4, 8, 16])
def test_split_known_answer_shapes(torch_cuda, split, osplit, ml, od):
    """tests/SplitTest.cpp's inputs: 8 points on a line, the 4 x 3 grid."""
    _check("sphere", _line(8), od, split, osplit, ml)
    _check("sphere", _grid4x3(), od, split, osplit, ml)


@pytest.mark.parametrize("split,osplit,ml", SPLITS, ids=SPLIT_IDS)
@pytest.mark.parametrize("seed,n,od", [(1, 5000, 16), (2, 20000, 64), (3, 777, 256)])
def test_split_random_spheres(torch_cuda, split, osplit, ml, seed, n, od):
    rng = np.random.default_rng(seed)
    xyzr = np.zeros((4, n))
    xyzr[:3] = rng.normal(size=(3, n)) * np.array([[3.0], [1.0], [0.3]])
    xyzr[3] = 0.02
    _check("sphere", xyzr, od, split, osplit, ml)


@pytest.mark.parametrize("split,osplit,ml", SPLITS, ids=SPLIT_IDS)
@pytest.mark.parametrize("precision", FORMS)
def test_split_wavy_sheets(torch_cuda, split, osplit, ml, precision):
    a, b = W.wavy_sheets(64)
    _check("poly", b, 32, split, osplit, ml, precision=precision)
    ka, _ = O.elements_poly(a)
    kb, _ = O.elements_poly(b)
    with CollisionWorld(32, precision=precision) as w:
        oa = _build(w, "poly", a, split)
        ob = _build(w, "poly", b, split)
        oa.broadphase(ob)
        assert np.array_equal(sort_pairs(oa.results()[:, [1, 3]]), sort_pairs(O.brute_pairs(ka, kb)))


@pytest.mark.parametrize("split,osplit,ml", SPLITS, ids=SPLIT_IDS)
def test_split_exact_split_point(torch_cuda, split, osplit, ml):
    """Projections on the split point itself and a sum whose rounding depends
    on the order (2^53 + 1 + 1 - 2^53): the split point must be the
    reference's, summed in its order."""
    xs = [2.0 ** 53, 1.0, 1.0, -(2.0 ** 53)] + [0.25 * k for k in range(-20, 21)] + [0.5, 0.5, 0.0]
    n = len(xs)
    xyzr = np.zeros((4, n))
    xyzr[0] = xs
    xyzr[3] = 0.125
    for od in (2, 4, 8):
        _check("sphere", xyzr, od, split, osplit, ml)
    _check("sphere", _line(9), 2, split, osplit, ml)  # the middle point is the mean


@pytest.mark.parametrize("split,osplit,ml", SPLITS, ids=SPLIT_IDS)
def test_split_full_size(torch_cuda, split, osplit, ml):
    """2 x 2,097,152 triangles, 256 patches: the reference order and splits
    of the oracle, and the pairs of the clustering split (digest)."""
    import torch

    a, b = W.wavy_sheets(1024)
    kb, cb = O.elements_poly(b)
    idx, sp = O.split_permutations(kb, cb, 8, ml=ml)
    ta = torch.from_numpy(a).cuda()
    tb = torch.from_numpy(b).cuda()
    with CollisionWorld(256) as w:
        ob = w.create_collision_object()
        ob.set_entity_data_poly(tb, split_algorithm=split)
        oa = w.create_collision_object()
        oa.set_entity_data_poly(ta, split_algorithm=split)
        oa.init_broadphase()
        ob.init_broadphase()
        parts = [ob.patch_elements(k) for k in range(256)]
        assert np.array_equal(np.concatenate(parts).astype(np.uint32), idx)
        assert np.array_equal(np.cumsum([len(p) for p in parts])[:-1].astype(np.uint64), sp)
        oa.broadphase(ob)
        got = (oa.result_count(), oa.digest())
        ca = w.create_collision_object()
        cbo = w.create_collision_object()
        ca.set_entity_data_poly(ta)
        cbo.set_entity_data_poly(tb)
        ca.init_broadphase()
        cbo.init_broadphase()
        ca.broadphase(cbo)
        assert got == (ca.result_count(), ca.digest())


def test_split_needs_power_of_two_patches(torch_cuda):
    """geom_axis requires od = 2^depth (collision_object.cpp:122-124); the
    world already refuses other factors."""
    with pytest.raises(capi.BvhError):
        CollisionWorld(12)
