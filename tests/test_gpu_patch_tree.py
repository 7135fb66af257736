"""for_each_tree's patch-level tree (bvh_object_patch_tree): one leaf per
non-empty patch, left to right (collision_object_test.cpp:49-81 checks
count() == od and 0 for an empty object), in pre-order with relative
offsets as bvh_node stores them (node.hpp:318-322); leaf k-DOPs are the
patches' exact slabs, inner k-DOPs the unions of their children."""
import numpy as np
import pytest

import oracle as O
from distbvh_amd import CollisionWorld
from distbvh_amd import capi
from distbvh_amd import workloads as W
from helpers import FORMS, to_cuda

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU test needs a GPU"
    return torch


def _check_tree(obj):
    nodes, leaves = obj.patch_tree()
    pats = obj.local_patches()
    nonempty = [k for k, p in enumerate(pats) if p["size"] > 0]
    assert leaves.tolist() == nonempty
    assert len(nodes) == (2 * len(nonempty) - 1 if nonempty else 0)
    if not nodes:
        return
    assert nodes[0]["parent_offset"] == 0
    assert nodes[0]["entity_range"] == (0, len(nonempty))
    seen = 0
    for k, nd in enumerate(nodes):
        l, r = nd["child_offsets"]
        if k:
            p = k + nd["parent_offset"]
            assert 0 <= p < k and k in (p + nodes[p]["child_offsets"][0], p + nodes[p]["child_offsets"][1])
        if l == 0 and r == 0:
            first, last = nd["entity_range"]
            assert (first, last) == (seen, seen + 1)
            assert np.array_equal(nd["kdop"], pats[int(leaves[first])]["kdop"])
            seen += 1
        else:
            assert l == 1 and r > 1  # pre-order: the left child follows its parent
            a, b = nodes[k + l], nodes[k + r]
            assert nd["entity_range"] == (a["entity_range"][0], b["entity_range"][1])
            assert a["entity_range"][1] == b["entity_range"][0]
            want = np.empty(26)
            want[0::2] = np.minimum(a["kdop"][0::2], b["kdop"][0::2])
            want[1::2] = np.maximum(a["kdop"][1::2], b["kdop"][1::2])
            assert np.array_equal(nd["kdop"], want)
    assert seen == len(nonempty)


@pytest.mark.parametrize("precision", FORMS)
@pytest.mark.parametrize("split", [capi.BVH_SPLIT_CLUSTERING, capi.BVH_SPLIT_GEOM_AXIS,
                                   capi.BVH_SPLIT_ML_GEOM_AXIS])
@pytest.mark.parametrize("od", [1, 16, 64])
def test_patch_tree_wavy(torch_cuda, precision, split, od):
    _, b = W.wavy_sheets(64)
    with CollisionWorld(od, precision=precision) as w:
        o = w.create_collision_object()
        o.set_entity_data_poly(to_cuda(b), split_algorithm=split)
        o.init_broadphase()
        _check_tree(o)
        assert len(o.patch_tree()[1]) == od  # every patch non-empty


def test_patch_tree_sparse_and_empty(torch_cuda):
    rng = np.random.default_rng(4)
    xyzr = np.zeros((4, 20))
    xyzr[:3] = rng.uniform(size=(3, 20))
    xyzr[3] = 0.01
    with CollisionWorld(64) as w:
        o = w.create_collision_object()
        o.set_entity_data_sphere(to_cuda(xyzr))
        o.init_broadphase()
        _check_tree(o)
        assert 0 < len(o.patch_tree()[1]) <= 20
        o.set_entity_data_sphere(to_cuda(np.zeros((4, 0))))
        o.init_broadphase()
        assert o.patch_tree()[0] == [] and len(o.patch_tree()[1]) == 0
