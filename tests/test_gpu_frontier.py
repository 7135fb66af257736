"""GPU parity of the walks over trees of many patches.  The fp32 walk starts
at the patch roots whose slabs overlap the union of a wave's queries (up to
256 patches; above that it starts at the root), so the cases here are the
ones that start touches: empty patches (no root), patches of one element
(a leaf is the patch root), patch counts at 256 and above it, and the
query kinds (A x B both ways, self; cross-patch self, which walks from the
root), each against brute force in every tree form."""
import numpy as np
import pytest

import oracle as O
from distbvh_amd import CollisionWorld
from distbvh_amd import workloads as W
from helpers import FORMS, sort_pairs, to_cuda

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU test needs a GPU"
    return torch


@pytest.fixture(scope="module")
def meshes():
    """A batch of small height-field meshes (B) and the same meshes lifted a
    little (A), so that A x B has many pairs; element k-DOPs of both."""
    b, _ = W.patch_batch(npatch=48, median=64, lo=16, hi=256, lattice=4, seed=0xF1)
    a = b.copy()
    a[:, 2, :] += 0.004
    ka, _ = O.elements_poly(a)
    kb, _ = O.elements_poly(b)
    return a, b, ka, kb


def _patch_sizes(n, od, seed):
    """od patch sizes summing to n: empty patches (first, last and some
    inside), one-element patches next to them, the rest random."""
    rng = np.random.default_rng(seed)
    sizes = np.zeros(od, np.int64)
    special = {0: 0, 1: 1, 2: 0, 3: 1, 4: 1, od // 2: 0, od // 2 + 1: 1, od - 2: 1, od - 1: 0}
    free = [k for k in range(od) if k not in special]
    rest = n - sum(special.values())
    assert rest >= len(free)
    cuts = np.sort(rng.choice(np.arange(1, rest), size=len(free) - 1, replace=False))
    sizes[free] = np.diff(np.concatenate([[0], cuts, [rest]]))
    for k, v in special.items():
        sizes[k] = v
    assert sizes.sum() == n and (sizes[free] > 0).all()
    return sizes


@pytest.mark.parametrize("precision", FORMS)
@pytest.mark.parametrize("od", [64, 256, 2048])
def test_given_patches_with_empty_and_single_patches(torch_cuda, meshes, precision, od):
    from distbvh_amd import GIVEN
    a, b, ka, kb = meshes
    n = b.shape[2]
    sizes = _patch_sizes(n, od, seed=od)
    offsets = np.cumsum(sizes)[:-1]
    patch = np.repeat(np.arange(od), sizes)
    want_ab = sort_pairs(O.brute_pairs(ka, kb))
    want_ba = sort_pairs(O.brute_pairs(kb, ka))
    want_self = sort_pairs(O.brute_self(kb))
    want_cross = want_self[patch[want_self[:, 0].astype(np.int64)] != patch[want_self[:, 1].astype(np.int64)]]
    assert len(want_ab) > 0 and 0 < len(want_cross) < len(want_self)
    with CollisionWorld(od, precision=precision) as w:
        ob = w.create_collision_object()
        ob.set_patch_offsets(offsets)
        ob.set_entity_data_poly(to_cuda(b), split_algorithm=GIVEN)
        ob.init_broadphase()
        assert [p["size"] for p in ob.local_patches()] == [int(s) for s in sizes]
        oa = w.create_collision_object()
        oa.set_entity_data_poly(to_cuda(a))
        oa.init_broadphase()
        oa.broadphase(ob)  # target: the given patches
        got_ab = sort_pairs(oa.results()[:, [1, 3]])
        ob.broadphase(oa)  # target: the clustering patches
        got_ba = sort_pairs(ob.results()[:, [1, 3]])
        ob.self_collision()
        got_self = sort_pairs(ob.pairs())
        ob.self_collision_cross_patch()
        got_cross = sort_pairs(ob.pairs())
    assert np.array_equal(got_ab, want_ab)
    assert np.array_equal(got_ba, want_ba)
    assert np.array_equal(got_self, want_self)
    assert np.array_equal(got_cross, want_cross)

