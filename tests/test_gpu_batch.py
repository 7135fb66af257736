"""Batching form: a whole step (build + self collision of a batch of meshes)
captured into one HIP graph and replayed on new element data in place; the
autotuner.  Every replay is checked against brute force on the data
it ran on, with eager GPU work and synchronizations between replays."""
import numpy as np
import pytest

import oracle as O
from distbvh_amd import CollisionWorld
from distbvh_amd import capi
from distbvh_amd.capi import BvhError
from distbvh_amd import workloads as W
from helpers import FORMS, gpu_object, pair_digest, sort_pairs, to_cuda

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU test needs a GPU"
    return torch


def _batch(seed):
    a, _ = W.patch_batch(npatch=64, median=256, lo=16, hi=2048, lattice=4, seed=seed)
    return a


def _brute_self(a):
    kd, _ = O.elements_poly(a)
    return sort_pairs(O.brute_self(kd))


@pytest.mark.parametrize("precision", FORMS)
def test_graph_replays_match_brute_force(torch_cuda, precision):
    torch = torch_cuda
    a0 = _batch(0xC4)
    n = a0.shape[2]
    # same topology of the batch, new heights: the same element count
    datas = [a0]
    for k in range(3):
        d = a0.copy()
        d[:, 2, :] += 0.002 * np.sin(np.arange(n) * (0.37 + k))[None, :]
        datas.append(d)
    want = [_brute_self(d) for d in datas]
    assert all(len(w) > 0 for w in want)
    buf = to_cuda(a0)
    with CollisionWorld(64, leaf_size=4, precision=precision) as w:
        o = w.create_collision_object()

        def step():
            o.set_entity_data_poly(buf)
            o.init_broadphase()
            o.self_collision()

        step()  # eager: sizes every buffer
        assert o.result_count() == len(want[0])
        assert o.digest() == pair_digest(want[0])
        with w.capture() as g:
            step()
        try:
            for rep in range(2):
                for d, wp in zip(datas, want):
                    buf.copy_(to_cuda(d))
                    torch.cuda.synchronize()
                    junk = torch.randn(1 << 16, device="cuda").sum().item()  # eager work between replays
                    assert junk == junk
                    g.launch()
                    assert o.result_count() == len(wp)
                    assert o.digest() == pair_digest(wp)
            # the last replay's pairs, element for element
            assert np.array_equal(sort_pairs(o.pairs()), want[-1])
        finally:
            g.close()


def test_graph_overflow_is_reported(torch_cuda):
    torch = torch_cuda
    a0 = _batch(0xC5)
    buf = to_cuda(a0)
    with CollisionWorld(64, leaf_size=4) as w:
        o = w.create_collision_object()
        o.set_entity_data_poly(buf)
        o.init_broadphase()
        o.self_collision()
        cap_pairs = max(1 << 20, 8 * a0.shape[2])
        base = o.result_count()
        with w.capture() as g:
            o.set_entity_data_poly(buf)
            o.init_broadphase()
            o.self_collision()
        try:
            # every element a copy of element 0: all n (n - 1) / 2 pairs
            # overlap, far more than the pair buffer holds
            d = np.ascontiguousarray(np.broadcast_to(a0[:, :, :1], a0.shape))
            n = a0.shape[2]
            if n * (n - 1) // 2 <= cap_pairs:
                pytest.skip("batch too small to overflow the pair buffer")
            buf.copy_(to_cuda(d))
            torch.cuda.synchronize()
            g.launch()
            with pytest.raises(RuntimeError, match="outgrew"):
                o.result_count()
            # an eager run grows the buffers and gives the full set
            o.set_entity_data_poly(buf)
            o.init_broadphase()
            o.self_collision()
            assert o.result_count() > cap_pairs
            # the capture refers to freed buffers now and is refused
            with pytest.raises(RuntimeError, match="reallocated"):
                g.launch()
        finally:
            g.close()
        assert base > 0


def test_capture_refuses_buffer_growth(torch_cuda):
    a, _ = W.wavy_sheets(32)
    big, _ = W.wavy_sheets(64)
    with CollisionWorld(16, leaf_size=4) as w:
        o = w.create_collision_object()
        o.set_entity_data_poly(to_cuda(a))
        with pytest.raises(RuntimeError, match="grow inside a capture"):
            with w.capture():
                o.set_entity_data_poly(to_cuda(big))
        # host data cannot be captured either
        with pytest.raises(RuntimeError, match="device element data"):
            with w.capture():
                o.set_entity_data_poly(a)
        # the world still works eagerly
        o.set_entity_data_poly(to_cuda(big))
        o.self_collision()
        kd, _ = O.elements_poly(big)
        assert np.array_equal(sort_pairs(o.pairs()), sort_pairs(O.brute_self(kd)))


@pytest.mark.parametrize("precision", FORMS)
def test_autotune(torch_cuda, precision):
    a, b = W.wavy_sheets(64)
    ka, _ = O.elements_poly(a)
    kb, _ = O.elements_poly(b)
    want = sort_pairs(O.brute_pairs(ka, kb))
    with CollisionWorld(16, leaf_size=1, precision=precision) as w:
        oa = gpu_object(w, "poly", a, device=True)
        ob = gpu_object(w, "poly", b, device=True)
        t = oa.autotune(ob)
        assert t["batch"] in (1, 2, 4)
        assert t["block_size"] in (64, 128, 256)
        # the fp32-outward record walks take one record per round trip
        assert len(t["tried_ms"]) == (18 if precision == "fp64_records" else 6)
        assert all(x > 0 for x in t["tried_ms"]) and t["ms"] == min(t["tried_ms"])
        assert np.array_equal(sort_pairs(oa.results()[:, [1, 3]]), want)
        oa.broadphase(ob)  # the chosen setting stays
        assert np.array_equal(sort_pairs(oa.results()[:, [1, 3]]), want)
    # every launch shape gives the same pairs (self collision of the second sheet)
    kd = kb
    want_self = sort_pairs(O.brute_self(kd))
    with CollisionWorld(16, leaf_size=2, precision=precision) as w:
        o = gpu_object(w, "poly", b, device=True)
        t = o.autotune()
        assert np.array_equal(sort_pairs(o.pairs()), want_self)


@pytest.mark.parametrize("precision", FORMS)
def test_launch_shapes(torch_cuda, precision):
    """Every batch (records per round trip), block size and XCD remap gives
    the oracle's pairs, for two objects and for self collision."""
    a, b = W.wavy_sheets(48)
    ka, _ = O.elements_poly(a)
    kb, _ = O.elements_poly(b)
    want = sort_pairs(O.brute_pairs(ka, kb))
    want_self = sort_pairs(O.brute_self(kb))
    assert len(want) > 0 and len(want_self) > 0
    with CollisionWorld(16, leaf_size=1, precision=precision) as w:
        assert w.launch() == {"batch": 2, "block_size": 256, "xcd_remap": True}
        oa = gpu_object(w, "poly", a, device=True)
        ob = gpu_object(w, "poly", b, device=True)
        for batch in (1, 2, 4):
            for block in (64, 256):
                for xcd in (False, True):
                    w.set_launch(batch=batch, block_size=block, xcd_remap=xcd)
                    assert w.launch() == {"batch": batch, "block_size": block, "xcd_remap": xcd}
                    oa.broadphase(ob)
                    assert np.array_equal(sort_pairs(oa.results()[:, [1, 3]]), want), (batch, block, xcd)
                    ob.self_collision()
                    assert np.array_equal(sort_pairs(ob.pairs()), want_self), (batch, block, xcd)
        with pytest.raises(BvhError):
            w.set_launch(batch=3)
        with pytest.raises(BvhError):
            w.set_launch(block_size=96)


@pytest.mark.parametrize("precision", FORMS)
def test_launch_shapes_full_size(torch_cuda, precision):
    """2 x 2,097,152 triangles: the batches give one pair set (digest and
    count), and the same node visits (the visited set does not depend on the
    batch)."""
    import torch

    a, b = W.wavy_sheets(1024)
    ta = torch.from_numpy(a).cuda()
    tb = torch.from_numpy(b).cuda()
    with CollisionWorld(256, leaf_size=4, precision=precision) as w:
        oa = w.create_collision_object()
        ob = w.create_collision_object()
        oa.set_entity_data_poly(ta)
        ob.set_entity_data_poly(tb)
        oa.init_broadphase()
        ob.init_broadphase()
        seen = set()
        for batch in (1, 2, 4):
            w.set_launch(batch=batch)
            oa.broadphase(ob)
            st = oa.result_stats()
            seen.add((oa.result_count(), oa.digest(), st["node_visits"]))
        assert len(seen) == 1, seen


# ---- patches given by the batch, pairs between patches ---------------------------------------
def _mesh_of(counts):
    return np.repeat(np.arange(len(counts)), counts)


@pytest.mark.parametrize("precision", FORMS)
def test_given_patches_cross_patch_pairs(torch_cuda, precision):
    """One patch per mesh of the batch (BVH_SPLIT_GIVEN): the patches are the
    meshes, plain self collision is brute force, and cross-patch self
    collision is brute force without the pairs inside one mesh."""
    from distbvh_amd import GIVEN
    a, counts = W.patch_batch(npatch=64, median=256, lo=16, hi=2048, lattice=4, seed=0xC6)
    kd, _ = O.elements_poly(a)
    allp = sort_pairs(O.brute_self(kd))
    mesh = _mesh_of(counts)
    cross = allp[mesh[allp[:, 0].astype(np.int64)] != mesh[allp[:, 1].astype(np.int64)]]
    assert 0 < len(cross) < len(allp)
    offsets = np.cumsum(counts)[:-1]
    with CollisionWorld(64, precision=precision) as w:
        o = w.create_collision_object()
        o.set_patch_offsets(offsets)
        o.set_entity_data_poly(to_cuda(a), split_algorithm=GIVEN)
        o.init_broadphase()
        pats = o.local_patches()
        assert [p["size"] for p in pats] == [int(c) for c in counts]
        starts = np.concatenate([[0], offsets])
        for k in (0, 17, 63):
            assert np.array_equal(o.patch_elements(k), np.arange(starts[k], starts[k] + counts[k], dtype=np.uint64))
        o.self_collision()
        assert np.array_equal(sort_pairs(o.pairs()), allp)
        o.self_collision_cross_patch()
        got = sort_pairs(o.pairs())
        stats_cross = o.result_stats()
    assert np.array_equal(got, cross)
    if precision != "fp64_records":
        # the vertex-sharing pairs of a mesh are never candidates in cross-patch
        # mode, so they are not re-checked either
        assert stats_cross["rechecked_pairs"] < len(allp) - len(cross)


def test_cross_patch_graph_replays(torch_cuda):
    """The cross-patch step captured in a graph and replayed on new heights:
    every replay is the brute-force cross-mesh set of its data."""
    torch = torch_cuda
    from distbvh_amd import GIVEN
    a0, counts = W.patch_batch(npatch=64, median=256, lo=16, hi=2048, lattice=4, seed=0xC7)
    n = a0.shape[2]
    mesh = _mesh_of(counts)
    datas = [a0]
    for k in range(2):
        d = a0.copy()
        d[:, 2, :] += 0.003 * np.cos(np.arange(n) * (0.11 + k))[None, :]
        datas.append(d)

    def want(d):
        kd, _ = O.elements_poly(d)
        p = sort_pairs(O.brute_self(kd))
        return p[mesh[p[:, 0].astype(np.int64)] != mesh[p[:, 1].astype(np.int64)]]

    wants = [want(d) for d in datas]
    buf = to_cuda(a0)
    with CollisionWorld(64) as w:
        o = w.create_collision_object()
        o.set_patch_offsets(np.cumsum(counts)[:-1])

        def step():
            o.set_entity_data_poly(buf, split_algorithm=GIVEN)
            o.init_broadphase()
            o.self_collision_cross_patch()

        step()
        with w.capture() as g:
            step()
        try:
            for d, wp in zip(datas, wants):
                buf.copy_(to_cuda(d))
                torch.cuda.synchronize()
                g.launch()
                assert o.result_count() == len(wp)
                assert o.digest() == pair_digest(wp)
        finally:
            g.close()
