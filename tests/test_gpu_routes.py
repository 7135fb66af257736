"""GPU tests of the entry routes beside the plain polytope path: the generic
ContactEntity route (bvh_set_entity_data_snapshot), caller global ids in the
result records, device-sorted results (BVH_FLAG_SORT_RESULTS), and the
capture / stream / argument rules of the host mirror.  Each result is checked
against the CPU oracle on the same inputs."""
import numpy as np
import pytest

import oracle as O
from distbvh_amd import CollisionWorld
from distbvh_amd import workloads as W
from distbvh_amd.capi import BvhError
from helpers import FORMS, gpu_object, pair_digest, sort_pairs, sort_records, to_cuda

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU test needs a GPU"
    return torch


def _gids(n, seed):
    """Distinct non-identity 64-bit global ids (high bits set, shuffled)."""
    rng = np.random.default_rng(seed)
    base = rng.permutation(n).astype(np.uint64) * np.uint64(7919) + np.uint64(0xABCD000000000000)
    return base


# ---- the generic ContactEntity route (snapshot.hpp:68-149) --------------------------------
@pytest.mark.parametrize("device", [False, True])
@pytest.mark.parametrize("bits,od", [(63, 16), (30, 32)])
def test_snapshot_route_matches_polytope_route(torch_cuda, device, bits, od):
    """Oracle k-DOPs and centroids handed in as snapshots give the codes,
    permutation, splits, sorted k-DOPs, patches and records of the oracle
    (and of the polytope route on the same elements)."""
    a, b = W.wavy_sheets(48)
    ka, ca = O.elements_poly(a)
    kb, cb = O.elements_poly(b)
    ra = O.Object(ka, ca, od=od, morton_bits=bits)
    rb = O.Object(kb, cb, od=od, morton_bits=bits)
    ref = sort_records(ra.collide(rb))
    assert len(ref) > 0
    with CollisionWorld(od, morton_bits=bits) as w:
        oa = w.create_collision_object()
        ob = w.create_collision_object()
        if device:
            oa.set_entity_data_snapshot(to_cuda(ka), to_cuda(ca))
            ob.set_entity_data_snapshot(to_cuda(kb), to_cuda(cb))
        else:
            oa.set_entity_data_snapshot(ka, ca)
            ob.set_entity_data_snapshot(kb, cb)
        oa.init_broadphase()
        ob.init_broadphase()
        codes, perm = oa.sorted()
        assert np.array_equal(codes, ra.codes)
        assert np.array_equal(perm, ra.perm)
        assert np.array_equal(oa.splits(), ra.splits)
        assert np.array_equal(oa.sorted_kdops(), ka[perm])
        for k, p in enumerate(oa.local_patches()):
            assert p["size"] == ra.pend[k] - ra.pbegin[k]
            if p["size"]:
                assert np.array_equal(p["kdop"], ra.patch_kdops[k])
                assert np.array_equal(p["centroid"], ra.patch_centroids[k])
        oa.broadphase(ob)
        got = sort_records(oa.results())
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("precision", FORMS)
def test_snapshot_route_refit_and_self(torch_cuda, precision):
    """Snapshots through refit and self collision, both forms, against brute
    force (the fp32 form re-checks on the caller's fp64 k-DOPs)."""
    s = W.sphere_cloud(12000, radius=4e-3, ncomp=8, sigma=0.05, shift=(5.0, -3.0, 2.0), seed=31)
    kd, ce = O.elements_sphere(s)
    s2 = s.copy()
    s2[2] += 1e-3
    kd2, _ = O.elements_sphere(s2)
    with CollisionWorld(8, precision=precision) as w:
        o = w.create_collision_object()
        o.set_entity_data_snapshot(to_cuda(kd), to_cuda(ce))
        o.init_broadphase()
        o.self_collision()
        assert np.array_equal(sort_pairs(o.pairs()), sort_pairs(O.brute_self(kd)))
        o.refit(to_cuda(kd2))
        o.self_collision()
        assert np.array_equal(sort_pairs(o.pairs()), sort_pairs(O.brute_self(kd2)))


# ---- global ids in the records (snapshot.hpp:68-125) ------------------------------------
@pytest.mark.parametrize("kind", ["poly", "sphere", "snapshot"])
@pytest.mark.parametrize("device", [False, True])
def test_records_carry_global_ids(torch_cuda, kind, device):
    if kind == "sphere":
        a = W.sphere_cloud(6000, radius=6e-3, ncomp=4, sigma=0.05, shift=(0, 0, 0), seed=41)
        b = W.sphere_cloud(6000, radius=6e-3, ncomp=4, sigma=0.05, shift=(0, 0, 0), seed=42)
        ka, ca = O.elements_sphere(a)
        kb, cb = O.elements_sphere(b)
    else:
        a, b = W.wavy_sheets(40)
        ka, ca = O.elements_poly(a)
        kb, cb = O.elements_poly(b)
    ga = _gids(ka.shape[0], 1)
    gb = _gids(kb.shape[0], 2)
    od = 8
    ra = O.Object(ka, ca, gids=ga, od=od)
    rb = O.Object(kb, cb, gids=gb, od=od)
    ref = sort_records(ra.collide(rb))
    assert len(ref) > 0
    conv = to_cuda if device else (lambda x: x)
    with CollisionWorld(od) as w:
        oa = w.create_collision_object()
        ob = w.create_collision_object()
        for o, d, k, c, g in ((oa, a, ka, ca, ga), (ob, b, kb, cb, gb)):
            if kind == "poly":
                o.set_entity_data_poly(conv(d), global_ids=conv(g))
            elif kind == "sphere":
                o.set_entity_data_sphere(conv(d), global_ids=conv(g))
            else:
                o.set_entity_data_snapshot(conv(k), conv(c), global_ids=conv(g))
            o.init_broadphase()
        oa.broadphase(ob)
        got = sort_records(oa.results())
        pairs = sort_pairs(oa.pairs())
    assert np.array_equal(got, ref)
    # the ids are the callers': element_p = ga[input index], element_q = gb[...]
    want = sort_records(np.stack([np.zeros(len(pairs), np.uint64), ga[pairs[:, 0].astype(np.int64)],
                                  np.zeros(len(pairs), np.uint64), gb[pairs[:, 1].astype(np.int64)]], axis=1))
    assert np.array_equal(np.sort(got[:, 1]), np.sort(want[:, 1]))
    assert np.array_equal(np.sort(got[:, 3]), np.sort(want[:, 3]))


# ---- device-sorted results -----------------------------------------------------------------
@pytest.mark.parametrize("precision", FORMS)
def test_sort_results_flag_gives_the_oracle_order(torch_cuda, precision):
    """BVH_FLAG_SORT_RESULTS: the pairs and records come in ascending (input
    index of element_p, input index of element_q) order, the order of the
    oracle's records, element for element (no host-side sort)."""
    a, b = W.wavy_sheets(64)
    ka, ca = O.elements_poly(a)
    kb, cb = O.elements_poly(b)
    od = 16
    ref = O.Object(ka, ca, od=od).collide(O.Object(kb, cb, od=od))
    assert len(ref) > 2
    with CollisionWorld(od, precision=precision, sort_results=True) as w:
        oa = gpu_object(w, "poly", a, device=True)
        ob = gpu_object(w, "poly", b, device=True)
        oa.broadphase(ob)
        pairs = oa.pairs()
        recs = oa.results()
    keys = (pairs[:, 0] << np.uint64(32)) | pairs[:, 1]
    assert np.all(keys[1:] > keys[:-1])
    assert np.array_equal(recs, ref)


def test_sort_results_self_collision(torch_cuda):
    s = W.sphere_cloud(20000, radius=5e-3, ncomp=8, sigma=0.05, shift=(0, 0, 0), seed=51)
    kd, _ = O.elements_sphere(s)
    brute = sort_pairs(O.brute_self(kd))
    with CollisionWorld(8, sort_results=True) as w:
        o = gpu_object(w, "sphere", s, device=True)
        o.self_collision()
        got = o.pairs()
    assert np.array_equal(got, brute)


# ---- capture rules -----------------------------------------------------------------------
def test_replay_restores_the_captured_query(torch_cuda):
    """A graph captured on a.broadphase(b), an eager a.self_collision() in
    between (which allocates nothing), then a replay: the results are the
    captured query's (partner b, patch ids of b), equal to the oracle's."""
    a, b = W.wavy_sheets(48)
    ka, ca = O.elements_poly(a)
    kb, cb = O.elements_poly(b)
    od = 8
    ref = sort_records(O.Object(ka, ca, od=od).collide(O.Object(kb, cb, od=od)))
    self_ref = sort_pairs(O.brute_self(ka))
    ta, tb = to_cuda(a), to_cuda(b)
    with CollisionWorld(od) as w:
        oa = w.create_collision_object()
        ob = w.create_collision_object()

        def step():
            oa.set_entity_data_poly(ta)
            ob.set_entity_data_poly(tb)
            oa.init_broadphase()
            ob.init_broadphase()
            oa.broadphase(ob)

        step()
        oa.self_collision()  # sizes the buffers of the eager query below
        step()
        with w.capture() as g:
            step()
        try:
            # nothing to read before the graph has run
            assert oa.result_count() == 0
            oa.self_collision()
            assert np.array_equal(sort_pairs(oa.pairs()), self_ref)
            g.launch()
            assert np.array_equal(sort_records(oa.results()), ref)
            assert oa.patch_pairs().shape[0] > 0
            oa.self_collision()
            g.launch()
            assert np.array_equal(sort_records(oa.results()), ref)
        finally:
            g.close()


def test_allocation_in_another_world_keeps_a_graph_launchable(torch_cuda):
    """Device allocations of one world do not make another world's graph
    stale; an allocation in the graph's own world does."""
    a, b = W.wavy_sheets(32)
    ta, tb = to_cuda(a), to_cuda(b)
    with CollisionWorld(4) as w, CollisionWorld(4) as other:
        oa = w.create_collision_object()
        ob = w.create_collision_object()

        def step():
            oa.set_entity_data_poly(ta)
            ob.set_entity_data_poly(tb)
            oa.init_broadphase()
            ob.init_broadphase()
            oa.broadphase(ob)

        step()
        want = (oa.result_count(), oa.digest())
        with w.capture() as g:
            step()
        try:
            x = gpu_object(other, "poly", b, device=True)  # allocates in the other world
            x.self_collision()
            g.launch()
            assert (oa.result_count(), oa.digest()) == want
            big, _ = W.wavy_sheets(96)
            oa.set_entity_data_poly(to_cuda(big))  # grows this world's buffers
            with pytest.raises(BvhError):
                g.launch()
        finally:
            g.close()


# ---- host mirror argument and stream rules --------------------------------------------------
def test_device_tensors_are_checked(torch_cuda):
    torch = torch_cuda
    a, _ = W.wavy_sheets(16)
    with CollisionWorld(2) as w:
        o = w.create_collision_object()
        with pytest.raises(ValueError):
            o.set_entity_data_poly(torch.from_numpy(a).float().cuda())  # float32
        aos = torch.from_numpy(np.ascontiguousarray(a.transpose(2, 0, 1))).cuda()  # (n, 3, 3)
        with pytest.raises(ValueError):
            o.set_entity_data_poly(aos.permute(1, 2, 0))  # planes-shaped view of element-major data
        with pytest.raises(ValueError):
            o.set_entity_data_poly(torch.from_numpy(a[:, :2]).cuda())  # two coordinates
        with pytest.raises(ValueError):
            o.set_entity_data_poly(to_cuda(a), global_ids=torch.zeros(5, dtype=torch.int64, device="cuda"))
        with pytest.raises(ValueError):
            o.set_entity_data_poly(to_cuda(a), global_ids=torch.zeros(a.shape[2], dtype=torch.float64,
                                                                      device="cuda"))
        # the contiguous float64 tensor is accepted
        o.set_entity_data_poly(to_cuda(a))
        o.init_broadphase()
        o.self_collision()
        assert o.result_count() == len(O.brute_self(O.elements_poly(a)[0]))


def test_world_stream_waits_for_torch_stream(torch_cuda):
    """Element data written by torch on a side stream and handed to a world
    (which runs on its own stream) without any synchronization: the world's
    stream waits for torch's work, so the pairs are those of the data."""
    torch = torch_cuda
    a, b = W.wavy_sheets(96)
    ka, _ = O.elements_poly(a)
    kb, _ = O.elements_poly(b)
    want = sort_pairs(O.brute_pairs(ka, kb))
    side = torch.cuda.Stream()
    with CollisionWorld(8) as w:
        oa = w.create_collision_object()
        ob = w.create_collision_object()
        with torch.cuda.stream(side):
            x = torch.rand(4096, 4096, device="cuda")
            for _ in range(8):
                x = torch.sin(x @ x)  # keeps the side stream busy; stays in [-1, 1]
            zero = (x[0, 0] - x[0, 0]).double()  # 0, known only when the side stream gets there
            ta = torch.from_numpy(a).cuda() + zero
            tb = torch.from_numpy(b).cuda() + zero
            oa.set_entity_data_poly(ta)
            ob.set_entity_data_poly(tb)
            oa.init_broadphase()
            ob.init_broadphase()
            oa.broadphase(ob)
            got = sort_pairs(oa.pairs())
    assert np.array_equal(got, want)
    assert pair_digest(got) == pair_digest(want)
