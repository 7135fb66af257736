"""The wire formats on one GPU (distbvh_amd/csrc/wire.hip): patch records,
element payloads and result messages packed on the device (and from / to
host memory), their byte layout, and round trips -- including a ghost
object rebuilt from the packed elements of the patches another object's
broadphase names, which must find the same element pairs."""
import numpy as np
import pytest

import oracle as O
from distbvh_amd import CollisionWorld, capi
from distbvh_amd import workloads as W
from helpers import FORMS, gpu_object, sort_pairs, sort_records, to_cuda

pytestmark = pytest.mark.gpu

PATCH = np.dtype([("gid", "<u8"), ("size", "<u8"), ("kdop", "<f8", (26,)), ("centroid", "<f8", (4,))])


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU test needs a GPU"
    return torch


def _gids(n, seed):
    rng = np.random.default_rng(seed)
    return rng.permutation(n).astype(np.uint64) * np.uint64(31) + np.uint64(0x5150000000000000)


@pytest.mark.parametrize("device", [False, True])
def test_patch_records(torch_cuda, device):
    a, _ = W.wavy_sheets(48)
    ka, ca = O.elements_poly(a)
    od = 16
    ref = O.Object(ka, ca, od=od)
    with CollisionWorld(od) as w:
        o = gpu_object(w, "poly", a, device=True)
        buf = o.pack_patches(device=device)
        host = buf.cpu().numpy() if device else buf
        assert host.nbytes == od * 256
        rec = host.view(PATCH)
        back = w.unpack_patches(buf)
        lp = o.local_patches()
    assert len(back) == od
    for k in range(od):
        assert rec[k]["gid"] == k and back[k]["global_id"] == k
        assert rec[k]["size"] == ref.pend[k] - ref.pbegin[k] == back[k]["size"] == lp[k]["size"]
        assert rec[k]["centroid"][3] == 0.0
        assert np.array_equal(rec[k]["centroid"][:3], lp[k]["centroid"])
        assert np.array_equal(back[k]["centroid"], lp[k]["centroid"])
        assert np.array_equal(rec[k]["kdop"], lp[k]["kdop"])
        if rec[k]["size"]:
            assert np.array_equal(rec[k]["kdop"], ref.patch_kdops[k])
            assert np.array_equal(rec[k]["centroid"][:3], ref.patch_centroids[k])


@pytest.mark.parametrize("kind", ["poly", "sphere", "snapshot"])
@pytest.mark.parametrize("device", [False, True])
def test_element_payload_round_trip(torch_cuda, kind, device):
    if kind == "sphere":
        d = W.sphere_cloud(5000, radius=5e-3, ncomp=4, sigma=0.05, shift=(1.0, 2.0, 3.0), seed=61)
        n = d.shape[1]
    else:
        d, _ = W.wavy_sheets(40)
        n = d.shape[2]
    g = _gids(n, 62)
    od = 8
    with CollisionWorld(od) as w:
        o = w.create_collision_object()
        if kind == "poly":
            o.set_entity_data_poly(to_cuda(d), global_ids=to_cuda(g))
            stride, ekind = 8 + 72, capi.BVH_ELEM_POLY
        elif kind == "sphere":
            o.set_entity_data_sphere(to_cuda(d), global_ids=to_cuda(g))
            stride, ekind = 8 + 32, capi.BVH_ELEM_SPHERE
        else:
            kd, ce = O.elements_poly(d)
            o.set_entity_data_snapshot(to_cuda(kd), to_cuda(ce), global_ids=to_cuda(g))
            stride, ekind = 8 + 8 * 29, capi.BVH_ELEM_SNAPSHOT
        o.init_broadphase()
        assert o.element_record_bytes() == stride
        patches = [5, 0, 3]
        idx = np.concatenate([o.patch_elements(k) for k in patches]).astype(np.int64)
        buf = o.pack_patch_elements(patches, device=device)
        assert (buf.numel() if device else buf.nbytes) == len(idx) * stride
        data, cent, gids = w.unpack_elements(buf, len(idx), ekind, 3)
        if device:
            data = data.cpu().numpy()
            cent = cent.cpu().numpy() if cent is not None else None
            gids = gids.cpu().numpy().view(np.uint64)
    assert np.array_equal(gids, g[idx])
    if kind == "poly":
        assert np.array_equal(data, d[:, :, idx])
    elif kind == "sphere":
        assert np.array_equal(data, d[:, idx])
    else:
        assert np.array_equal(data, kd[idx]) and np.array_equal(cent, ce[idx])
    # the raw bytes: global id, then the element's doubles
    raw = (buf.cpu().numpy() if device else buf).view(np.uint64).reshape(len(idx), stride // 8)
    assert np.array_equal(raw[:, 0], g[idx])


@pytest.mark.parametrize("precision", FORMS)
def test_ghost_from_packed_patches_finds_the_same_pairs(torch_cuda, precision):
    """What the reference's ghosting does with the payload: the patches of A
    that B's broadphase names are packed, unpacked into a ghost object of a
    second world, and the ghost collided with B there: the element pairs (by
    global id) are A's."""
    a, b = W.wavy_sheets(64)
    ga = _gids(a.shape[2], 71)
    od = 16
    with CollisionWorld(od, precision=precision) as w, CollisionWorld(od, precision=precision) as w2:
        oa = w.create_collision_object()
        oa.set_entity_data_poly(to_cuda(a), global_ids=to_cuda(ga))
        oa.init_broadphase()
        ob = gpu_object(w, "poly", b, device=True)
        oa.broadphase(ob)
        want = sort_pairs(np.stack([oa.results()[:, 1], oa.results()[:, 3]], axis=1))
        pp = oa.patch_pairs()
        named = sorted(set(int(p) for p in pp[:, 0]))
        assert 0 < len(named) <= od
        buf = oa.pack_patch_elements(named, device=True)
        cnt = buf.numel() // oa.element_record_bytes()
        data, _, gids = w2.unpack_elements(buf, cnt, capi.BVH_ELEM_POLY, 3)
        ghost = w2.create_collision_object()
        ghost.set_entity_data_poly(data, global_ids=gids)
        ghost.init_broadphase()
        ob2 = gpu_object(w2, "poly", b, device=True)
        ghost.broadphase(ob2)
        recs = ghost.results()
        got = sort_pairs(np.stack([recs[:, 1], recs[:, 3]], axis=1))
    assert len(want) > 0
    assert np.array_equal(got, want)


@pytest.mark.parametrize("device", [False, True])
def test_result_message_round_trip(torch_cuda, device):
    a, b = W.wavy_sheets(48)
    with CollisionWorld(8) as w:
        oa = gpu_object(w, "poly", a, device=True, gids=_gids(a.shape[2], 81))
        ob = gpu_object(w, "poly", b, device=True, gids=_gids(b.shape[2], 82))
        oa.broadphase(ob)
        recs = oa.results()
        msg = oa.pack_results(device=device)
        back = w.unpack_results(msg)
        host = msg.cpu().numpy() if device else msg
        hdr = host[:16].view(np.uint64)
        assert hdr[0] == 32 and hdr[1] == 32 * len(recs)
        assert np.array_equal(host[16:].view(np.uint64).reshape(-1, 4), recs)
        assert np.array_equal(back, recs)
        assert len(recs) > 0
        bad = host.copy()
        bad[:8] = np.frombuffer(np.uint64(24).tobytes(), np.uint8)  # stride 24
        with pytest.raises(capi.BvhError):
            w.unpack_results(bad)
        with pytest.raises(capi.BvhError):
            w.unpack_results(host[:-8])  # truncated
