"""The serving loop on the GPU (distbvh_amd/serve.py): requests coalesced into
shared steps (one narrowphase launch for every waiting request, one world and
one device-memory budget for all of them) give every request the same pairs
as the same requests served one at a time, and those pairs are the FP64
oracle's: brute-force 26-DOP overlap of A against the moved B for small
meshes, the restated reference pipeline for a mesh of more than a million
triangles."""
import numpy as np
import pytest

import oracle as O
from distbvh_amd import serve as S
from distbvh_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU test needs a GPU"
    return torch


KW = dict(rate=3000.0, duration=0.03, k=6, lo=2e3, hi=1e4, variants=3, seed=0x51, record=True)


def _expected(pool, seed, variants, m, v, brute=True):
    a, b = pool[m]
    mb = S.moved(b, seed * 1000 + 97 * m + v)
    ka, ca = O.elements_poly(a)
    kb, cb = O.elements_poly(mb)
    if brute:
        exp = O.brute_pairs(ka, kb)
    else:
        recs = O.Object(ka, ca, od=64, nthreads=16).collide(O.Object(kb, cb, od=64, nthreads=16))
        exp = recs[:, [1, 3]]
    return len(exp), S.pair_digest(exp)


def test_serving_alone_and_in_the_mix(torch_cuda):
    alone = S.serve(max_batch=1, **KW)
    mix = S.serve(max_batch=16, **KW)
    assert alone["requests"] == mix["requests"] == 90
    assert alone["max_step_requests"] == 1
    assert mix["max_step_requests"] > 1  # requests were coalesced
    assert alone["per_request"] == mix["per_request"]
    assert mix["refits"] + mix["rebuilds"] == mix["requests"]
    pool = S.mesh_pool(KW["k"], KW["lo"], KW["hi"], KW["seed"])
    for r in range(12):
        m = mix["picks"][r]
        assert mix["per_request"][r] == _expected(pool, KW["seed"], KW["variants"], m, r % KW["variants"]), r


def test_serving_production_sizes_against_the_oracle(torch_cuda):
    """Mesh pairs of 10k, 100k and 1.2M triangles per object, coalesced: every
    request's pairs equal the alone run's, and each distinct request (mesh,
    variant) of the stream equals the FP64 oracle's: brute force for the
    smallest mesh, the restated reference pipeline for the others."""
    kw = dict(rate=400.0, duration=0.06, variants=2, seed=0x53, record=True, sizes=[1e4, 1e5, 1.2e6])
    alone = S.serve(max_batch=1, **kw)
    mix = S.serve(max_batch=16, **kw)
    assert mix["mesh_triangles"][2] >= 1_000_000
    assert alone["per_request"] == mix["per_request"]
    pool = S.mesh_pool(3, 0, 0, kw["seed"], sizes=kw["sizes"])
    seen = {}
    for r, m in enumerate(mix["picks"]):
        key = (m, r % kw["variants"])
        if key not in seen:
            seen[key] = _expected(pool, kw["seed"], kw["variants"], m, key[1], brute=m == 0)
        assert mix["per_request"][r] == seen[key], (r, key)
    assert any(m == 2 for m, _ in seen)


def test_serving_latency_report(torch_cuda):
    out = S.serve(max_batch=8, rate=500.0, duration=0.04, k=4, lo=2e3, hi=2e4, variants=2, seed=0x52)
    assert out["requests"] == 20
    assert 0.0 <= out["p50_ms"] <= out["p99_ms"] <= out["max_ms"]
    assert out["pairs"] > 0 and out["pairs_per_s"] > 0
    assert out["steps"] <= out["requests"]


def test_serving_under_a_small_budget(torch_cuda):
    """A budget that holds about two mesh pairs: pairs are evicted (their
    device memory released, the objects reused) and rebuilt on their next
    request, requests wait for room, and every request's pairs stay those of
    the run alone with an ample budget."""
    alone = S.serve(max_batch=1, **KW)
    tight = S.serve(max_batch=16, budget_bytes=2 * 3 << 20, **KW)
    assert tight["evictions"] > 0
    assert tight["per_request"] == alone["per_request"]


def test_object_release(torch_cuda):
    """bvh_object_release frees the object's memory, invalidates results
    against it, and the object builds again on its next set_entity_data."""
    from distbvh_amd import CollisionWorld, capi

    a, b = W.wavy_sheets(32)
    with CollisionWorld(8) as w:
        oa = w.create_collision_object()
        ob = w.create_collision_object()
        oa.set_entity_data_poly(a)
        ob.set_entity_data_poly(b)
        oa.init_broadphase()
        ob.init_broadphase()
        oa.broadphase(ob)
        ref = (oa.result_count(), oa.digest())
        assert ref[0] > 0 and ob.device_bytes() > 0
        ob.release()
        assert ob.device_bytes() == 0
        assert oa.result_count() == 0
        with pytest.raises(capi.BvhError):
            oa.broadphase(ob)
        ob.set_entity_data_poly(b)
        ob.init_broadphase()
        oa.broadphase(ob)
        assert (oa.result_count(), oa.digest()) == ref
