"""The serving loop on the GPU (distbvh_amd/serve.py): requests served by
several slots at once (each a host thread with its own world and stream, so
their work overlaps on the device) give every request the same pairs as the
same requests served one at a time, and those pairs are the FP64 oracle's
(brute-force 26-DOP overlap of A against the moved B)."""
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


def test_serving_alone_and_in_the_mix(torch_cuda):
    alone = S.serve(slots=1, **KW)
    mix = S.serve(slots=3, **KW)
    assert alone["requests"] == mix["requests"] == 90
    assert mix["slots"] == 3
    assert alone["per_request"] == mix["per_request"]
    assert mix["refits"] + mix["rebuilds"] == mix["requests"] + KW["k"]  # + the warm-up builds
    # the FP64 oracle for the first requests
    pool = S.mesh_pool(KW["k"], KW["lo"], KW["hi"], KW["seed"])
    nreq = mix["requests"]
    picks = np.minimum((W.uniform(KW["seed"] ^ 0x2, nreq) * KW["k"]).astype(np.int64), KW["k"] - 1)
    for r in range(8):
        m = int(picks[r])
        a, b = pool[m]
        mb = S.moved(b, KW["seed"] * 1000 + 97 * m + r % KW["variants"])
        ka, _ = O.elements_poly(a)
        kb, _ = O.elements_poly(mb)
        exp = O.brute_pairs(ka, kb)
        assert mix["per_request"][r] == (len(exp), S._pair_digest(exp)), r


def test_serving_latency_report(torch_cuda):
    out = S.serve(slots=2, rate=500.0, duration=0.04, k=4, lo=2e3, hi=2e4, variants=2, seed=0x52)
    assert out["requests"] == 20 and out["slots"] == 2
    assert 0.0 <= out["p50_ms"] <= out["p99_ms"] <= out["max_ms"]
    assert out["pairs"] > 0 and out["pairs_per_s"] > 0


def test_serving_under_a_small_budget(torch_cuda):
    """A budget that holds about one mesh pair per slot: pairs are evicted
    (their device memory released, the objects reused) and rebuilt on their
    next request, and every request's pairs stay those of the run alone."""
    alone = S.serve(slots=1, **KW)
    tight = S.serve(slots=2, budget_bytes=2 * 3 << 20, **KW)
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
