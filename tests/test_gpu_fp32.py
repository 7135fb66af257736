"""GPU parity of the fp32-outward precision form: every result must be the
fp64 form's (and the oracle's) exactly, because the fp32 tree only prunes and
every undecided candidate is re-tested in fp64."""
import numpy as np
import pytest

import oracle as O
from distbvh_amd import CollisionWorld
from distbvh_amd import workloads as W
from helpers import gpu_object, sort_pairs, sort_records

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU test needs a GPU"
    return torch


def _run(kind, a, b, od, leaf, precision, eps=0.0, bits=63):
    with CollisionWorld(od, morton_bits=bits, leaf_size=leaf, precision=precision) as w:
        oa = gpu_object(w, kind, a, eps, device=True)
        ob = gpu_object(w, kind, b, eps, device=True)
        oa.broadphase(ob)
        recs = sort_records(oa.results())
        pp = oa.patch_pairs()
        stats = oa.result_stats()
        patches = oa.local_patches()
    return recs, pp, stats, patches


@pytest.mark.parametrize("nq,od,leaf", [(64, 16, 1), (128, 32, 4), (128, 1, 2)])
def test_fp32_wavy_sheets_match_oracle(torch_cuda, nq, od, leaf):
    a, b = W.wavy_sheets(nq)
    ka, ca = O.elements_poly(a)
    kb, cb = O.elements_poly(b)
    ra = O.Object(ka, ca, od=od, morton_bits=63, leaf_size=leaf)
    rb = O.Object(kb, cb, od=od, morton_bits=63, leaf_size=leaf)
    ref = sort_records(ra.collide(rb))
    recs, pp, stats, patches = _run("poly", a, b, od, leaf, "fp32")
    recs64, pp64, stats64, patches64 = _run("poly", a, b, od, leaf, "fp64")
    assert len(ref) > 0
    assert np.array_equal(recs, ref)
    assert np.array_equal(pp, pp64)
    # patch k-DOPs in the fp32 form are the exact fp64 unions
    for p32, p64, k in zip(patches, patches64, range(od)):
        assert p32["size"] == p64["size"]
        assert np.array_equal(p32["kdop"], p64["kdop"]), k
        assert np.array_equal(p32["kdop"], ra.patch_kdops[k]) or p32["size"] == 0
    assert stats["candidate_pairs"] >= len(recs)
    assert stats["rechecked_pairs"] <= stats["candidate_pairs"]
    # outward rounding moves a slab by at most 1.5 fp32 ulps
    assert 0.0 < stats["max_slab_error_ulps"] <= 2.0


@pytest.mark.parametrize("shift", [(0.0, 0.0, 0.0), (1e4, -1e4, 1e4), (3e6, 3e6, -3e6)])
def test_fp32_spheres_far_from_origin(torch_cuda, shift):
    a = W.sphere_cloud(20000, radius=5e-3, ncomp=8, sigma=0.05, shift=shift, seed=21)
    b = W.sphere_cloud(20000, radius=5e-3, ncomp=8, sigma=0.05, shift=shift, seed=22)
    ka, _ = O.elements_sphere(a)
    kb, _ = O.elements_sphere(b)
    brute = sort_pairs(O.brute_pairs(ka, kb))
    recs, pp, stats, _ = _run("sphere", a, b, 8, 1, "fp32")
    assert len(brute) > 0
    assert np.array_equal(sort_pairs(recs[:, [1, 3]]), brute)


@pytest.mark.parametrize("leaf", [1, 4])
def test_fp32_recheck_decides_touching_pairs(torch_cuda, leaf):
    """Spheres within +-4 fp64 ulps of touching: fp32 cannot separate them,
    so they go through the fp64 re-check, which must reproduce the fp64
    answer pair by pair."""
    a, b = W.touching_spheres(4096, max_ulps=4)
    ka, _ = O.elements_sphere(a)
    kb, _ = O.elements_sphere(b)
    brute = sort_pairs(O.brute_pairs(ka, kb))
    recs, _, stats, _ = _run("sphere", a, b, 4, leaf, "fp32")
    recs64, _, _, _ = _run("sphere", a, b, 4, leaf, "fp64")
    got = sort_pairs(recs[:, [1, 3]])
    assert np.array_equal(got, brute)
    assert np.array_equal(recs, recs64)
    # the negative gaps overlap, the others do not: both kinds are present
    same = brute[brute[:, 0] == brute[:, 1]]
    assert 0 < len(same) < 4096
    assert stats["rechecked_pairs"] >= 4096 - len(same)


def test_fp32_self_collision(torch_cuda):
    s = W.sphere_cloud(30000, radius=5e-3, ncomp=8, sigma=0.05, shift=(2e3, 0.0, -1e3), seed=23)
    kd, _ = O.elements_sphere(s)
    brute = sort_pairs(O.brute_self(kd))
    with CollisionWorld(8, leaf_size=2, precision="fp32") as w:
        o = gpu_object(w, "sphere", s, device=True)
        o.self_collision()
        got = sort_pairs(o.pairs())
    assert len(brute) > 0
    assert np.array_equal(got, brute)


def test_fp32_collision_known_answers(torch_cuda):
    # tests/CollisionTest.cpp:43-81 in the fp32 form
    g = W.hex_grid(4, 4, 4)
    cube = W.hex_grid(1, 1, 1)
    with CollisionWorld(1, precision="fp32") as w:
        oa = gpu_object(w, "poly", g)
        ob = gpu_object(w, "poly", cube)
        oa.broadphase(ob)
        assert oa.result_count() == 64
    with CollisionWorld(1, precision="fp32") as w:
        o = gpu_object(w, "poly", W.hex_grid(2, 2, 1))
        o.self_collision()
        assert o.result_count() == 0
    with CollisionWorld(1, precision="fp32") as w:
        both = np.concatenate([W.hex_grid(2, 2, 2), cube], axis=2)
        o = gpu_object(w, "poly", both)
        o.self_collision()
        assert o.result_count() == 8


def test_fp32_full_size_matches_fp64(torch_cuda):
    """The benchmark workload at full size (2 x 2,097,152 triangles): the fp32
    form's records are the fp64 form's."""
    a, b = W.wavy_sheets(1024)
    recs, pp, stats, _ = _run("poly", a, b, 256, 4, "fp32")
    recs64, pp64, _, _ = _run("poly", a, b, 256, 4, "fp64")
    assert len(recs64) > 0
    assert np.array_equal(recs, recs64)
    assert np.array_equal(pp, pp64)
    assert stats["candidate_pairs"] >= len(recs)


def test_fp32_mesh_self_collision_recheck_stripes(torch_cuda):
    """Mesh self collision re-checks every pair of triangles that share a
    vertex (about six per triangle), more than the re-check list holds at
    first (max(65536, n) entries over 64 stripes): the query overflows, grows
    every stripe to the fullest one's count and runs again.  Pairs must equal
    the oracle's brute force on both runs and the fp64 form's."""
    a, _ = W.wavy_sheets(128)
    kd, _ = O.elements_poly(a)
    brute = sort_pairs(O.brute_self(kd))
    with CollisionWorld(16, leaf_size=1, precision="fp32") as w:
        o = gpu_object(w, "poly", a, device=True)
        o.self_collision()
        first = sort_pairs(o.pairs())
        stats = o.result_stats()
        o.self_collision()  # buffers already large enough
        second = sort_pairs(o.pairs())
    with CollisionWorld(16, leaf_size=1, precision="fp64") as w:
        o = gpu_object(w, "poly", a, device=True)
        o.self_collision()
        got64 = sort_pairs(o.pairs())
    assert len(brute) > 0
    assert stats["rechecked_pairs"] > max(65536, kd.shape[0])
    assert np.array_equal(first, brute)
    assert np.array_equal(second, brute)
    assert np.array_equal(got64, brute)


# ---- the precision workload (16,777,216 spheres, r = 1e-3, 64-component mixture) ------------
def test_precision_workload_generator_against_brute_force(torch_cuda):
    """The precision workload's generator exactly as configured (64
    components, sigma 0.01, r 1e-3, cloud translated by (1e4, -1e4, 1e4)),
    at 32,768 spheres: self collision in the fp32-outward and fp64 forms,
    both equal to the oracle's brute-force pairs."""
    s = W.sphere_cloud(32768)  # defaults: ncomp=64, sigma=0.01, radius=1e-3, shift=(1e4, -1e4, 1e4)
    kd, _ = O.elements_sphere(s)
    brute = sort_pairs(O.brute_self(kd))
    assert len(brute) > 0
    for prec in ("fp32", "fp64"):
        with CollisionWorld(64, precision=prec) as w:
            o = gpu_object(w, "sphere", s, device=True)
            o.self_collision()
            got = sort_pairs(o.pairs())
            stats = o.result_stats()
        assert np.array_equal(got, brute), prec
        if prec == "fp32":
            assert stats["candidate_pairs"] >= len(brute)
            assert 0.0 < stats["max_slab_error_ulps"] <= 2.0


def test_precision_workload_full_size(torch_cuda):
    """16,777,216 spheres, self collision: the fp32-outward form's pair count
    and digest equal the fp64 form's, the largest fp32 slab error is at most
    2 ulps and every fp64 pair was a candidate."""
    import torch
    s = W.sphere_cloud()
    ts = torch.from_numpy(s).cuda()
    del s
    res = {}
    for prec in ("fp64", "fp32"):
        with CollisionWorld(64, precision=prec) as w:
            o = w.create_collision_object()
            o.set_entity_data_sphere(ts)
            o.init_broadphase()
            o.self_collision()
            res[prec] = (o.result_count(), o.digest(), o.result_stats())
        torch.cuda.empty_cache()
    n64, d64, _ = res["fp64"]
    n32, d32, st = res["fp32"]
    print("precision workload: %d pairs, fp32 candidates %d, re-checked %d, excess false-positive rate %.3g, "
          "max slab error %.3f ulps" % (n64, st["candidate_pairs"], st["rechecked_pairs"],
                                        (st["candidate_pairs"] - n64) / max(n64, 1), st["max_slab_error_ulps"]))
    assert n64 > 10 ** 9
    assert (n32, d32) == (n64, d64)
    assert st["candidate_pairs"] >= n64
    assert 0.0 < st["max_slab_error_ulps"] <= 2.0


def test_batch_workload_full_size_fp32_matches_fp64(torch_cuda):
    """The 4,096-mesh batch (lognormal sizes, ~19M triangles) in one object,
    self collision: the fp32-outward form's pairs (count and digest) are the
    fp64 form's."""
    import torch
    a, counts = W.patch_batch()
    ta = torch.from_numpy(a).cuda()
    del a
    res = {}
    for prec in ("fp64", "fp32"):
        with CollisionWorld(4096, precision=prec) as w:
            o = w.create_collision_object()
            o.set_entity_data_poly(ta)
            o.init_broadphase()
            o.self_collision()
            res[prec] = (o.result_count(), o.digest())
        torch.cuda.empty_cache()
    assert res["fp64"][0] > 0
    assert res["fp32"] == res["fp64"]


# ---- object-local frames ---------------------------------------------------------------------
@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_object_frames_far_apart(torch_cuda, precision):
    """Each object's fp32 slabs are relative to its own origin.  An object
    built first at the origin must not coarsen the slabs of two objects
    built later 1e4 away: their pairs (A x B across frames, and A's self
    collision) equal the oracle's brute force, and the fp32 candidates stay
    within a few percent of the pairs (one world frame at the first object
    would round those slabs to fp32 ulps of about 1e-3, the spheres'
    radius)."""
    c = W.sphere_cloud(4000, radius=1e-3, ncomp=4, sigma=0.01, shift=(0.0, 0.0, 0.0), seed=0x61)
    a = W.sphere_cloud(12000, radius=1e-3, ncomp=8, sigma=0.01, shift=(1e4, -1e4, 1e4), seed=0x62)
    b = W.sphere_cloud(12000, radius=1e-3, ncomp=8, sigma=0.01, shift=(1e4 + 0.0013, -1e4 + 0.0007, 1e4), seed=0x62)
    ka, _ = O.elements_sphere(a)
    kb, _ = O.elements_sphere(b)
    want_ab = sort_pairs(O.brute_pairs(ka, kb))
    want_aa = sort_pairs(O.brute_self(ka))
    assert len(want_ab) > 0 and len(want_aa) > 0
    with CollisionWorld(8, precision=precision, serial_builds=True) as w:
        oc = gpu_object(w, "sphere", c, device=True)  # built first, at the origin
        oa = gpu_object(w, "sphere", a, device=True)
        ob = gpu_object(w, "sphere", b, device=True)
        oa.broadphase(ob)
        got_ab = sort_pairs(oa.pairs())
        st_ab = oa.result_stats()
        oa.self_collision()
        got_aa = sort_pairs(oa.pairs())
        st_aa = oa.result_stats()
        oc.self_collision()
    assert np.array_equal(got_ab, want_ab)
    assert np.array_equal(got_aa, want_aa)
    for st, want in ((st_ab, want_ab), (st_aa, want_aa)):
        assert len(want) <= st["candidate_pairs"] <= 1.05 * len(want) + 16
        assert 0.0 < st["max_slab_error_ulps"] <= 2.0
