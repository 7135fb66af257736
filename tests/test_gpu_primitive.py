"""The primitive kernels on the GPU (distbvh_amd/csrc/primitive.hip) against
the oracle (oracle/bvh_oracle.c, pinned by tests/test_oracle_primitive.py):
bvh_closest_point bit for bit, and the exact triangle-triangle and
sphere-sphere tests (bvh_*_pairs_intersect, bvh_result_filter_exact)
decision for decision, from host and device memory."""
import numpy as np
import pytest

import oracle as O
from distbvh_amd import CollisionWorld
from distbvh_amd import workloads as W
from helpers import FORMS, gpu_object, sort_pairs, to_cuda

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU test needs a GPU"
    return torch


def test_closest_point_known_answer(torch_cuda):
    tri = np.array([[0.0, 1.0, 0.0], [-1.0, 0.0, 0.0], [1.0, 0.0, 0.0]])
    pts = np.array([[0.0, 2.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 0.0], [0.0, -1.0, 0.0]])
    want = np.array([[0.0, 1.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 0.0], [0.0, 0.0, 0.0]])
    with CollisionWorld(1) as w:
        assert np.array_equal(w.closest_point(pts, tri), want)


@pytest.mark.parametrize("seed,n", [(1, 1000), (2, 100003), (3, 1 << 22)])
def test_closest_point_matches_oracle(torch_cuda, seed, n):
    rng = np.random.default_rng(seed)
    tri = rng.normal(size=(3, 3))
    pts = rng.normal(size=(n, 3)) * 3.0
    # every region: the vertices, edge midpoints and points on the edges' lines
    pts[:6] = np.concatenate([tri, (tri + np.roll(tri, 1, axis=0)) / 2])
    want = O.closest_point(pts, tri)
    with CollisionWorld(1) as w:
        assert np.array_equal(w.closest_point(pts, tri), want)
        got = w.closest_point(torch_cuda.from_numpy(pts).cuda(), tri).cpu().numpy()
        assert np.array_equal(got, want)


def test_closest_point_degenerate(torch_cuda):
    """A segment-like triangle and repeated vertices: same values (and NaNs)
    as the oracle."""
    rng = np.random.default_rng(9)
    pts = rng.normal(size=(5000, 3))
    for tri in (np.array([[0.0, 0, 0], [1, 0, 0], [2, 0, 0]]), np.array([[1.0, 1, 1], [1, 1, 1], [0, 1, 0]])):
        want = O.closest_point(pts, tri)
        with CollisionWorld(1) as w:
            got = w.closest_point(pts, tri)
        assert np.array_equal(np.isnan(got), np.isnan(want))
        assert np.array_equal(np.nan_to_num(got), np.nan_to_num(want))


# ---- exact triangle-triangle / sphere-sphere tests -----------------------------------------
def _grid_triangles(n, seed):
    """Integer triangles on a 7^3 grid, with shared vertices, shared edges
    and coplanar pairs (the oracle test's generator)."""
    rng = np.random.default_rng(seed)
    P = rng.integers(-3, 4, size=(n, 3, 3))
    Q = rng.integers(-3, 4, size=(n, 3, 3))
    Q[: n // 4, 0] = P[: n // 4, 0]
    Q[n // 4: n // 2, :2] = P[n // 4: n // 2, :2]
    P[n // 2: 5 * n // 8, :, 2] = 0
    Q[n // 2: 5 * n // 8, :, 2] = 0
    return P.reshape(n, 9).astype(np.float64), Q.reshape(n, 9).astype(np.float64)


def _float_triangles(n, seed):
    rng = np.random.default_rng(seed)
    c = rng.normal(size=(n, 1, 3)) * 1e3
    P = c + rng.normal(size=(n, 3, 3))
    Q = c + 0.8 * rng.normal(size=(n, 3, 3))
    return P.reshape(n, 9), Q.reshape(n, 9)


@pytest.mark.parametrize("device", [False, True])
def test_triangle_pairs_match_oracle(torch_cuda, device):
    with CollisionWorld(1) as w:
        for P, Q in (_grid_triangles(100000, 21), _float_triangles(100000, 22)):
            want = O.tri_tri(P, Q)
            assert 0.1 < want.mean() < 0.9
            if device:
                got = w.triangle_pairs_intersect(to_cuda(P), to_cuda(Q)).cpu().numpy()
            else:
                got = w.triangle_pairs_intersect(P, Q)
            assert np.array_equal(got, want)


@pytest.mark.parametrize("device", [False, True])
def test_sphere_pairs_match_oracle(torch_cuda, device):
    rng = np.random.default_rng(23)
    A = rng.integers(-24, 25, size=(50000, 4)).astype(np.float64) / 4.0
    B = rng.integers(-24, 25, size=(50000, 4)).astype(np.float64) / 4.0
    A[:, 3] = np.abs(A[:, 3])
    B[:, 3] = np.abs(B[:, 3])
    want = O.sphere_sphere(A, B)
    with CollisionWorld(1) as w:
        got = (w.sphere_pairs_intersect(to_cuda(A), to_cuda(B)).cpu().numpy() if device
               else w.sphere_pairs_intersect(A, B))
    assert np.array_equal(got, want)


def _tris_of(planes, idx):
    """(k, 9) triangles of SoA vertex planes (3, 3, n) at input indices idx."""
    return np.ascontiguousarray(planes[:, :, idx].transpose(2, 0, 1).reshape(-1, 9))


@pytest.mark.parametrize("precision", FORMS)
@pytest.mark.parametrize("sort", [False, True])
def test_filter_exact_contact_sheets(torch_cuda, precision, sort):
    """k-DOP pairs of two wavy sheets, then the primitive narrowphase: the kept
    pairs are exactly the oracle's k-DOP pairs whose triangles meet, in
    their order."""
    a, b = W.wavy_sheets(96)
    ka, _ = O.elements_poly(a)
    kb, _ = O.elements_poly(b)
    cand = sort_pairs(O.brute_pairs(ka, kb))
    keep = O.tri_tri(_tris_of(a, cand[:, 0].astype(np.int64)), _tris_of(b, cand[:, 1].astype(np.int64)))
    want = cand[keep.astype(bool)]
    assert 0 < len(want) < len(cand)
    with CollisionWorld(8, precision=precision, sort_results=sort) as w:
        oa = gpu_object(w, "poly", a, device=True)
        ob = gpu_object(w, "poly", b, device=True)
        oa.broadphase(ob)
        assert oa.result_count() == len(cand)
        kept = oa.filter_exact()
        got = oa.pairs()
        recs = oa.results()
    assert kept == len(want) == len(got) == len(recs)
    if sort:
        assert np.array_equal(got, want)
    assert np.array_equal(sort_pairs(got), want)
    assert np.array_equal(np.sort(recs[:, 1]), np.sort(want[:, 0]))


def test_filter_exact_mesh_self_and_spheres(torch_cuda):
    """Self collision of one sheet keeps every pair of triangles that share a
    vertex or an edge (closed triangles touch); sphere self collision keeps
    the pairs whose balls meet."""
    a, _ = W.wavy_sheets(64)
    ka, _ = O.elements_poly(a)
    cand = sort_pairs(O.brute_self(ka))
    keep = O.tri_tri(_tris_of(a, cand[:, 0].astype(np.int64)), _tris_of(a, cand[:, 1].astype(np.int64)))
    want = cand[keep.astype(bool)]
    s = W.sphere_cloud(20000, radius=5e-3, ncomp=8, sigma=0.05, shift=(3.0, -2.0, 1.0), seed=24)
    ks, _ = O.elements_sphere(s)
    scand = sort_pairs(O.brute_self(ks))
    sph = np.ascontiguousarray(s.T)
    skeep = O.sphere_sphere(sph[scand[:, 0].astype(np.int64)], sph[scand[:, 1].astype(np.int64)])
    swant = scand[skeep.astype(bool)]
    assert 0 < len(swant) < len(scand)
    with CollisionWorld(8) as w:
        o = gpu_object(w, "poly", a, device=True)
        o.self_collision()
        o.filter_exact()
        assert np.array_equal(sort_pairs(o.pairs()), want)
        os_ = gpu_object(w, "sphere", s, device=True)
        os_.self_collision()
        os_.filter_exact()
        assert np.array_equal(sort_pairs(os_.pairs()), swant)
