"""bvh_closest_point on the GPU (distbvh_amd/csrc/primitive.hip) against the
oracle (oracle/bvh_oracle.c orc_closest_point, pinned by
tests/test_oracle_primitive.py), bit for bit, from host and device memory."""
import numpy as np
import pytest

import oracle as O
from distbvh_amd import CollisionWorld

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
