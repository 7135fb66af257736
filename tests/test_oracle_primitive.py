"""The oracle's closest point on a triangle (oracle/bvh_oracle.c
orc_closest_point) against the reference's known answer
(tests/PrimitiveTest.cpp:66-87) and against what the geometry fixes: the
result lies on the triangle and no point of the triangle is closer."""
import numpy as np

import oracle as O

TRI = np.array([[0.0, 1.0, 0.0], [-1.0, 0.0, 0.0], [1.0, 0.0, 0.0]])


def test_closest_point_known_answer():
    pts = np.array([[0.0, 2.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 0.0], [0.0, -1.0, 0.0]])
    want = np.array([[0.0, 1.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 0.0], [0.0, 0.0, 0.0]])
    assert np.array_equal(O.closest_point(pts, TRI), want)


def _barycentric(q, a, b, c):
    m = np.stack([b - a, c - a], axis=1)
    uv, *_ = np.linalg.lstsq(m, q - a, rcond=None)
    return 1.0 - uv.sum(), uv[0], uv[1]


def test_closest_point_is_the_minimum():
    rng = np.random.default_rng(3)
    for _ in range(20):
        tri = rng.normal(size=(3, 3))
        a, b, c = tri
        pts = rng.normal(size=(200, 3)) * 2.0
        got = O.closest_point(pts, tri)
        # a dense sample of the triangle
        s = rng.uniform(size=(4000, 2))
        s = np.where(s.sum(1, keepdims=True) > 1, 1 - s, s)
        sample = a + s[:, :1] * (b - a) + s[:, 1:] * (c - a)
        for p, q in zip(pts, got):
            w = _barycentric(q, a, b, c)
            assert min(w) > -1e-9 and abs(sum(w) - 1) < 1e-9  # on the triangle
            d = np.linalg.norm(p - q)
            assert d <= np.min(np.linalg.norm(sample - p, axis=1)) + 1e-9


def test_closest_point_regions():
    """Points in the plane inside the triangle map to themselves; points
    beyond a vertex map to that vertex."""
    a, b, c = TRI
    inside = np.array([a * 0.2 + b * 0.3 + c * 0.5, a * 0.6 + b * 0.2 + c * 0.2])
    assert np.allclose(O.closest_point(inside, TRI), inside, atol=1e-15)
    beyond = np.array([a + (a - (b + c) / 2) * 3, b + (b - (a + c) / 2) * 3, c + (c - (a + b) / 2) * 3])
    assert np.array_equal(O.closest_point(beyond, TRI), TRI)
