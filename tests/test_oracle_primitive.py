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


# ---- triangle-triangle and sphere-sphere (no reference counterpart) ----------------------
# The exact reference below is an independent method: the separating-axis
# test over the 17 candidate axes of two triangles (both normals, the nine
# edge-edge cross products and the six in-plane edge normals), evaluated in
# exact integer arithmetic on integer (or integer-scaled dyadic) coordinates.
# Two closed triangles are disjoint iff some axis strictly separates their
# projections.
def _cross(a, b):
    return np.stack([a[..., 1] * b[..., 2] - a[..., 2] * b[..., 1],
                     a[..., 2] * b[..., 0] - a[..., 0] * b[..., 2],
                     a[..., 0] * b[..., 1] - a[..., 1] * b[..., 0]], axis=-1)


def _sat_intersect(P, Q):
    """P, Q: (n, 3, 3) integer arrays (int64 or object of Python ints).
    Exact separating-axis test for closed non-degenerate triangles."""
    eP = np.stack([P[:, 1] - P[:, 0], P[:, 2] - P[:, 1], P[:, 0] - P[:, 2]], axis=1)
    eQ = np.stack([Q[:, 1] - Q[:, 0], Q[:, 2] - Q[:, 1], Q[:, 0] - Q[:, 2]], axis=1)
    nP = _cross(eP[:, 0], P[:, 2] - P[:, 0])
    nQ = _cross(eQ[:, 0], Q[:, 2] - Q[:, 0])
    axes = [nP, nQ]
    for i in range(3):
        for j in range(3):
            axes.append(_cross(eP[:, i], eQ[:, j]))
    for i in range(3):
        axes.append(_cross(nP, eP[:, i]))
        axes.append(_cross(nQ, eQ[:, i]))
    sep = np.zeros(len(P), bool)
    for ax in axes:
        pp = np.einsum("nvc,nc->nv", P, ax)
        qq = np.einsum("nvc,nc->nv", Q, ax)
        sep |= (pp.max(axis=1) < qq.min(axis=1)) | (qq.max(axis=1) < pp.min(axis=1))
    return ~sep


def _nondegenerate(T):
    n = _cross(T[:, 1] - T[:, 0], T[:, 2] - T[:, 0])
    return np.any(n != 0, axis=1)


def test_tri_tri_closed_forms():
    base = np.array([[0, 0, 0], [4, 0, 0], [0, 4, 0]], np.float64)
    cases = [
        (base, base, 1),                                              # identical
        (base, base + [10, 0, 0], 0),                                 # apart, coplanar
        (base, base + [0, 0, 1], 0),                                  # parallel planes
        (base, np.array([[4, 0, 0], [8, 0, 0], [4, 4, 0]]), 1),       # coplanar, shared vertex
        (base, np.array([[4, 0, 0], [0, 4, 0], [4, 4, 0]]), 1),       # coplanar, shared edge
        (base, np.array([[1, 1, 0], [2, 1, 0], [1, 2, 0]]), 1),       # coplanar, one inside the other
        (base, np.array([[3, 3, 0], [6, 3, 0], [3, 6, 0]]), 0),       # coplanar, corner just outside
        (base, np.array([[2, 2, 0], [6, 2, 0], [2, 6, 0]]), 1),       # coplanar, touching at the hypotenuse
        (base, np.array([[1, 1, -1], [1, 1, 1], [5, 5, 1]]), 1),      # an edge pierces the interior
        (base, np.array([[1, 1, 0], [2, 1, 0], [1, 1, 3]]), 1),       # an edge lies in the face (T shape)
        (base, np.array([[0, 0, 0], [-1, 0, 1], [0, -1, 1]]), 1),     # shared vertex, off-plane
        (base, np.array([[1, 1, 1], [2, 1, 1], [1, 2, 3]]), 0),       # above the face
        (base, np.array([[1, 1, 0.5], [2, 1, 0.5], [1, 1, 3]]), 0),   # vertex just above the face
        (base, np.array([[5, 5, -1], [5, 5, 1], [9, 9, 0]]), 0),      # crosses the plane outside
        (base, np.array([[2, 2, -1], [2, 2, 1], [9, 9, 0]]), 1),      # crosses the plane on the hypotenuse
        (base, np.array([[0, 0, 1], [4, 0, 1], [2, 0, 1]]), 0),       # degenerate (collinear) triangle
    ]
    for k, (p, q, want) in enumerate(cases):
        assert O.tri_tri(p, q)[0] == want, k
        assert O.tri_tri(q, p)[0] == want, k


def test_tri_tri_integer_grid_against_exact_separating_axes():
    """Small integer coordinates: every orientation determinant of the rule
    is exact in fp64, so the oracle must give the exact answer, including
    every degenerate contact (shared vertices and edges, coplanar overlaps,
    vertices on faces) that a 7^3 grid produces."""
    rng = np.random.default_rng(11)
    n = 40000
    P = rng.integers(-3, 4, size=(n, 3, 3))
    Q = rng.integers(-3, 4, size=(n, 3, 3))
    # many pairs sharing vertices or lying in one coordinate plane
    Q[: n // 4, 0] = P[: n // 4, 0]
    Q[n // 4: n // 2, :2] = P[n // 4: n // 2, :2]
    P[n // 2: 5 * n // 8, :, 2] = 0
    Q[n // 2: 5 * n // 8, :, 2] = 0
    got = O.tri_tri(P.astype(np.float64), Q.astype(np.float64)).astype(bool)
    ok = _nondegenerate(P) & _nondegenerate(Q)
    assert not np.any(got[~ok])  # the declared rule: degenerate triangles meet nothing
    want = _sat_intersect(P[ok], Q[ok])
    assert 0.2 < want.mean() < 0.9
    assert np.array_equal(got[ok], want)
    # symmetric
    assert np.array_equal(O.tri_tri(Q.astype(np.float64), P.astype(np.float64)).astype(bool), got)


def test_tri_tri_dyadic_random_against_exact_separating_axes():
    """Random coordinates with 20 fractional bits (exactly representable;
    the determinants now round in fp64): away from exact ties the fp64 rule
    must agree with the exact answer on every pair."""
    rng = np.random.default_rng(12)
    n = 3000
    scale = 1 << 20
    c = rng.normal(size=(n, 1, 3))
    Pi = np.rint((c + rng.normal(size=(n, 3, 3))) * scale).astype(np.int64)
    Qi = np.rint((c + 0.8 * rng.normal(size=(n, 3, 3))) * scale).astype(np.int64)
    got = O.tri_tri(Pi / scale, Qi / scale).astype(bool)
    want = _sat_intersect(Pi.astype(object), Qi.astype(object))
    assert 0.2 < want.mean() < 0.9
    assert np.array_equal(got, want)


def test_sphere_sphere_closed_forms_and_exact():
    a = np.array([[0.0, 0.0, 0.0, 2.0]] * 4)
    b = np.array([[3.0, 4.0, 0.0, 3.0],      # |d| = 5 = r1 + r2: touching
                  [3.0, 4.0, 0.0, 2.75],     # apart by 0.25
                  [0.0, 0.0, 0.0, 0.5],      # concentric
                  [-3.0, 0.0, 4.0, 3.5]])    # overlapping
    assert list(O.sphere_sphere(a, b)) == [1, 0, 1, 1]
    assert list(O.sphere_sphere(b, a)) == [1, 0, 1, 1]
    rng = np.random.default_rng(13)
    n = 20000
    A = rng.integers(-24, 25, size=(n, 4))
    B = rng.integers(-24, 25, size=(n, 4))
    A[:, 3] = np.abs(A[:, 3])
    B[:, 3] = np.abs(B[:, 3])
    d = A[:, :3] - B[:, :3]
    want = (d * d).sum(axis=1) <= (A[:, 3] + B[:, 3]) ** 2
    assert 0.1 < want.mean() < 0.9
    assert np.array_equal(O.sphere_sphere(A / 4.0, B / 4.0).astype(bool), want)
