"""sort_and_sweep_local / max_variant_axis on the GPU (distbvh_amd/csrc/
sweep.hip) against the oracle (oracle/bvh_oracle.c orc_sort_and_sweep /
orc_max_variant_axis, pinned by tests/test_oracle_golden.py: the reference's
known answers, brute force and the closed-form pair order).  The pairs are
compared as arrays: the same pairs in the same order (A ascending, then B in
sorted order)."""
import numpy as np
import pytest

import oracle as O
from distbvh_amd import CollisionWorld
from distbvh_amd import capi
from distbvh_amd import workloads as W
from helpers import to_cuda

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU test needs a GPU"
    return torch


def _patch(k):
    p = np.empty(26)
    p[0::2], p[1::2] = k[:, 0::2].min(0), k[:, 1::2].max(0)
    return p


def _grid(nx, ny, nz, base=0):
    kd, ce = O.elements_poly(W.hex_grid(nx, ny, nz))
    return kd, ce, np.arange(base, base + kd.shape[0], dtype=np.uint64)


def _sheets(nq, seed=0):
    a, b = W.wavy_sheets(nq)
    ka, ca = O.elements_poly(a)
    kb, cb = O.elements_poly(b)
    rng = np.random.default_rng(seed)
    ga = rng.permutation(ka.shape[0]).astype(np.uint64) + np.uint64(1000)
    gb = rng.permutation(kb.shape[0]).astype(np.uint64) + np.uint64(5)
    return ka, ca, ga, kb, cb, gb


def test_sweep_known_answer(torch_cuda):
    """KokkosTest.cpp:87-139: 12 elements along x against a unit cube, both
    ways round: 12 pairs, the oracle's."""
    ka, ca, ga = _grid(12, 1, 1)
    kb, cb, gb = _grid(1, 1, 1)
    pa = O.kdop_from_vertices(W.hex_grid(12, 1, 1).transpose(2, 0, 1).reshape(-1, 3))
    pb = kb[0]
    with CollisionWorld(1) as w:
        got = w.sort_and_sweep(pa, ka, ga, pb, kb, gb, 0)
        assert len(got) == 12
        assert np.array_equal(got, O.sort_and_sweep(pa, ka, ga, pb, kb, gb, 0))
        got = w.sort_and_sweep(pb, kb, gb, pa, ka, ga, 0)
        assert len(got) == 12
        assert np.array_equal(got, O.sort_and_sweep(pb, kb, gb, pa, ka, ga, 0))


def test_max_variant_axis_known(torch_cuda):
    """KokkosTest.cpp:37-85 and exact integer ties (the first axis wins)."""
    _, cb, _ = _grid(1, 1, 1)
    with CollisionWorld(1) as w:
        for dims, axis in [((12, 1, 1), 0), ((1, 12, 1), 1), ((1, 1, 12), 2)]:
            _, ca, _ = _grid(*dims)
            assert w.max_variant_axis(ca, cb) == axis == O.max_variant_axis(ca, cb)
        g = np.array([(i, j, 0) for i in range(4) for j in range(4)], np.float64)
        assert w.max_variant_axis(g, g[:1]) == 0
        assert w.max_variant_axis(g[:, [2, 0, 1]], g[:1, [2, 0, 1]]) == 1
        assert w.max_variant_axis(g[:0], g) == 0


def test_max_variant_axis_random(torch_cuda):
    """Random centroid sets with distinct variances, host and device memory,
    sizes across one to many reduction blocks."""
    rng = np.random.default_rng(3)
    with CollisionWorld(1) as w:
        for na, nb in [(1, 0), (5, 7), (1000, 3), (70000, 90000)]:
            for scale in ([3.0, 1.0, 0.5], [0.2, 0.1, 2.0], [1.0, 4.0, 1.0]):
                ca = rng.normal(size=(na, 3)) * scale + 3.0
                cb = rng.normal(size=(nb, 3)) * scale - 1.0
                exp = O.max_variant_axis(ca, cb)
                assert w.max_variant_axis(ca, cb) == exp
                assert w.max_variant_axis(to_cuda(ca), to_cuda(cb)) == exp


@pytest.mark.parametrize("nq,axis", [(8, 0), (24, 1), (24, 2), (24, 5), (64, None)])
def test_sweep_sheets(torch_cuda, nq, axis):
    """Wavy contact sheets: the oracle's pairs in its order, host and device
    memory, with and without global ids; the pair set is brute force's."""
    ka, ca, ga, kb, cb, gb = _sheets(nq)
    pa, pb = _patch(ka), _patch(kb)
    if axis is None:
        axis = O.max_variant_axis(ca, cb)
    exp = O.sort_and_sweep(pa, ka, ga, pb, kb, gb, axis)
    assert len(exp) > 0
    with CollisionWorld(1) as w:
        got = w.sort_and_sweep(pa, ka, ga, pb, kb, gb, axis)
        assert np.array_equal(got, exp)
        dev = w.sort_and_sweep(pa, to_cuda(ka), to_cuda(ga.astype(np.int64)), pb, to_cuda(kb),
                               to_cuda(gb.astype(np.int64)), axis)
        assert np.array_equal(dev.cpu().numpy().astype(np.uint64), exp)
        loc = w.sort_and_sweep(pa, ka, None, pb, kb, None, axis)
        la = np.arange(ka.shape[0], dtype=np.uint64)
        lb = np.arange(kb.shape[0], dtype=np.uint64)
        assert np.array_equal(loc, O.sort_and_sweep(pa, ka, la, pb, kb, lb, axis))
    brute = O.brute_pairs(ka, kb)
    assert np.array_equal(np.unique(loc, axis=0), brute)


def test_sweep_edge_cases(torch_cuda):
    """Empty sides, one element, a zero-width sweep axis (0 * inf = NaN
    scaled values, every b visited), identical elements (ties in the sort
    keep B's order), and a pair buffer that is too small."""
    ka, ca, ga, kb, cb, gb = _sheets(8, seed=1)
    pa, pb = _patch(ka), _patch(kb)
    with CollisionWorld(1) as w:
        assert w.sort_and_sweep(pa, ka[:0], ga[:0], pb, kb, gb, 0).shape == (0, 2)
        assert w.sort_and_sweep(pa, ka, ga, pb, kb[:0], gb[:0], 0).shape == (0, 2)
        one = w.sort_and_sweep(pa, ka[:1], ga[:1], pb, kb, gb, 1)
        assert np.array_equal(one, O.sort_and_sweep(pa, ka[:1], ga[:1], pb, kb, gb, 1))
        # patch slabs of zero width on the sweep axis: conv = inf, the scaled
        # values are inf or NaN (truncated to 0), every b is visited
        fpa, fpb = pa.copy(), pb.copy()
        fpa[4:6] = fpb[4:6] = pa[4]
        got = w.sort_and_sweep(fpa, ka, ga, fpb, kb, gb, 2)
        assert len(got) > 0
        assert np.array_equal(got, O.sort_and_sweep(fpa, ka, ga, fpb, kb, gb, 2))
        assert len(got) == len(O.brute_pairs(ka, kb))
        # duplicated B elements: equal keys keep their order
        kd = np.repeat(kb[:40], 3, axis=0)
        gd = np.arange(kd.shape[0], dtype=np.uint64)
        got = w.sort_and_sweep(pa, ka, ga, _patch(kd), kd, gd, 0)
        assert np.array_equal(got, O.sort_and_sweep(pa, ka, ga, _patch(kd), kd, gd, 0))
        # the pair buffer must hold every pair
        import ctypes as C
        n = C.c_uint64(0)
        pa_, pb_ = np.ascontiguousarray(pa), np.ascontiguousarray(pb)
        lib = capi.load()
        st = lib.bvh_sort_and_sweep(w.handle, C.c_void_p(pa_.ctypes.data), C.c_void_p(ka.ctypes.data), None,
                                    ka.shape[0], C.c_void_p(pb_.ctypes.data), C.c_void_p(kb.ctypes.data), None,
                                    kb.shape[0], 0, capi.BVH_MEM_HOST, None, 0, C.byref(n))
        assert st == 0 and n.value > 1
        out = np.empty((n.value - 1, 2), np.uint64)
        st = lib.bvh_sort_and_sweep(w.handle, C.c_void_p(pa_.ctypes.data), C.c_void_p(ka.ctypes.data), None,
                                    ka.shape[0], C.c_void_p(pb_.ctypes.data), C.c_void_p(kb.ctypes.data), None,
                                    kb.shape[0], 0, capi.BVH_MEM_HOST, C.c_void_p(out.ctypes.data), n.value - 1,
                                    C.byref(n))
        assert st == capi.BVH_E_OVERFLOW
        with pytest.raises(capi.BvhError):
            w.sort_and_sweep(pa, ka, ga, pb, kb, gb, 13)


def test_sweep_large(torch_cuda):
    """2 x 32,768 triangles (128 x 128 quads), device memory: the oracle's
    pairs, in order."""
    ka, ca, ga, kb, cb, gb = _sheets(128, seed=2)
    pa, pb = _patch(ka), _patch(kb)
    axis = O.max_variant_axis(ca, cb)
    exp = O.sort_and_sweep(pa, ka, ga, pb, kb, gb, axis)
    with CollisionWorld(1) as w:
        assert w.max_variant_axis(to_cuda(ca), to_cuda(cb)) == axis
        got = w.sort_and_sweep(pa, to_cuda(ka), to_cuda(ga.astype(np.int64)), pb, to_cuda(kb),
                               to_cuda(gb.astype(np.int64)), axis)
        assert np.array_equal(got.cpu().numpy().astype(np.uint64), exp)


@pytest.mark.parametrize("nv,eps", [(3, 0.0), (3, 0.01), (8, 0.0), (1, 0.0)])
def test_element_kdops(torch_cuda, nv, eps):
    """bvh_element_kdops (kdop::from_vertices for a batch) equals the
    oracle's element k-DOPs bit for bit, host and device memory."""
    rng = np.random.default_rng(nv)
    n = 3001
    verts = rng.normal(size=(nv, 3, n)) * np.array([2.0, 1.0, 0.5])[None, :, None] + 5.0
    exp, _ = O.elements_poly(verts, eps)
    with CollisionWorld(1) as w:
        assert np.array_equal(w.element_kdops(verts, eps), exp)
        assert np.array_equal(w.element_kdops(to_cuda(verts), eps).cpu().numpy(), exp)
        assert w.element_kdops(verts[:, :, :0]).shape == (0, 26)
