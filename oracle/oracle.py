"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker or the timed CPU
baseline, never as part of the product path.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None

u8p = C.POINTER(C.c_uint8)
dp = C.POINTER(C.c_double)
u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)


class OrcTree(C.Structure):
    _fields_ = [
        ("nnodes", C.c_size_t), ("nleaves", C.c_size_t),
        ("kdop", dp), ("left", C.POINTER(C.c_int64)), ("right", C.POINTER(C.c_int64)),
        ("first", u64p), ("last", u64p),
        ("leaf_kdop", dp), ("leaf_centroid", dp), ("leaf_gid", u64p),
    ]


class OrcPairs(C.Structure):
    _fields_ = [("data", u64p), ("n", C.c_size_t), ("cap", C.c_size_t)]


class OrcCfg(C.Structure):
    _fields_ = [("od", C.c_size_t), ("morton_bits", C.c_int), ("leaf_size", C.c_size_t),
                ("nthreads", C.c_int), ("sample", C.c_double), ("split", C.c_int)]


# split_algorithm, types.hpp:52-57
SPLIT_GEOM_AXIS, SPLIT_ML_GEOM_AXIS, SPLIT_CLUSTERING = 0, 1, 2


class OrcTimes(C.Structure):
    _fields_ = [(k, C.c_double) for k in (
        "t_bounds", "t_morton", "t_sort", "t_cluster", "t_patches", "t_tree",
        "t_broadphase", "t_narrowphase")] + [("patch_pairs", C.c_size_t),
                                             ("patch_pairs_done", C.c_size_t)]


class OrcObject(C.Structure):
    _fields_ = [("n", C.c_size_t), ("kdops", dp), ("centroids", dp), ("gids", u64p),
                ("codes", u64p), ("perm", u32p), ("splits", u64p), ("pbegin", u64p),
                ("pend", u64p), ("patch_kdops", dp), ("patch_centroids", dp)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle library missing: run `python oracle/build.py`")
        L = C.CDLL(LIB_PATH)
        sig = {
            "orc_normal": (C.c_double, [C.c_int, C.c_int]),
            "orc_project": (C.c_double, [dp, C.c_int]),
            "orc_kdop_from_vertices": (None, [dp, C.c_int, C.c_double, dp]),
            "orc_kdop_from_sphere": (None, [dp, C.c_double, dp]),
            "orc_overlap": (C.c_int, [dp, dp]),
            "orc_longest_axis": (C.c_int, [dp]),
            "orc_elements_poly": (None, [dp, C.c_int, C.c_size_t, C.c_double, dp, dp]),
            "orc_elements_sphere": (None, [dp, C.c_size_t, dp, dp]),
            "orc_compute_bounds": (None, [dp, C.c_size_t, dp, dp]),
            "orc_expand32": (C.c_uint32, [C.c_uint32]),
            "orc_expand64": (C.c_uint64, [C.c_uint64]),
            "orc_pdep64": (C.c_uint64, [C.c_uint64, C.c_uint64]),
            "orc_morton32": (C.c_uint32, [C.c_uint32] * 3),
            "orc_morton64": (C.c_uint64, [C.c_uint64] * 3),
            "orc_float_to_uint": (C.c_uint64, [C.c_double]),
            "orc_quantize32": (None, [dp, dp, dp, u32p]),
            "orc_quantize64": (None, [dp, dp, dp, u64p]),
            "orc_morton_codes32": (None, [dp, C.c_size_t, dp, dp, u32p]),
            "orc_morton_codes64": (None, [dp, C.c_size_t, dp, dp, u64p]),
            "orc_radix_sort32": (None, [u32p, u32p, C.c_size_t]),
            "orc_radix_sort64": (None, [u64p, u32p, C.c_size_t]),
            "orc_prefix_sum_u32": (None, [u32p, C.c_size_t]),
            "orc_clz32": (C.c_int, [C.c_uint32]),
            "orc_clz64": (C.c_int, [C.c_uint64]),
            "orc_cluster_splits32": (None, [u32p, C.c_size_t, C.c_size_t, u64p]),
            "orc_cluster_splits64": (None, [u64p, C.c_size_t, C.c_size_t, u64p]),
            "orc_patches": (None, [u64p, C.c_size_t, C.c_size_t, dp, dp, u64p, u64p, dp, dp]),
            "orc_tree_build_top_down": (C.POINTER(OrcTree), [dp, dp, u64p, C.c_size_t, C.c_size_t]),
            "orc_tree_free": (None, [C.POINTER(OrcTree)]),
            "orc_pairs_new": (C.POINTER(OrcPairs), []),
            "orc_pairs_free": (None, [C.POINTER(OrcPairs)]),
            "orc_query_tree": (None, [C.POINTER(OrcTree), dp, C.c_uint64, C.POINTER(OrcPairs)]),
            "orc_self_collision_set": (None, [C.POINTER(OrcTree), C.POINTER(OrcPairs)]),
            "orc_potential_collision_set": (None, [C.POINTER(OrcTree), C.POINTER(OrcTree), C.c_int,
                                                   C.POINTER(OrcPairs)]),
            "orc_max_variant_axis": (C.c_int, [dp, C.c_size_t, dp, C.c_size_t]),
            "orc_sort_and_sweep": (None, [dp, dp, u64p, C.c_size_t, dp, dp, u64p, C.c_size_t, C.c_int,
                                          C.POINTER(OrcPairs)]),
            "orc_brute_pairs": (C.POINTER(OrcPairs), [dp, C.c_size_t, dp, C.c_size_t, C.c_int]),
            "orc_brute_self": (C.POINTER(OrcPairs), [dp, C.c_size_t, C.c_int]),
            "orc_object_build": (None, [C.POINTER(OrcObject), C.POINTER(OrcCfg), C.POINTER(OrcTimes)]),
            "orc_collide": (C.POINTER(OrcPairs), [C.POINTER(OrcObject), C.POINTER(OrcObject),
                                                  C.POINTER(OrcCfg), C.POINTER(OrcTimes)]),
            "orc_now": (C.c_double, []),
            "orc_split_geom_axis": (None, [dp, dp, C.c_size_t, C.c_int, u32p, u64p]),
            "orc_closest_point": (None, [dp, C.c_size_t, dp, dp]),
            "orc_tri_tri": (C.c_int, [dp, dp]),
            "orc_sphere_sphere": (C.c_int, [dp, dp]),
            "orc_tri_tri_batch": (None, [dp, dp, C.c_size_t, C.POINTER(C.c_uint8)]),
            "orc_sphere_sphere_batch": (None, [dp, dp, C.c_size_t, C.POINTER(C.c_uint8)]),
            "orc_split_ml_geom_axis": (None, [dp, dp, C.c_size_t, C.c_int, u32p, u64p]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def _pairs_to_numpy(pp, width=2):
    p = pp.contents
    cnt = p.n
    if cnt == 0:
        out = np.zeros((0, width), np.uint64)
    else:
        out = np.ctypeslib.as_array(p.data, shape=(cnt * 2,)).copy().reshape(-1, width)
    lib().orc_pairs_free(pp)
    return out


# ---- k-DOP ---------------------------------------------------------------
def project(v, axis):
    v = _c(v, np.float64)
    return lib().orc_project(_p(v, dp), axis)


def kdop_from_vertices(verts, eps=0.0):
    v = _c(verts, np.float64).reshape(-1, 3)
    out = np.empty(26, np.float64)
    lib().orc_kdop_from_vertices(_p(v, dp), v.shape[0], eps, _p(out, dp))
    return out


def kdop_from_sphere(c, r):
    c = _c(c, np.float64)
    out = np.empty(26, np.float64)
    lib().orc_kdop_from_sphere(_p(c, dp), r, _p(out, dp))
    return out


def overlap(a, b):
    return bool(lib().orc_overlap(_p(_c(a, np.float64), dp), _p(_c(b, np.float64), dp)))


def longest_axis(k):
    return lib().orc_longest_axis(_p(_c(k, np.float64), dp))


def elements_poly(verts, eps=0.0):
    """verts: (nv, 3, n) SoA planes -> kdops (n, 26), centroids (n, 3)."""
    v = _c(verts, np.float64)
    nv, _, n = v.shape
    kd = np.empty((n, 26), np.float64)
    ce = np.empty((n, 3), np.float64)
    lib().orc_elements_poly(_p(v, dp), nv, n, eps, _p(kd, dp), _p(ce, dp))
    return kd, ce


def elements_sphere(xyzr):
    x = _c(xyzr, np.float64)
    n = x.shape[1]
    kd = np.empty((n, 26), np.float64)
    ce = np.empty((n, 3), np.float64)
    lib().orc_elements_sphere(_p(x, dp), n, _p(kd, dp), _p(ce, dp))
    return kd, ce


# ---- bounds / Morton ------------------------------------------------------
def compute_bounds(kdops):
    kd = _c(kdops, np.float64)
    mn = np.empty(3, np.float64)
    inv = np.empty(3, np.float64)
    lib().orc_compute_bounds(_p(kd, dp), kd.shape[0], _p(mn, dp), _p(inv, dp))
    return mn, inv


def morton_codes(centroids, mn, inv, bits=63):
    ce = _c(centroids, np.float64)
    n = ce.shape[0]
    mn = _c(mn, np.float64)
    inv = _c(inv, np.float64)
    if bits == 30:
        out = np.empty(n, np.uint32)
        lib().orc_morton_codes32(_p(ce, dp), n, _p(mn, dp), _p(inv, dp), _p(out, u32p))
    else:
        out = np.empty(n, np.uint64)
        lib().orc_morton_codes64(_p(ce, dp), n, _p(mn, dp), _p(inv, dp), _p(out, u64p))
    return out


def radix_sort(keys, idx=None):
    """1-bit LSD radix sort (reference algorithm). Returns (sorted_keys, perm)."""
    if keys.dtype == np.uint32:
        k = keys.copy()
        f, t = lib().orc_radix_sort32, u32p
    else:
        k = _c(keys, np.uint64).copy()
        f, t = lib().orc_radix_sort64, u64p
    n = k.shape[0]
    i = np.arange(n, dtype=np.uint32) if idx is None else _c(idx, np.uint32).copy()
    f(_p(k, t), _p(i, u32p), n)
    return k, i


def prefix_sum(v):
    a = _c(v, np.uint32).copy()
    lib().orc_prefix_sum_u32(_p(a, u32p), a.shape[0])
    return a


def cluster_splits(sorted_codes, nsplits):
    n = sorted_codes.shape[0]
    out = np.zeros(max(nsplits, 1), np.uint64)
    if sorted_codes.dtype == np.uint32:
        lib().orc_cluster_splits32(_p(sorted_codes, u32p), n, nsplits, _p(out, u64p))
    else:
        c = _c(sorted_codes, np.uint64)
        lib().orc_cluster_splits64(_p(c, u64p), n, nsplits, _p(out, u64p))
    return out[:nsplits]


def split_permutations(kdops, centroids, depth, ml=False):
    """split_permutations (ml=False) or split_permutations_ml (ml=True) with
    split::mean and axis::longest: (indices, splits) as element_permutations
    holds them (split/split.hpp:285-323)."""
    kd = _c(kdops, np.float64)
    ce = _c(centroids, np.float64)
    n = kd.shape[0]
    idx = np.empty(max(n, 1), np.uint32)
    ns = (1 << depth) - 1 if depth > 0 else 0
    sp = np.zeros(max(ns, 1), np.uint64)
    f = lib().orc_split_ml_geom_axis if ml else lib().orc_split_geom_axis
    f(_p(kd, dp), _p(ce, dp), n, depth, _p(idx, u32p), _p(sp, u64p))
    return idx[:n], sp[:ns]


def closest_point(points, tri):
    """bvh::kokkos::closest_point restated: (n, 3) points, tri (3, 3)."""
    p = _c(np.asarray(points, np.float64).reshape(-1, 3), np.float64)
    t = _c(np.asarray(tri, np.float64).reshape(9), np.float64)
    out = np.empty_like(p)
    if len(p):
        lib().orc_closest_point(_p(p, dp), len(p), _p(t, dp), _p(out, dp))
    return out


def tri_tri(p, q):
    """Closed triangles p, q ((n, 9) or (n, 3, 3): a, b, c) intersect: (n,)
    uint8 (the rule of orc_tri_tri, no reference counterpart)."""
    p = _c(np.asarray(p, np.float64).reshape(-1, 9), np.float64)
    q = _c(np.asarray(q, np.float64).reshape(-1, 9), np.float64)
    out = np.zeros(len(p), np.uint8)
    if len(p):
        lib().orc_tri_tri_batch(_p(p, dp), _p(q, dp), len(p), out.ctypes.data_as(C.POINTER(C.c_uint8)))
    return out


def sphere_sphere(a, b):
    """Closed balls a, b ((n, 4): x, y, z, r) intersect: (n,) uint8."""
    a = _c(np.asarray(a, np.float64).reshape(-1, 4), np.float64)
    b = _c(np.asarray(b, np.float64).reshape(-1, 4), np.float64)
    out = np.zeros(len(a), np.uint8)
    if len(a):
        lib().orc_sphere_sphere_batch(_p(a, dp), _p(b, dp), len(a), out.ctypes.data_as(C.POINTER(C.c_uint8)))
    return out


def patches(splits, od, sorted_kdops, sorted_centroids):
    n = sorted_kdops.shape[0]
    sp = _c(splits, np.uint64) if len(splits) else np.zeros(1, np.uint64)
    sk = _c(sorted_kdops, np.float64)
    sc = _c(sorted_centroids, np.float64)
    b = np.empty(od, np.uint64)
    e = np.empty(od, np.uint64)
    pk = np.empty((od, 26), np.float64)
    pc = np.empty((od, 3), np.float64)
    lib().orc_patches(_p(sp, u64p), od, n, _p(sk, dp), _p(sc, dp), _p(b, u64p), _p(e, u64p),
                      _p(pk, dp), _p(pc, dp))
    return b, e, pk, pc


# ---- trees / queries ------------------------------------------------------
class Tree:
    def __init__(self, kdops, centroids, gids=None, leaf_size=1):
        kd = _c(kdops, np.float64)
        ce = _c(centroids, np.float64)
        n = kd.shape[0]
        g = np.arange(n, dtype=np.uint64) if gids is None else _c(gids, np.uint64)
        self._keep = (kd, ce, g)
        self.ptr = lib().orc_tree_build_top_down(_p(kd, dp), _p(ce, dp), _p(g, u64p), n, leaf_size)

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().orc_tree_free(self.ptr)
            self.ptr = None

    @property
    def nnodes(self):
        return self.ptr.contents.nnodes

    def query(self, kdop, gid):
        pp = lib().orc_pairs_new()
        lib().orc_query_tree(self.ptr, _p(_c(kdop, np.float64), dp), gid, pp)
        return _pairs_to_numpy(pp)


def self_collision_set(tree):
    pp = lib().orc_pairs_new()
    lib().orc_self_collision_set(tree.ptr, pp)
    return _pairs_to_numpy(pp)


def potential_collision_set(lt, rt, as_written=False):
    pp = lib().orc_pairs_new()
    lib().orc_potential_collision_set(lt.ptr, rt.ptr, int(as_written), pp)
    return _pairs_to_numpy(pp)


def max_variant_axis(ca, cb):
    a = _c(ca, np.float64)
    b = _c(cb, np.float64)
    return lib().orc_max_variant_axis(_p(a, dp), a.shape[0], _p(b, dp), b.shape[0])


def sort_and_sweep(pa, ka, ga, pb, kb, gb, axis):
    pp = lib().orc_pairs_new()
    ka = _c(ka, np.float64); kb = _c(kb, np.float64)
    ga = _c(ga, np.uint64); gb = _c(gb, np.uint64)
    lib().orc_sort_and_sweep(_p(_c(pa, np.float64), dp), _p(ka, dp), _p(ga, u64p), ka.shape[0],
                             _p(_c(pb, np.float64), dp), _p(kb, dp), _p(gb, u64p), kb.shape[0],
                             axis, pp)
    return _pairs_to_numpy(pp)


def brute_pairs(ka, kb, nthreads=8):
    ka = _c(ka, np.float64)
    kb = _c(kb, np.float64)
    pp = lib().orc_brute_pairs(_p(ka, dp), ka.shape[0], _p(kb, dp), kb.shape[0], nthreads)
    return _pairs_to_numpy(pp)


def brute_self(k, nthreads=8):
    k = _c(k, np.float64)
    pp = lib().orc_brute_self(_p(k, dp), k.shape[0], nthreads)
    return _pairs_to_numpy(pp)


# ---- pipeline -------------------------------------------------------------
class Object:
    """One object built the reference way (bounds, Morton, 1-bit radix sort,
    cluster splits, patches; or the geom_axis / ml_geom_axis splitters with
    split=SPLIT_GEOM_AXIS / SPLIT_ML_GEOM_AXIS, od a power of two)."""

    def __init__(self, kdops, centroids, gids=None, od=1, morton_bits=63, leaf_size=1,
                 nthreads=1, sample=1.0, times=None, split=SPLIT_CLUSTERING):
        self.kdops = _c(kdops, np.float64)
        self.centroids = _c(centroids, np.float64)
        n = self.kdops.shape[0]
        self.gids = np.arange(n, dtype=np.uint64) if gids is None else _c(gids, np.uint64)
        self.codes = np.empty(n, np.uint64)
        self.perm = np.empty(n, np.uint32)
        self.splits = np.zeros(max(od - 1, 1), np.uint64)
        self.pbegin = np.empty(od, np.uint64)
        self.pend = np.empty(od, np.uint64)
        self.patch_kdops = np.empty((od, 26), np.float64)
        self.patch_centroids = np.empty((od, 3), np.float64)
        self.cfg = OrcCfg(od, morton_bits, leaf_size, nthreads, sample, split)
        self.s = OrcObject(n, _p(self.kdops, dp), _p(self.centroids, dp), _p(self.gids, u64p),
                           _p(self.codes, u64p), _p(self.perm, u32p), _p(self.splits, u64p),
                           _p(self.pbegin, u64p), _p(self.pend, u64p), _p(self.patch_kdops, dp),
                           _p(self.patch_centroids, dp))
        self.times = times if times is not None else OrcTimes()
        lib().orc_object_build(C.byref(self.s), C.byref(self.cfg), C.byref(self.times))
        self.splits = self.splits[:max(od - 1, 0)]

    def collide(self, other, times=None):
        """Records (k, 4) uint64 {patch_p, element_p, patch_q, element_q}
        sorted by (element_p, element_q)."""
        tm = times if times is not None else self.times
        pp = lib().orc_collide(C.byref(self.s), C.byref(other.s), C.byref(self.cfg), C.byref(tm))
        return _pairs_to_numpy(pp).reshape(-1, 4)
