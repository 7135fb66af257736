"""Host-side mirror of distBVH's collision API over the C ABI.

Names follow the reference: CollisionWorld ~ bvh::collision_world
(collision_world.hpp:54-114), CollisionObject ~ bvh::collision_object
(collision_object.hpp:55-283).  Element data is a numpy array (host) or a
torch CUDA tensor (device, read in place); see include/distbvh.h for layouts.
"""
import ctypes as C

import numpy as np

from . import capi
from .capi import call

CLUSTERING = capi.BVH_SPLIT_CLUSTERING
GEOM_AXIS = capi.BVH_SPLIT_GEOM_AXIS
ML_GEOM_AXIS = capi.BVH_SPLIT_ML_GEOM_AXIS


def _ptr_mem(a, dtype):
    """(pointer, memory kind, keepalive) of a numpy array or a CUDA tensor."""
    if a is None:
        return None, capi.BVH_MEM_HOST, None
    if hasattr(a, "data_ptr") and getattr(a, "is_cuda", False):
        return C.c_void_p(a.data_ptr()), capi.BVH_MEM_DEVICE, a
    arr = np.ascontiguousarray(a, dtype=dtype)
    return C.c_void_p(arr.ctypes.data), capi.BVH_MEM_HOST, arr


class CollisionWorld:
    """collision_world(od, config)."""

    def __init__(self, overdecomposition=1, morton_bits=63, leaf_size=1, precision="fp64",
                 sort_results=False, timing=False, stream=None, device=0, serial_builds=False):
        cfg = capi.WorldConfig()
        capi.load().bvh_world_config_init(C.byref(cfg))
        cfg.morton_bits = morton_bits
        cfg.leaf_size = leaf_size
        cfg.precision = capi.BVH_PREC_FP32_OUTWARD if precision == "fp32" else capi.BVH_PREC_FP64
        cfg.flags = (capi.BVH_FLAG_SORT_RESULTS if sort_results else 0) | \
                    (capi.BVH_FLAG_TIMING if timing else 0) | \
                    (capi.BVH_FLAG_SERIAL_BUILDS if serial_builds else 0)
        cfg.stream = stream
        cfg.device = device
        h = C.c_void_p()
        call("bvh_world_create", overdecomposition, C.byref(cfg), C.byref(h))
        self.handle = h
        self.objects = []
        self.od = overdecomposition

    def close(self):
        if getattr(self, "handle", None):
            capi.load().bvh_world_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def overdecomposition_factor(self):
        v = C.c_size_t()
        call("bvh_world_overdecomposition", self.handle, C.byref(v))
        return v.value

    def create_collision_object(self):
        h = C.c_void_p()
        call("bvh_create_collision_object", self.handle, C.byref(h))
        o = CollisionObject(self, h)
        self.objects.append(o)
        return o

    def num_collision_objects(self):
        v = C.c_size_t()
        call("bvh_num_collision_objects", self.handle, C.byref(v))
        return v.value

    def start_iteration(self):
        call("bvh_start_iteration", self.handle)

    def finish_iteration(self):
        call("bvh_finish_iteration", self.handle)

    def synchronize(self):
        call("bvh_synchronize", self.handle)

    def launch(self):
        """bvh_get_launch: the narrowphase launch shape."""
        t = capi.Tuning()
        call("bvh_get_launch", self.handle, C.byref(t))
        return {"batch": t.batch, "block_size": t.block_size, "xcd_remap": bool(t.xcd_remap)}

    def set_launch(self, batch=None, block_size=None, xcd_remap=None):
        """bvh_set_launch: change any of batch (1, 2, 4), block size (64,
        128, 256) and XCD remap; the others keep their values."""
        cur = self.launch()
        t = capi.Tuning()
        t.batch = cur["batch"] if batch is None else batch
        t.block_size = cur["block_size"] if block_size is None else block_size
        t.xcd_remap = int(cur["xcd_remap"] if xcd_remap is None else xcd_remap)
        call("bvh_set_launch", self.handle, C.byref(t))

    def stage_times(self):
        names = (C.c_char_p * 64)()
        ms = (C.c_double * 64)()
        cnt = C.c_size_t()
        call("bvh_stage_times", self.handle, names, ms, 64, C.byref(cnt))
        return [(names[i].decode(), ms[i]) for i in range(min(cnt.value, 64))]

    # stand-alone primitives on device tensors (torch)
    def capture(self):
        """Context manager capturing the world's stream into a StepGraph:
        `with world.capture() as g: ...step...`, then `g.launch()`."""
        return StepGraph(self)

    def sort(self, keys, vals=None):
        fn = "bvh_sort_u64" if keys.element_size() == 8 else "bvh_sort_u32"
        call(fn, self.handle, C.c_void_p(keys.data_ptr()),
             C.c_void_p(vals.data_ptr()) if vals is not None else None, keys.numel())

    def exclusive_scan(self, vals):
        call("bvh_exclusive_scan_u32", self.handle, C.c_void_p(vals.data_ptr()), vals.numel())

    def closest_point(self, points, tri):
        """bvh_closest_point: the closest point of triangle tri ((3, 3): a, b,
        c) to each of the (n, 3) points (numpy: host memory; CUDA tensor:
        device memory, the result a new tensor on the same device)."""
        t = np.ascontiguousarray(np.asarray(tri, np.float64).reshape(9))
        if hasattr(points, "data_ptr"):
            import torch
            out = torch.empty_like(points)
            call("bvh_closest_point", self.handle, C.c_void_p(points.data_ptr()), points.shape[0],
                 C.c_void_p(t.ctypes.data), C.c_void_p(out.data_ptr()), capi.BVH_MEM_DEVICE)
            return out
        p = np.ascontiguousarray(points, np.float64)
        out = np.empty_like(p)
        call("bvh_closest_point", self.handle, C.c_void_p(p.ctypes.data), p.shape[0], C.c_void_p(t.ctypes.data),
             C.c_void_p(out.ctypes.data), capi.BVH_MEM_HOST)
        return out

    def element_kdops(self, verts, epsilon=0.0):
        """bvh_element_kdops: the (n, 26) k-DOPs of (nv, 3, n) vertex planes
        (numpy: host memory; CUDA tensor: device memory, result on device)."""
        nv, _, n = tuple(verts.shape)
        if hasattr(verts, "data_ptr"):
            import torch
            out = torch.empty((n, 26), dtype=torch.float64, device=verts.device)
            call("bvh_element_kdops", self.handle, C.c_void_p(verts.data_ptr()), nv, n, float(epsilon),
                 capi.BVH_MEM_DEVICE, C.c_void_p(out.data_ptr()))
            return out
        v = np.ascontiguousarray(verts, np.float64)
        out = np.empty((n, 26), np.float64)
        call("bvh_element_kdops", self.handle, C.c_void_p(v.ctypes.data), nv, n, float(epsilon),
             capi.BVH_MEM_HOST, C.c_void_p(out.ctypes.data))
        return out

    def max_variant_axis(self, ca, cb):
        """bvh_max_variant_axis: the axis of largest variance of the (na, 3)
        and (nb, 3) centroids together (numpy or CUDA tensors, both alike)."""
        ax = C.c_int32(-1)
        if hasattr(ca, "data_ptr"):
            call("bvh_max_variant_axis", self.handle, C.c_void_p(ca.data_ptr()), ca.shape[0],
                 C.c_void_p(cb.data_ptr()), cb.shape[0], capi.BVH_MEM_DEVICE, C.byref(ax))
            return ax.value
        a = np.ascontiguousarray(ca, np.float64)
        b = np.ascontiguousarray(cb, np.float64)
        call("bvh_max_variant_axis", self.handle, C.c_void_p(a.ctypes.data), a.shape[0], C.c_void_p(b.ctypes.data),
             b.shape[0], capi.BVH_MEM_HOST, C.byref(ax))
        return ax.value

    def sort_and_sweep(self, pa, ka, ga, pb, kb, gb, axis):
        """bvh_sort_and_sweep: the (k, 2) uint64 pairs (id in A, id in B) of
        overlapping elements of patches A and B, in the reference's callback
        order per element of A.  pa / pb: the patches' 26-DOPs; ka / kb:
        (n, 26) element k-DOPs; ga / gb: global ids or None (local
        indices).  numpy arrays (host memory) or CUDA tensors (device
        memory; the pairs are then a CUDA tensor)."""
        pa = np.ascontiguousarray(np.asarray(pa, np.float64).reshape(26))
        pb = np.ascontiguousarray(np.asarray(pb, np.float64).reshape(26))
        n = C.c_uint64(0)
        dev = hasattr(ka, "data_ptr")
        if dev:
            import torch
            ptr = lambda t: C.c_void_p(t.data_ptr() if t is not None else None)
            args = (ptr(ka), ka.shape[0], C.c_void_p(pb.ctypes.data), ptr(kb), ptr(gb), kb.shape[0])
            call("bvh_sort_and_sweep", self.handle, C.c_void_p(pa.ctypes.data), args[0], ptr(ga), args[1], *args[2:],
                 int(axis), capi.BVH_MEM_DEVICE, None, 0, C.byref(n))
            out = torch.empty((n.value, 2), dtype=torch.int64, device=ka.device)
            if n.value:
                call("bvh_sort_and_sweep", self.handle, C.c_void_p(pa.ctypes.data), args[0], ptr(ga), args[1],
                     *args[2:], int(axis), capi.BVH_MEM_DEVICE, C.c_void_p(out.data_ptr()), n.value, C.byref(n))
            return out
        ka = np.ascontiguousarray(ka, np.float64)
        kb = np.ascontiguousarray(kb, np.float64)
        ga = None if ga is None else np.ascontiguousarray(ga, np.uint64)
        gb = None if gb is None else np.ascontiguousarray(gb, np.uint64)
        ptr = lambda a: C.c_void_p(a.ctypes.data if a is not None else None)
        args = (C.c_void_p(pa.ctypes.data), ptr(ka), ptr(ga), ka.shape[0], C.c_void_p(pb.ctypes.data), ptr(kb),
                ptr(gb), kb.shape[0], int(axis), capi.BVH_MEM_HOST)
        call("bvh_sort_and_sweep", self.handle, *args, None, 0, C.byref(n))
        out = np.empty((n.value, 2), np.uint64)
        if n.value:
            call("bvh_sort_and_sweep", self.handle, *args, C.c_void_p(out.ctypes.data), n.value, C.byref(n))
        return out


class StepGraph:
    """A step captured into a HIP graph (bvh_capture_begin / _end)."""

    def __init__(self, world):
        self.world = world
        self.handle = None

    def __enter__(self):
        call("bvh_capture_begin", self.world.handle)
        return self

    def __exit__(self, exc_type, exc, tb):
        h = C.c_void_p()
        if exc_type is not None:
            capi.load().bvh_capture_end(self.world.handle, C.byref(h))
            if h.value:
                capi.load().bvh_graph_destroy(h)
            return False
        call("bvh_capture_end", self.world.handle, C.byref(h))
        self.handle = h
        return False

    def launch(self):
        call("bvh_graph_launch", self.handle)

    def close(self):
        if self.handle is not None:
            capi.load().bvh_graph_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class CollisionObject:
    """collision_object: owned by its world."""

    def __init__(self, world, handle):
        self.world = world
        self.handle = handle
        self._keep = None
        self.n = 0

    def id(self):
        v = C.c_size_t()
        call("bvh_object_id", self.handle, C.byref(v))
        return v.value

    # ---- set_entity_data -------------------------------------------------------
    def set_entity_data_poly(self, verts, epsilon=0.0, global_ids=None, split_algorithm=CLUSTERING):
        """verts: (nv, 3, n) float64 SoA planes (numpy or CUDA tensor)."""
        nv, _, n = tuple(verts.shape)
        p, mem, keep = _ptr_mem(verts, np.float64)
        g, gmem, gkeep = _ptr_mem(global_ids, np.uint64)
        if global_ids is not None and gmem != mem:
            raise ValueError("global ids and elements must live in the same memory")
        # the previous arrays stay referenced until the library has finished
        # the build that reads them
        call("bvh_set_entity_data_poly", self.handle, p, nv, n, g, float(epsilon), mem, split_algorithm)
        self._keep = (keep, gkeep)
        self.n = n

    def set_entity_data_sphere(self, xyzr, global_ids=None, split_algorithm=CLUSTERING):
        """xyzr: (4, n) float64 planes x, y, z, r."""
        n = tuple(xyzr.shape)[1]
        p, mem, keep = _ptr_mem(xyzr, np.float64)
        g, _, gkeep = _ptr_mem(global_ids, np.uint64)
        call("bvh_set_entity_data_sphere", self.handle, p, n, g, mem, split_algorithm)
        self._keep = (keep, gkeep)
        self.n = n

    def set_entity_data_snapshot(self, kdops, centroids, global_ids=None, split_algorithm=CLUSTERING):
        """kdops: (n, 26) float64, centroids: (n, 3) float64."""
        n = tuple(kdops.shape)[0]
        p, mem, keep = _ptr_mem(kdops, np.float64)
        c, _, ckeep = _ptr_mem(centroids, np.float64)
        g, _, gkeep = _ptr_mem(global_ids, np.uint64)
        call("bvh_set_entity_data_snapshot", self.handle, p, c, n, g, mem, split_algorithm)
        self._keep = (keep, ckeep, gkeep)
        self.n = n

    def refit(self, data):
        """New element data of the last set_entity_data's kind and count
        (same array shape); the tree keeps its topology and every k-DOP is
        recomputed (bvh_refit)."""
        p, mem, keep = _ptr_mem(data, np.float64)
        call("bvh_refit", self.handle, p, mem)
        self._keep = (keep,) + tuple(self._keep[1:])

    def device_bytes(self):
        v = C.c_uint64()
        call("bvh_object_device_bytes", self.handle, C.byref(v))
        return v.value

    def release(self):
        """bvh_object_release: free the object's device memory; it is empty
        until its next set_entity_data."""
        call("bvh_object_release", self.handle)
        self._keep = None
        self.n = 0

    def root_kdop(self):
        k = np.zeros(26)
        call("bvh_object_root_kdop", self.handle, C.c_void_p(k.ctypes.data))
        return k

    def autotune(self, other=None):
        """bvh_autotune: the fastest batch / block size / XCD remap for the
        query self x other (self collision when other is None)."""
        t = capi.Tuning()
        call("bvh_autotune", self.handle, (other or self).handle, C.byref(t))
        return {"batch": t.batch, "block_size": t.block_size, "xcd_remap": bool(t.xcd_remap),
                "ms": t.ms, "tried_ms": list(t.tried_ms[:t.tried])}

    def init_broadphase(self):
        call("bvh_init_broadphase", self.handle)

    def broadphase(self, other):
        call("bvh_broadphase", self.handle, other.handle)

    def self_collision(self):
        call("bvh_self_collision", self.handle)

    def end_phase(self):
        call("bvh_end_phase", self.handle)

    # ---- results ---------------------------------------------------------------
    def result_count(self):
        v = C.c_uint64()
        call("bvh_result_count", self.handle, C.byref(v))
        return v.value

    def results(self):
        """(k, 4) uint64 records {patch_p, element_p, patch_q, element_q}."""
        n = self.result_count()
        out = np.zeros((max(n, 1), 4), np.uint64)
        call("bvh_result_records", self.handle, C.c_void_p(out.ctypes.data), n, capi.BVH_MEM_HOST)
        return out[:n]

    def for_each_result(self, fn):
        """for_each_result<detailed_narrowphase_result>(f), collision_object.hpp:177-187."""
        for r in self.results():
            fn(r)

    def pairs(self):
        """(k, 2) input-index pairs."""
        n = self.result_count()
        out = np.zeros(max(n, 1), np.uint64)
        call("bvh_result_pairs", self.handle, C.c_void_p(out.ctypes.data), n, capi.BVH_MEM_HOST)
        out = out[:n]
        return np.stack([out >> np.uint64(32), out & np.uint64(0xffffffff)], axis=1)

    def result_stats(self):
        v = C.c_uint64()
        pp = C.c_uint64()
        cand = C.c_uint64()
        rck = C.c_uint64()
        err = C.c_double()
        call("bvh_result_stats", self.handle, C.byref(v), C.byref(pp), C.byref(cand), C.byref(rck), C.byref(err))
        return {"node_visits": v.value, "patch_pairs": pp.value, "candidate_pairs": cand.value,
                "rechecked_pairs": rck.value, "max_slab_error_ulps": err.value}

    def digest(self):
        """Order-independent (sum, xor) digest of the pair keys (device)."""
        d = (C.c_uint64 * 2)()
        call("bvh_result_digest", self.handle, d)
        return int(d[0]), int(d[1])

    def pairs_device(self):
        p = C.c_void_p()
        n = C.c_uint64()
        call("bvh_result_pairs_device", self.handle, C.byref(p), C.byref(n))
        return p.value, n.value

    def patch_pairs(self):
        n = C.c_uint64()
        call("bvh_patch_pair_count", self.handle, C.byref(n))
        out = np.zeros((max(n.value, 1), 2), np.uint64)
        call("bvh_patch_pairs", self.handle, C.c_void_p(out.ctypes.data), n.value)
        return out[:n.value]

    def local_patches(self):
        od = self.world.od
        arr = (capi.PatchInfo * od)()
        call("bvh_local_patches", self.handle, arr, od)
        return [{"global_id": p.global_id, "size": p.size, "kdop": np.array(p.kdop[:]),
                 "centroid": np.array(p.centroid[:])} for p in arr]

    def patch_tree(self):
        """bvh_object_patch_tree: the patch-level tree for_each_tree hands out,
        pre-order: (nodes, leaf_patches) with nodes a list of dicts
        (parent_offset, child_offsets, kdop, entity_range)."""
        nn = C.c_size_t()
        nl = C.c_size_t()
        call("bvh_object_patch_tree", self.handle, None, 0, C.byref(nn), None, 0, C.byref(nl))
        nodes = (capi.TreeNode * max(nn.value, 1))()
        leaves = np.zeros(max(nl.value, 1), np.uint64)
        call("bvh_object_patch_tree", self.handle, nodes, nn.value, C.byref(nn), C.c_void_p(leaves.ctypes.data),
             nl.value, C.byref(nl))
        out = [{"parent_offset": t.parent_offset, "child_offsets": tuple(t.child_offsets),
                "kdop": np.array(t.kdop[:]), "entity_range": tuple(t.entity_range)} for t in nodes[:nn.value]]
        return out, leaves[:nl.value]

    def patch_elements(self, k):
        cnt = C.c_uint64()
        out = np.zeros(max(self.n, 1), np.uint64)
        call("bvh_patch_elements", self.handle, k, C.c_void_p(out.ctypes.data), self.n, C.byref(cnt))
        return out[:cnt.value]

    # ---- stage outputs -----------------------------------------------------------
    def bounds(self):
        mn = np.zeros(3)
        inv = np.zeros(3)
        call("bvh_object_bounds", self.handle, C.c_void_p(mn.ctypes.data), C.c_void_p(inv.ctypes.data))
        return mn, inv

    def sorted(self):
        codes = np.zeros(max(self.n, 1), np.uint64)
        perm = np.zeros(max(self.n, 1), np.uint32)
        call("bvh_object_sorted", self.handle, C.c_void_p(codes.ctypes.data), C.c_void_p(perm.ctypes.data),
             capi.BVH_MEM_HOST)
        return codes[:self.n], perm[:self.n]

    def splits(self):
        od = self.world.od
        out = np.zeros(max(od - 1, 1), np.uint64)
        call("bvh_object_splits", self.handle, C.c_void_p(out.ctypes.data), max(od - 1, 0))
        return out[:od - 1]

    def sorted_kdops(self):
        out = np.zeros((max(self.n, 1), 26))
        call("bvh_object_sorted_kdops", self.handle, C.c_void_p(out.ctypes.data), capi.BVH_MEM_HOST)
        return out[:self.n]

    def tree(self):
        nn = C.c_size_t()
        root = C.c_size_t()
        call("bvh_object_tree", self.handle, None, None, None, 0, C.byref(nn), C.byref(root))
        m = nn.value
        kd = np.zeros((max(m, 1), 26))
        ch = np.zeros((max(m, 1), 2), np.int64)
        rg = np.zeros((max(m, 1), 2), np.uint64)
        call("bvh_object_tree", self.handle, C.c_void_p(kd.ctypes.data), C.c_void_p(ch.ctypes.data),
             C.c_void_p(rg.ctypes.data), m, C.byref(nn), C.byref(root))
        return {"kdop": kd[:m], "children": ch[:m], "range": rg[:m], "root": root.value}
