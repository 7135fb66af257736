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
GIVEN = capi.BVH_SPLIT_GIVEN


_TORCH_OK = {np.dtype(np.float64): ("float64",), np.dtype(np.uint64): ("int64", "uint64"),
             np.dtype(np.uint32): ("int32", "uint32")}


def _check_shape(got, want, what):
    if want is None:
        return
    if len(got) != len(want) or any(w is not None and g != w for g, w in zip(got, want)):
        spec = "(" + ", ".join("n" if w is None else str(w) for w in want) + ")"
        raise ValueError("%s: shape %s, expected %s" % (what, tuple(got), spec))


def _is_cuda_tensor(a):
    return hasattr(a, "data_ptr") and getattr(a, "is_cuda", False)


def _device_ptr(a, dtype, shape, what):
    """_ptr_mem of an argument that must be a CUDA tensor."""
    if not _is_cuda_tensor(a):
        raise ValueError("%s: expected a CUDA tensor" % what)
    return _ptr_mem(a, dtype, shape, what)


def _ptr_mem(a, dtype, shape=None, what="array"):
    """(pointer, memory kind, keepalive) of a numpy array or a CUDA tensor.

    A CUDA tensor is read in place, so it must already be what the C ABI
    reads: the element type (float64; 64-bit integers for global ids), a
    contiguous layout and the expected shape (None entries are free).
    Anything else raises ValueError instead of being read as another layout.
    numpy input is converted (a contiguous copy of the right type) and its
    shape checked."""
    if a is None:
        return None, capi.BVH_MEM_HOST, None
    dt = np.dtype(dtype)
    if _is_cuda_tensor(a):
        tname = str(a.dtype).replace("torch.", "")
        if tname not in _TORCH_OK[dt]:
            raise ValueError("%s: CUDA tensor of dtype %s, expected %s" % (what, tname, " or ".join(_TORCH_OK[dt])))
        if not a.is_contiguous():
            raise ValueError("%s: CUDA tensor is not contiguous (a permuted or strided view); pass "
                             ".contiguous()" % what)
        _check_shape(tuple(a.shape), shape, what)
        return C.c_void_p(a.data_ptr()), capi.BVH_MEM_DEVICE, a
    arr = np.ascontiguousarray(a, dtype=dt)
    _check_shape(arr.shape, shape, what)
    return C.c_void_p(arr.ctypes.data), capi.BVH_MEM_HOST, arr


class CollisionWorld:
    """collision_world(od, config)."""

    def __init__(self, overdecomposition=1, morton_bits=63, leaf_size=1, precision="fp64",
                 sort_results=False, timing=False, stream=None, device=0, serial_builds=False):
        cfg = capi.WorldConfig()
        capi.load().bvh_world_config_init(C.byref(cfg))
        cfg.morton_bits = morton_bits
        cfg.leaf_size = leaf_size
        forms = {"fp64": capi.BVH_PREC_FP64, "fp32": capi.BVH_PREC_FP32_OUTWARD,
                 "fp64_records": capi.BVH_PREC_FP64_RECORDS}
        if precision not in forms:
            raise ValueError("precision must be one of %s" % sorted(forms))
        cfg.precision = forms[precision]
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
        sp = C.c_void_p()
        call("bvh_world_stream", h, C.byref(sp))
        self._stream_ptr = sp.value or 0
        self._ext_stream = None

    def after_torch(self):
        """Orders the world's stream after the work torch has enqueued so far
        on its current stream (an event the world's stream waits for), so
        device tensors written there are complete before the library reads
        them.  Called before every call that reads a CUDA tensor."""
        import torch

        cur = torch.cuda.current_stream()
        if cur.cuda_stream == self._stream_ptr:
            return
        if self._ext_stream is None:
            self._ext_stream = torch.cuda.ExternalStream(self._stream_ptr, device=cur.device)
        ev = torch.cuda.Event()
        ev.record(cur)
        self._ext_stream.wait_event(ev)

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
    def broadphase_batch(self, objects, others):
        """bvh_broadphase_batch: objects[i].broadphase(others[i]) for every i,
        coalesced (one narrowphase launch for all pairs)."""
        k = len(objects)
        if len(others) != k:
            raise ValueError("objects and others differ in length")
        a = (C.c_void_p * max(k, 1))(*[o.handle.value if hasattr(o.handle, "value") else o.handle for o in objects])
        b = (C.c_void_p * max(k, 1))(*[o.handle.value if hasattr(o.handle, "value") else o.handle for o in others])
        call("bvh_broadphase_batch", a, b, k)

    def capture(self):
        """Context manager capturing the world's stream into a StepGraph:
        `with world.capture() as g: ...step...`, then `g.launch()`."""
        return StepGraph(self)

    def sort(self, keys, vals=None):
        """bvh_sort_u64 / bvh_sort_u32 in place on CUDA tensors: 64- or 32-bit
        integer keys and optional 32-bit integer values of the same length."""
        kt = np.uint64 if keys.element_size() == 8 else np.uint32
        kp, _, _ = _device_ptr(keys, kt, (None,), "keys")
        vp = None
        if vals is not None:
            vp, _, _ = _device_ptr(vals, np.uint32, (keys.numel(),), "values")
        self.after_torch()
        call("bvh_sort_u64" if kt is np.uint64 else "bvh_sort_u32", self.handle, kp, vp, keys.numel())

    def exclusive_scan(self, vals):
        """bvh_exclusive_scan_u32 in place on a CUDA tensor of 32-bit integers."""
        vp, _, _ = _device_ptr(vals, np.uint32, (None,), "values")
        self.after_torch()
        call("bvh_exclusive_scan_u32", self.handle, vp, vals.numel())

    def closest_point(self, points, tri):
        """bvh_closest_point: the closest point of triangle tri ((3, 3): a, b,
        c) to each of the (n, 3) points (numpy: host memory; CUDA tensor:
        device memory, the result a new tensor on the same device)."""
        t = np.ascontiguousarray(np.asarray(tri, np.float64).reshape(9))
        if hasattr(points, "data_ptr"):
            import torch
            _device_ptr(points, np.float64, (None, 3), "points")
            out = torch.empty_like(points)
            self.after_torch()
            call("bvh_closest_point", self.handle, C.c_void_p(points.data_ptr()), points.shape[0],
                 C.c_void_p(t.ctypes.data), C.c_void_p(out.data_ptr()), capi.BVH_MEM_DEVICE)
            return out
        p = np.ascontiguousarray(points, np.float64)
        out = np.empty_like(p)
        call("bvh_closest_point", self.handle, C.c_void_p(p.ctypes.data), p.shape[0], C.c_void_p(t.ctypes.data),
             C.c_void_p(out.ctypes.data), capi.BVH_MEM_HOST)
        return out

    def _pairs_intersect(self, fn, p, q, width):
        if _is_cuda_tensor(p):
            import torch
            pp, _, _ = _device_ptr(p, np.float64, (None, width), "first of the pairs")
            qp, _, _ = _device_ptr(q, np.float64, (p.shape[0], width), "second of the pairs")
            out = torch.empty(p.shape[0], dtype=torch.uint8, device=p.device)
            self.after_torch()
            call(fn, self.handle, pp, qp, p.shape[0], capi.BVH_MEM_DEVICE, C.c_void_p(out.data_ptr()))
            return out
        a = np.ascontiguousarray(np.asarray(p, np.float64).reshape(-1, width))
        b = np.ascontiguousarray(np.asarray(q, np.float64).reshape(-1, width))
        _check_shape(b.shape, (a.shape[0], width), "second of the pairs")
        out = np.zeros(a.shape[0], np.uint8)
        call(fn, self.handle, C.c_void_p(a.ctypes.data), C.c_void_p(b.ctypes.data), a.shape[0], capi.BVH_MEM_HOST,
             C.c_void_p(out.ctypes.data))
        return out

    def triangle_pairs_intersect(self, p, q):
        """bvh_triangle_pairs_intersect: 1 where closed triangles p[i], q[i]
        ((n, 9): vertices a, b, c) meet (numpy, or CUDA tensors)."""
        return self._pairs_intersect("bvh_triangle_pairs_intersect", p, q, 9)

    def sphere_pairs_intersect(self, a, b):
        """bvh_sphere_pairs_intersect: 1 where closed balls a[i], b[i] ((n, 4):
        x, y, z, r) meet."""
        return self._pairs_intersect("bvh_sphere_pairs_intersect", a, b, 4)

    # ---- wire formats -------------------------------------------------------------
    @staticmethod
    def _buf_ptr(buf):
        if _is_cuda_tensor(buf):
            if not buf.is_contiguous():
                raise ValueError("wire buffer: CUDA tensor is not contiguous")
            return C.c_void_p(buf.data_ptr()), capi.BVH_MEM_DEVICE, buf.numel() * buf.element_size()
        b = np.ascontiguousarray(buf).view(np.uint8)
        return C.c_void_p(b.ctypes.data), capi.BVH_MEM_HOST, b.nbytes

    def unpack_patches(self, buf):
        """bvh_unpack_patches: 256-byte patch records (numpy bytes or a CUDA
        uint8 tensor) -> list of dicts like CollisionObject.local_patches()."""
        p, mem, nbytes = self._buf_ptr(buf)
        if mem == capi.BVH_MEM_DEVICE:
            self.after_torch()
        cnt = C.c_uint64()
        call("bvh_unpack_patches", self.handle, p, nbytes, mem, None, 0, C.byref(cnt))
        arr = (capi.PatchInfo * max(cnt.value, 1))()
        call("bvh_unpack_patches", self.handle, p, nbytes, mem, arr, cnt.value, C.byref(cnt))
        return [{"global_id": q.global_id, "size": q.size, "kdop": np.array(q.kdop[:]),
                 "centroid": np.array(q.centroid[:])} for q in arr[:cnt.value]]

    def unpack_elements(self, buf, count, kind, verts_per_element=3):
        """bvh_unpack_elements: count element records -> (data, centroids,
        global ids) in set_entity_data's layout: (3 nv, n) -> reshaped (nv,
        3, n) planes for polytopes, (4, n) for spheres, (n, 26) k-DOPs and
        (n, 3) centroids for snapshots.  numpy in, numpy out; a CUDA tensor
        in, CUDA tensors out."""
        p, mem, nbytes = self._buf_ptr(buf)
        nv = int(verts_per_element)
        if kind == capi.BVH_ELEM_POLY:
            dshape = (nv, 3, count)
        elif kind == capi.BVH_ELEM_SPHERE:
            dshape = (4, count)
        else:
            dshape = (count, 26)
        if mem == capi.BVH_MEM_DEVICE:
            import torch
            self.after_torch()
            data = torch.empty(dshape, dtype=torch.float64, device=buf.device)
            cent = torch.empty((count, 3), dtype=torch.float64, device=buf.device) \
                if kind == capi.BVH_ELEM_SNAPSHOT else None
            gids = torch.empty(count, dtype=torch.int64, device=buf.device)
            ptr = lambda t: C.c_void_p(t.data_ptr()) if t is not None else None
        else:
            data = np.zeros(dshape, np.float64)
            cent = np.zeros((count, 3), np.float64) if kind == capi.BVH_ELEM_SNAPSHOT else None
            gids = np.zeros(count, np.uint64)
            ptr = lambda a: C.c_void_p(a.ctypes.data) if a is not None else None
        call("bvh_unpack_elements", self.handle, p, count, kind, nv, ptr(data), ptr(cent), ptr(gids), mem)
        return data, cent, gids

    def unpack_results(self, buf):
        """bvh_unpack_results: a result message -> (k, 4) uint64 records."""
        p, mem, nbytes = self._buf_ptr(buf)
        if mem == capi.BVH_MEM_DEVICE:
            self.after_torch()
        cnt = C.c_uint64()
        call("bvh_unpack_results", self.handle, p, nbytes, mem, None, 0, C.byref(cnt))
        out = np.zeros((max(cnt.value, 1), 4), np.uint64)
        if cnt.value:
            host = buf.cpu().numpy() if mem == capi.BVH_MEM_DEVICE else buf
            hp, _, hn = self._buf_ptr(host)
            call("bvh_unpack_results", self.handle, hp, hn, capi.BVH_MEM_HOST, C.c_void_p(out.ctypes.data),
                 cnt.value, C.byref(cnt))
        return out[:cnt.value]

    def element_kdops(self, verts, epsilon=0.0):
        """bvh_element_kdops: the (n, 26) k-DOPs of (nv, 3, n) vertex planes
        (numpy: host memory; CUDA tensor: device memory, result on device)."""
        nv, _, n = tuple(verts.shape)
        if hasattr(verts, "data_ptr"):
            import torch
            _device_ptr(verts, np.float64, (None, 3, None), "vertices")
            out = torch.empty((n, 26), dtype=torch.float64, device=verts.device)
            self.after_torch()
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
            _device_ptr(ca, np.float64, (None, 3), "centroids of A")
            _device_ptr(cb, np.float64, (None, 3), "centroids of B")
            self.after_torch()
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
            _device_ptr(ka, np.float64, (None, 26), "k-DOPs of A")
            _device_ptr(kb, np.float64, (None, 26), "k-DOPs of B")
            if ga is not None:
                _device_ptr(ga, np.uint64, (ka.shape[0],), "global ids of A")
            if gb is not None:
                _device_ptr(gb, np.uint64, (kb.shape[0],), "global ids of B")
            self.after_torch()
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
    def _same_memory(self, mems):
        kinds = {m for m in mems if m is not None}
        if len(kinds) > 1:
            raise ValueError("element data, centroids and global ids must live in the same memory")
        if capi.BVH_MEM_DEVICE in kinds:
            self.world.after_torch()

    def set_entity_data_poly(self, verts, epsilon=0.0, global_ids=None, split_algorithm=CLUSTERING):
        """verts: (nv, 3, n) float64 SoA planes (numpy or CUDA tensor)."""
        shape = tuple(verts.shape)
        if len(shape) != 3 or shape[1] != 3:
            raise ValueError("vertices: shape %s, expected (nv, 3, n)" % (shape,))
        nv, _, n = shape
        p, mem, keep = _ptr_mem(verts, np.float64, (nv, 3, n), "vertices")
        g, gmem, gkeep = _ptr_mem(global_ids, np.uint64, (n,), "global ids")
        self._same_memory([mem, gmem if global_ids is not None else None])
        # the previous arrays stay referenced until the library has finished
        # the build that reads them
        call("bvh_set_entity_data_poly", self.handle, p, nv, n, g, float(epsilon), mem, split_algorithm)
        self._keep = (keep, gkeep)
        self._shapes = ((nv, 3, n),)
        self.n = n

    def set_entity_data_sphere(self, xyzr, global_ids=None, split_algorithm=CLUSTERING):
        """xyzr: (4, n) float64 planes x, y, z, r."""
        shape = tuple(xyzr.shape)
        if len(shape) != 2 or shape[0] != 4:
            raise ValueError("spheres: shape %s, expected (4, n)" % (shape,))
        n = shape[1]
        p, mem, keep = _ptr_mem(xyzr, np.float64, (4, n), "spheres")
        g, gmem, gkeep = _ptr_mem(global_ids, np.uint64, (n,), "global ids")
        self._same_memory([mem, gmem if global_ids is not None else None])
        call("bvh_set_entity_data_sphere", self.handle, p, n, g, mem, split_algorithm)
        self._keep = (keep, gkeep)
        self._shapes = ((4, n),)
        self.n = n

    def set_entity_data_snapshot(self, kdops, centroids, global_ids=None, split_algorithm=CLUSTERING):
        """kdops: (n, 26) float64, centroids: (n, 3) float64."""
        n = tuple(kdops.shape)[0]
        p, mem, keep = _ptr_mem(kdops, np.float64, (n, 26), "k-DOPs")
        c, cmem, ckeep = _ptr_mem(centroids, np.float64, (n, 3), "centroids")
        g, gmem, gkeep = _ptr_mem(global_ids, np.uint64, (n,), "global ids")
        self._same_memory([mem, cmem, gmem if global_ids is not None else None])
        call("bvh_set_entity_data_snapshot", self.handle, p, c, n, g, mem, split_algorithm)
        self._keep = (keep, ckeep, gkeep)
        self._shapes = ((n, 26),)
        self.n = n

    def refit(self, data):
        """New element data of the last set_entity_data's kind and count
        (same array shape); the tree keeps its topology and every k-DOP is
        recomputed (bvh_refit)."""
        want = getattr(self, "_shapes", (None,))[0]
        p, mem, keep = _ptr_mem(data, np.float64, want, "refit data")
        self._same_memory([mem])
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

    def self_collision_cross_patch(self):
        """bvh_self_collision_cross_patch: self pairs between different patches."""
        call("bvh_self_collision_cross_patch", self.handle)

    def set_patch_offsets(self, offsets):
        """bvh_object_set_patch_offsets: od - 1 input-order offsets for
        set_entity_data(..., split_algorithm=GIVEN)."""
        off = np.ascontiguousarray(np.asarray(offsets, np.uint64).reshape(-1))
        call("bvh_object_set_patch_offsets", self.handle, C.c_void_p(off.ctypes.data), len(off))

    # ---- wire formats -------------------------------------------------------------
    def _packed(self, fn, args, device):
        nb = C.c_uint64()
        call(fn, self.handle, *args, None, 0, C.byref(nb), capi.BVH_MEM_HOST)
        if device:
            import torch
            out = torch.empty(nb.value, dtype=torch.uint8, device="cuda")
            call(fn, self.handle, *args, C.c_void_p(out.data_ptr()), nb.value, C.byref(nb), capi.BVH_MEM_DEVICE)
            return out
        out = np.zeros(nb.value, np.uint8)
        call(fn, self.handle, *args, C.c_void_p(out.ctypes.data), nb.value, C.byref(nb), capi.BVH_MEM_HOST)
        return out

    def pack_patches(self, device=False):
        """bvh_pack_patches: od 256-byte patch records (numpy uint8, or a CUDA
        tensor with device=True)."""
        return self._packed("bvh_pack_patches", (), device)

    def element_record_bytes(self):
        v = C.c_uint64()
        call("bvh_element_record_bytes", self.handle, C.byref(v))
        return v.value

    def pack_patch_elements(self, patches, device=False):
        """bvh_pack_patch_elements: the element records of the listed patches."""
        ids = np.ascontiguousarray(np.asarray(patches, np.uint64).reshape(-1))
        return self._packed("bvh_pack_patch_elements", (C.c_void_p(ids.ctypes.data), len(ids)), device)

    def pack_results(self, device=False):
        """bvh_pack_results: the result message (stride, byte count, records)."""
        return self._packed("bvh_pack_results", (), device)

    def filter_exact(self):
        """bvh_result_filter_exact: keep the result pairs whose triangles or
        spheres intersect (closed); returns the number kept."""
        k = C.c_uint64()
        call("bvh_result_filter_exact", self.handle, C.byref(k))
        return k.value

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
