"""ctypes declarations of the C ABI in include/distbvh.h.

The library is distbvh_amd/libdistbvh.so, built in-tree by
distbvh_amd/build.py.  There is no fallback: if the library is missing, or a
call fails, an exception is raised.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdistbvh.so")

BVH_OK = 0
BVH_E_INVALID = 1
BVH_E_OVERFLOW = 3
BVH_SPLIT_GEOM_AXIS, BVH_SPLIT_ML_GEOM_AXIS, BVH_SPLIT_CLUSTERING, BVH_SPLIT_GIVEN = 0, 1, 2, 3
BVH_MEM_HOST, BVH_MEM_DEVICE = 0, 1
BVH_PREC_FP64, BVH_PREC_FP32_OUTWARD, BVH_PREC_FP64_RECORDS = 0, 1, 2
BVH_FLAG_SORT_RESULTS = 0x1
BVH_FLAG_TIMING = 0x2
BVH_FLAG_SERIAL_BUILDS = 0x4
BVH_ELEM_POLY, BVH_ELEM_SPHERE, BVH_ELEM_SNAPSHOT = 0, 1, 2


class BvhError(RuntimeError):
    def __init__(self, func, status, text):
        super().__init__("%s failed: %s (%s)" % (func, text, status))
        self.status = status


class WorldConfig(C.Structure):
    _fields_ = [("morton_bits", C.c_uint32), ("leaf_size", C.c_uint32), ("precision", C.c_uint32),
                ("flags", C.c_uint32), ("stream", C.c_void_p), ("device", C.c_int32),
                ("reserved", C.c_uint32 * 7)]


class Record(C.Structure):
    _fields_ = [("patch_p", C.c_uint64), ("element_p", C.c_uint64), ("patch_q", C.c_uint64),
                ("element_q", C.c_uint64)]


class TreeNode(C.Structure):
    _fields_ = [("parent_offset", C.c_int64), ("child_offsets", C.c_int64 * 2), ("kdop", C.c_double * 26),
                ("entity_range", C.c_uint64 * 2)]


class Tuning(C.Structure):
    _fields_ = [("batch", C.c_uint32), ("block_size", C.c_uint32), ("xcd_remap", C.c_uint32),
                ("tried", C.c_uint32), ("ms", C.c_double), ("tried_ms", C.c_double * 24)]


class PatchInfo(C.Structure):
    _fields_ = [("global_id", C.c_uint64), ("size", C.c_uint64), ("kdop", C.c_double * 26),
                ("centroid", C.c_double * 3)]


_vp = C.c_void_p
_sz = C.c_size_t
_u32 = C.c_uint32
_u64 = C.c_uint64
_d = C.c_double
_st = C.c_int

# name -> argument types (all return bvh_status unless listed in _RESTYPES)
SIGNATURES = {
    "bvh_status_string": [_st],
    "bvh_last_error": [],
    "bvh_abi_version": [],
    "bvh_world_config_init": [C.POINTER(WorldConfig)],
    "bvh_world_create": [_sz, C.POINTER(WorldConfig), C.POINTER(_vp)],
    "bvh_world_destroy": [_vp],
    "bvh_world_overdecomposition": [_vp, C.POINTER(_sz)],
    "bvh_create_collision_object": [_vp, C.POINTER(_vp)],
    "bvh_num_collision_objects": [_vp, C.POINTER(_sz)],
    "bvh_start_iteration": [_vp],
    "bvh_finish_iteration": [_vp],
    "bvh_synchronize": [_vp],
    "bvh_world_stream": [_vp, C.POINTER(_vp)],
    "bvh_stage_times": [_vp, C.POINTER(C.c_char_p), C.POINTER(_d), _sz, C.POINTER(_sz)],
    "bvh_object_id": [_vp, C.POINTER(_sz)],
    "bvh_set_entity_data_poly": [_vp, _vp, _u32, _u64, _vp, _d, _u32, _u32],
    "bvh_set_entity_data_sphere": [_vp, _vp, _u64, _vp, _u32, _u32],
    "bvh_set_entity_data_snapshot": [_vp, _vp, _vp, _u64, _vp, _u32, _u32],
    "bvh_refit": [_vp, _vp, _u32],
    "bvh_object_set_patch_offsets": [_vp, _vp, _sz],
    "bvh_object_device_bytes": [_vp, C.POINTER(_u64)],
    "bvh_object_root_kdop": [_vp, _vp],
    "bvh_autotune": [_vp, _vp, _vp],
    "bvh_set_launch": [_vp, _vp],
    "bvh_get_launch": [_vp, _vp],
    "bvh_capture_begin": [_vp],
    "bvh_capture_end": [_vp, _vp],
    "bvh_graph_launch": [_vp],
    "bvh_graph_destroy": [_vp],
    "bvh_init_broadphase": [_vp],
    "bvh_broadphase": [_vp, _vp],
    "bvh_broadphase_patches": [_vp, _vp],
    "bvh_broadphase_batch": [_vp, _vp, _u64],
    "bvh_self_collision": [_vp],
    "bvh_self_collision_cross_patch": [_vp],
    "bvh_end_phase": [_vp],
    "bvh_result_count": [_vp, C.POINTER(_u64)],
    "bvh_result_records": [_vp, _vp, _u64, _u32],
    "bvh_result_pairs": [_vp, _vp, _u64, _u32],
    "bvh_result_digest": [_vp, _vp],
    "bvh_result_pairs_device": [_vp, C.POINTER(_vp), C.POINTER(_u64)],
    "bvh_result_stats": [_vp, C.POINTER(_u64), C.POINTER(_u64), C.POINTER(_u64), C.POINTER(_u64), C.POINTER(_d)],
    "bvh_patch_pair_count": [_vp, C.POINTER(_u64)],
    "bvh_patch_pairs": [_vp, _vp, _u64],
    "bvh_local_patches": [_vp, _vp, _sz],
    "bvh_object_overdecomposition": [_vp, C.POINTER(_sz)],
    "bvh_patch_elements": [_vp, _sz, _vp, _u64, C.POINTER(_u64)],
    "bvh_object_bounds": [_vp, _vp, _vp],
    "bvh_object_sorted": [_vp, _vp, _vp, _u32],
    "bvh_object_splits": [_vp, _vp, _sz],
    "bvh_object_sorted_kdops": [_vp, _vp, _u32],
    "bvh_object_tree": [_vp, _vp, _vp, _vp, _sz, C.POINTER(_sz), C.POINTER(_sz)],
    "bvh_sort_u64": [_vp, _vp, _vp, _u64],
    "bvh_sort_u64_noalloc": [_vp, _vp, _vp, _u64],
    "bvh_sort_u32": [_vp, _vp, _vp, _u64],
    "bvh_exclusive_scan_u32": [_vp, _vp, _u64],
    "bvh_closest_point": [_vp, _vp, _u64, _vp, _vp, _u32],
    "bvh_triangle_pairs_intersect": [_vp, _vp, _vp, _u64, _u32, _vp],
    "bvh_sphere_pairs_intersect": [_vp, _vp, _vp, _u64, _u32, _vp],
    "bvh_result_filter_exact": [_vp, C.POINTER(_u64)],
    "bvh_pack_patches": [_vp, _vp, _u64, C.POINTER(_u64), _u32],
    "bvh_unpack_patches": [_vp, _vp, _u64, _u32, _vp, _u64, C.POINTER(_u64)],
    "bvh_element_record_bytes": [_vp, C.POINTER(_u64)],
    "bvh_pack_patch_elements": [_vp, _vp, _u64, _vp, _u64, C.POINTER(_u64), _u32],
    "bvh_unpack_elements": [_vp, _vp, _u64, _u32, _u32, _vp, _vp, _vp, _u32],
    "bvh_pack_results": [_vp, _vp, _u64, C.POINTER(_u64), _u32],
    "bvh_unpack_results": [_vp, _vp, _u64, _u32, _vp, _u64, C.POINTER(_u64)],
    "bvh_object_release": [_vp],
    "bvh_element_kdops": [_vp, _vp, _u32, _u64, C.c_double, _u32, _vp],
    "bvh_max_variant_axis": [_vp, _vp, _u64, _vp, _u64, _u32, _vp],
    "bvh_sort_and_sweep": [_vp, _vp, _vp, _vp, _u64, _vp, _vp, _vp, _u64, C.c_int32, _u32, _vp, _u64, _vp],
    "bvh_object_patch_tree": [_vp, _vp, C.c_size_t, _vp, _vp, C.c_size_t, _vp],
}
_RESTYPES = {"bvh_status_string": C.c_char_p, "bvh_last_error": C.c_char_p,
             "bvh_abi_version": C.c_uint32, "bvh_world_config_init": None}

_lib = None


def load():
    """Load libdistbvh.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("distbvh_amd/libdistbvh.so is missing: run __graft_entry__.build() "
                               "or python distbvh_amd/build.py")
        L = C.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = _RESTYPES.get(name, _st)
        _lib = L
    return _lib


def call(name, *args):
    """Call a bvh_* entry point and raise BvhError on a non-OK status."""
    L = load()
    st = getattr(L, name)(*args)
    if st != BVH_OK:
        text = L.bvh_status_string(st).decode()
        detail = L.bvh_last_error().decode()
        raise BvhError(name, st, text + (": " + detail if detail else ""))
    return st
