"""Shared helpers of the test suite: run a workload through the GPU library
and through the CPU oracle, and compare."""
import numpy as np

import oracle as O  # oracle/oracle.py (checker)
from distbvh_amd import workloads as W


# the three tree forms: fp32-outward records (the default fp64 form, which
# also exports fp64 element k-DOPs, and the fp32-outward form), fp64 slabs in
# every record
FORMS = ["fp64", "fp32", "fp64_records"]


def to_cuda(a):
    import torch
    a = np.ascontiguousarray(a)
    if a.dtype == np.uint64:
        return torch.from_numpy(a.view(np.int64)).cuda()
    if a.dtype == np.uint32:
        return torch.from_numpy(a.view(np.int32)).cuda()
    return torch.from_numpy(a).cuda()


def from_cuda(t, dtype):
    return t.cpu().numpy().view(dtype)


def oracle_elements(kind, data, eps=0.0):
    if kind == "poly":
        return O.elements_poly(data, eps)
    if kind == "sphere":
        return O.elements_sphere(data)
    raise ValueError(kind)


def gpu_object(world, kind, data, eps=0.0, device=False, gids=None):
    obj = world.create_collision_object()
    arr = to_cuda(data) if device else data
    g = None if gids is None else (to_cuda(gids) if device else gids)
    if kind == "poly":
        obj.set_entity_data_poly(arr, epsilon=eps, global_ids=g)
    elif kind == "sphere":
        obj.set_entity_data_sphere(arr, global_ids=g)
    else:
        raise ValueError(kind)
    obj.init_broadphase()
    return obj


def sort_records(r):
    r = np.asarray(r, dtype=np.uint64).reshape(-1, 4)
    if len(r) == 0:
        return r
    order = np.lexsort((r[:, 2], r[:, 0], r[:, 3], r[:, 1]))
    return r[order]


def sort_pairs(p):
    p = np.asarray(p, dtype=np.uint64).reshape(-1, 2)
    if len(p) == 0:
        return p
    order = np.lexsort((p[:, 1], p[:, 0]))
    return p[order]


def mix64(z):
    z = np.asarray(z, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xbf58476d1ce4e5b9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94d049bb133111eb)
        return z ^ (z >> np.uint64(31))


def pair_digest(pairs):
    """Host restatement of bvh_result_digest over (k, 2) index pairs: the sum
    and the xor of two 64-bit mixes of key = a << 32 | b."""
    p = np.asarray(pairs, dtype=np.uint64).reshape(-1, 2)
    keys = (p[:, 0] << np.uint64(32)) | p[:, 1]
    with np.errstate(over="ignore"):
        s = int(mix64(keys).sum(dtype=np.uint64)) if len(keys) else 0
    x = int(np.bitwise_xor.reduce(mix64(keys ^ np.uint64(0x9e3779b97f4a7c15)))) if len(keys) else 0
    return s, x
