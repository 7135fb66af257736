"""The C ABI library loads, exports every entry point include/distbvh.h
declares, and rejects bad arguments before touching a GPU.  CPU only."""
import ctypes as C
import os
import re

import pytest

from distbvh_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "distbvh.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bvh_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = capi.load()
    names = declared_functions()
    assert len(names) >= 35
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # and the ctypes table covers them all
    assert set(names) == set(capi.SIGNATURES), set(names) ^ set(capi.SIGNATURES)


def test_abi_version_and_strings():
    L = capi.load()
    assert L.bvh_abi_version() == 1
    assert L.bvh_status_string(0) == b"ok"
    assert L.bvh_status_string(3) == b"buffer too small"


def _cfg(**kw):
    cfg = capi.WorldConfig()
    capi.load().bvh_world_config_init(C.byref(cfg))
    for k, v in kw.items():
        setattr(cfg, k, v)
    return cfg


def test_config_defaults():
    cfg = _cfg()
    assert cfg.morton_bits == 63 and cfg.leaf_size == 1 and cfg.precision == 0 and cfg.stream is None


@pytest.mark.parametrize("od,kw", [(3, {}), (0, {}), (1 << 15, {}), (2, {"morton_bits": 32}),
                                   (2, {"leaf_size": 0}), (2, {"leaf_size": 9}), (2, {"precision": 7})])
def test_world_create_rejects_bad_arguments(od, kw):
    # od must be a power of two (collision_world.cpp:59-65); options checked
    h = C.c_void_p()
    st = capi.load().bvh_world_create(od, C.byref(_cfg(**kw)), C.byref(h))
    assert st == 1 and h.value is None
    assert capi.load().bvh_last_error()


def test_null_arguments():
    L = capi.load()
    assert L.bvh_world_create(2, None, None) == 1
    assert L.bvh_init_broadphase(None) == 1
    assert L.bvh_broadphase(None, None) == 1
    n = C.c_uint64()
    assert L.bvh_result_count(None, C.byref(n)) == 1


def test_python_binding_raises():
    with pytest.raises(capi.BvhError):
        capi.call("bvh_world_create", 3, None, C.byref(C.c_void_p()))
