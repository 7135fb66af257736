"""Writes the golden fixtures of the test suite.

1. reference_known_answers.json -- the known-answer vectors that distBVH's
   own unit tests hold (inputs and expected outputs, with the test file and
   lines they come from).  These are data transcribed from the reference's
   test sources; nothing of the reference is run.
2. flat_sheets_oracle.json -- digests of the CPU oracle's outputs on the
   8,192-triangle flat-sheet pair (sorted Morton codes, permutation, cluster
   splits, pair records), a regression guard on the oracle itself.

Run from the repository root:  python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))

KNOWN = {
    "morton32": {
        "source": "tests/hash_test.cpp:43-101",
        "cases": [
            {"xyz": [0x3ff, 0x3ff, 0x3ff], "code": 0x3fffffff},
            {"xyz": [0x3ff, 0, 0], "code": 0x09249249},
            {"xyz": [0, 0x3ff, 0], "code": 0x12492492},
            {"xyz": [0, 0, 0x3ff], "code": 0x24924924},
        ],
    },
    "morton64": {
        "source": "tests/hash_test.cpp:103-160",
        "cases": [
            {"xyz": [0x1fffff, 0x1fffff, 0x1fffff], "code": 0x7fffffffffffffff},
            {"xyz": [0x1fffff, 0, 0], "code": 0x1249249249249249},
            {"xyz": [0, 0x1fffff, 0], "code": 0x2492492492492492},
            {"xyz": [0, 0, 0x1fffff], "code": 0x4924924924924924},
        ],
    },
    "quantize32_bounds": {
        "source": "tests/hash_test.cpp:163-178",
        "min": [-1.0, 0.0, 3.0], "max": [-0.5, 2.0, 4.0], "almost": 0.00001,
        "q_min": [0, 0, 0], "q_almost_max": [0x3ff, 0x3ff, 0x3ff],
    },
    "radix_sort": {
        "source": "tests/sort_test.cpp:9-30",
        "keys": [3, 7, 2, 1, 4, 9, 1, 3],
        "sorted": [1, 1, 2, 3, 3, 4, 7, 9],
        "perm": [3, 6, 2, 0, 7, 4, 1, 5],
    },
    "prefix_sum": {
        "source": "tests/UtilTest.cpp:42-55",
        "input": [3, 7, 2, 1, 4, 9, 1, 3],
        "output": [0, 3, 10, 12, 13, 17, 26, 27],
    },
    "clz": {
        "source": "tests/bits_test.cpp:39-57",
        "clz32_zero": 32, "clz64_zero": 64,
        "rule": "clz(1 << i) == width - 1 - i",
    },
    "kdop_vertices": {
        "source": "tests/KDopTest.cpp:81-152",
        "cases": [
            {"verts": [[1.0, 1.0, 1.0]], "cardinal": [[1.0, 1.0], [1.0, 1.0], [1.0, 1.0]], "longest_length": 0.0},
            {"verts": [[1.0, 1.0, 1.0], [1.0, 5.0, 1.0]], "longest": 1, "longest_length": 4.0},
            {"verts": [[1.0, 1.0, 1.0], [4.0, 5.0, 4.0]], "longest": 3, "longest_length": "10/sqrt(3)"},
            {"verts": [[-5.0, -1.0, 1.0], [1.0, 5.0, 1.0]], "longest": 7, "longest_length": "12/sqrt(2)"},
            {"verts": [[0.0, 0.0, 0.0], [1.0, 0.0, 0.0]], "longest": 0, "longest_length": 1.0,
             "zero_length_axes": [1, 2]},
        ],
    },
    "project": {
        "source": "tests/KDopTest.cpp:189-239",
        "cases": [
            {"v": [5.0, 0.9, 11.0], "axis": 0, "value": 5.0},
            {"v": [0.2, 5.0, 11.0], "axis": 1, "value": 5.0},
            {"v": [0.2, 13.0, 5.0], "axis": 2, "value": 5.0},
            {"v": [1.0, 1.0, 1.0], "axis": 3, "value": "3/sqrt(3)"},
            {"v": [-1.0, 0.0, 0.0], "axis": 3, "value": "-1/sqrt(3)"},
            {"v": [5.0, 0.9, 11.0], "axis": 3, "value": "16.9/sqrt(3)"},
            {"v": [1.0, 1.0, 1.0], "axis": 17, "value": "-1/sqrt(3)"},
            {"v": [-1.0, 0.0, 0.0], "axis": 17, "value": "1/sqrt(3)"},
            {"v": [5.0, 0.9, 11.0], "axis": 17, "value": "-15.1/sqrt(3)"},
        ],
    },
    "cube_overlap": {
        "source": "tests/KDopTest.cpp:241-302 (cube = center +- 1 on every axis)",
        "cases": [
            {"center": [3.0, 0.0, 0.0], "overlap": False},
            {"center": ["2-eps", 0.0, 0.0], "overlap": True},
            {"center": [0.0, 0.0, 0.0], "overlap": True},
            {"center": ["-2+eps", 0.0, 0.0], "overlap": True},
            {"center": [-3.0, 0.0, 0.0], "overlap": False},
            {"center": ["2-eps", "2-eps", 0.0], "overlap": True},
            {"center": ["-2+eps", "-2+eps", 0.0], "overlap": True},
            {"center": ["2-eps", "2-eps", "2-eps"], "overlap": True},
            {"center": ["-2+eps", "-2+eps", "-2+eps"], "overlap": True},
        ],
    },
    "sphere_overlap": {
        "source": "tests/KDopTest.cpp:304-315",
        "cases": [
            {"center": [3.0, 0.0, 0.0], "overlap": False},
            {"center": ["2-eps", 0.0, 0.0], "overlap": True},
        ],
    },
    "snapshot_morton_grid": {
        "source": "tests/snapshot_test.cpp:31-72",
        "bound_min": [-100.0, -100.0, -100.0], "bound_max": [100.0, 100.0, 100.0],
        "inv_diag": [0.005, 0.005, 0.005], "grid": 8, "step": 0x80,
    },
    "collision_counts": {
        "source": "tests/CollisionTest.cpp:43-81",
        "potential_4x4x4_vs_unit_cube": 64,
        "self_2x2x1": 0,
        "self_2x2x2_plus_unit_cube": 8,
    },
    "sort_and_sweep": {
        "source": "tests/KokkosTest.cpp:87-139 (disabled in the reference build; arithmetic valid)",
        "grid_12x1x1_vs_unit_cube_hits": 12,
    },
    "max_variant_axis": {
        "source": "tests/KokkosTest.cpp:37-85 (disabled in the reference build; arithmetic valid)",
        "grids_vs_unit_cube": [[[12, 1, 1], 0], [[1, 12, 1], 1], [[1, 1, 12], 2]],
    },
    "narrowphase_functor": {
        "source": "tests/collision_object_test.cpp:376-443 (one rank)",
        "od": 2, "object1": [1, 1, 1], "object2": [2, 3, 2], "results": 12,
    },
}


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def oracle_digests():
    sys.path.insert(0, ROOT)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from distbvh_amd import workloads as W

    a, b = W.flat_sheets(64)
    out = {"workload": "flat sheets 64x64 quads (8,192 triangles / object), seed 0xC1"}
    for eps in (0.0, 0.01):
        ka, ca = O.elements_poly(a, eps)
        kb, cb = O.elements_poly(b, eps)
        for bits in (63, 30):
            ra = O.Object(ka, ca, od=16, morton_bits=bits)
            rb = O.Object(kb, cb, od=16, morton_bits=bits)
            rec = ra.collide(rb)
            key = "eps%g_bits%d" % (eps, bits)
            out[key] = {
                "codes_a": digest(ra.codes), "perm_a": digest(ra.perm), "splits_a": ra.splits.tolist(),
                "codes_b": digest(rb.codes), "perm_b": digest(rb.perm), "splits_b": rb.splits.tolist(),
                "records": digest(rec), "nrecords": int(len(rec)),
            }
    return out


if __name__ == "__main__":
    with open(os.path.join(HERE, "reference_known_answers.json"), "w") as f:
        json.dump(KNOWN, f, indent=1)
    with open(os.path.join(HERE, "flat_sheets_oracle.json"), "w") as f:
        json.dump(oracle_digests(), f, indent=1)
    print("wrote golden fixtures")
