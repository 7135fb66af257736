"""Benchmark: build + broadphase + narrowphase of two 2M-triangle objects.

Workload (one step): two deformed contact sheets, 1024 x 1024 quads =
2,097,152 triangles each, fp64 vertices resident in HBM.  Sheet A:
z = 0.01 sin(2 pi 8 x) cos(2 pi 8 y); sheet B = A + (0, 0, 0.005) + seeded
N(0, 0.002^2) vertex noise.  Per step both objects are rebuilt
(26-DOP bounds, 63-bit Morton codes, radix sort, cluster splits into 256
patches, bottom-up tree with refit, leaf size 4) and A is collided with B
(patch broadphase + element narrowphase); the step ends when the pair count is
on the host.

Prints one JSON line (see the contract in the repository's README of the
benchmark driver): value = ms per step (lower is better), the roofline of the
dominant kernel from HIP events on the library's stream, and the CPU oracle
timed on this host's cores as a baseline.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def kernel_bytes(stage, n, npairs, n_other, fp32=False):
    """Algorithmic (compulsory) HBM bytes of one launch of a stage for an
    object of n triangles (3 fp64 vertices = 72 B).  Node slabs are 13
    (min, max) pairs: 208 B in fp64, 104 B in the fp32 form."""
    SLAB_BYTES = 104 if fp32 else 208
    if stage == "bounds":      # read vertices, write centroid
        return n * (72 + 24)
    if stage == "morton":      # read centroid, write code
        return n * (24 + 8)
    if stage == "sort":        # 8 passes x (read + write 8 B key + 4 B index)
        return n * 8 * 24
    if stage == "cluster":     # depth keys (read codes, write u32), one 8-bit pass
        return n * (8 + 4 + 16)
    if stage == "delta":       # read codes, write patch id + split height
        return n * (8 + 4 + 1 + 4)
    if stage == "tree":        # leaves: gather vertices, write slabs; internal:
        # read two children's slabs, write one, links + counters
        return n * (4 + 4 + 72 + SLAB_BYTES + 4) + (n - 1) * (2 * SLAB_BYTES + SLAB_BYTES + 12)
    if stage == "ropes":
        return (2 * n - 1) * 12
    if stage == "narrowphase":  # query slabs once, other tree once, pairs
        return n * (SLAB_BYTES + 4) + (2 * n_other - 1) * (SLAB_BYTES + 16) + 8 * npairs
    if stage == "broadphase":
        return 0
    return 0


def cpu_baseline(a, b, od, leaf):
    """The CPU oracle (restated reference path) on the full workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import build as oracle_build
    import oracle as O

    if not os.path.exists(O.LIB_PATH):
        oracle_build.build()
    try:
        nthreads = min(len(os.sched_getaffinity(0)), 16)
    except AttributeError:
        nthreads = min(os.cpu_count() or 1, 16)
    t0 = time.perf_counter()
    ka, ca = O.elements_poly(a)
    kb, cb = O.elements_poly(b)
    oa = O.Object(ka, ca, od=od, morton_bits=63, leaf_size=leaf, nthreads=nthreads)
    ob = O.Object(kb, cb, od=od, morton_bits=63, leaf_size=leaf, nthreads=nthreads)
    recs = oa.collide(ob)
    t1 = time.perf_counter()
    return {
        "value": round((t1 - t0) * 1e3, 3),
        "unit": "ms/step",
        "cores": nthreads,
        "kind": "port",
        "sample": "one full step of the same workload (2 x 2,097,152 triangles); element k-DOPs "
                  "and the per-patch-pair narrowphase on %d threads, bounds / Morton / 1-bit radix "
                  "sort / cluster / patches serial as in the reference" % nthreads,
        "pairs": int(len(recs)),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nq", type=int, default=1024, help="quads per sheet side")
    ap.add_argument("--od", type=int, default=256)
    ap.add_argument("--leaf", type=int, default=4)
    ap.add_argument("--precision", choices=["fp64", "fp32"], default="fp64",
                    help="fp64 tree, or the fp32-outward tree with an exact fp64 re-check")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="first copy a MIB-sized device buffer once (the reference kernel for "
                         "scaling FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(the sharding design is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import torch
    from distbvh_amd import CollisionWorld
    from distbvh_amd import workloads as W

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    if args.calibrate_copy:
        src = torch.empty(args.calibrate_copy << 20, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    a, b = W.wavy_sheets(args.nq)
    n = a.shape[2]
    ta = torch.from_numpy(a).cuda()
    tb = torch.from_numpy(b).cuda()
    stream = torch.cuda.current_stream()
    fp32 = args.precision == "fp32"
    world = CollisionWorld(args.od, morton_bits=63, leaf_size=args.leaf, timing=True,
                           stream=stream.cuda_stream, precision=args.precision)
    oa = world.create_collision_object()
    ob = world.create_collision_object()

    def step():
        oa.set_entity_data_poly(ta)
        ob.set_entity_data_poly(tb)
        oa.init_broadphase()
        ob.init_broadphase()
        oa.broadphase(ob)
        return oa.result_count()

    for _ in range(args.warmup):
        step()
        world.stage_times()
    torch.cuda.synchronize()

    stage_sum = {}
    stage_cnt = {}
    total_ms = 0.0
    npairs = 0
    for _ in range(args.steps):
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        npairs = step()
        e1.record(stream)
        torch.cuda.synchronize()
        total_ms += e0.elapsed_time(e1)
        for name, ms in world.stage_times():
            stage_sum[name] = stage_sum.get(name, 0.0) + ms
            stage_cnt[name] = stage_cnt.get(name, 0) + 1
    ms_per_step = total_ms / args.steps
    qstats = oa.result_stats()
    waves = (n + 63) // 64
    stages = {k: stage_sum[k] / args.steps for k in stage_sum}           # ms per step
    per_launch = {k: stage_sum[k] / stage_cnt[k] for k in stage_sum}      # ms per launch

    # dominant kernel: largest share of the step
    dom = max(stages, key=lambda k: stages[k])
    nbytes = kernel_bytes(dom, n, npairs, n, fp32)
    t_launch = per_launch[dom] * 1e-3
    achieved = nbytes / t_launch / 1e9 if t_launch > 0 else 0.0
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "algorithmic_bytes": int(nbytes)}
    step_bytes = sum(kernel_bytes(k, n, npairs, n, fp32) * stage_cnt[k] / args.steps for k in stages)

    out = {
        "metric": "ms/step build+broad+narrow @2M tris/object, %HBM peak; batched p99 latency",
        "value": round(ms_per_step, 4),
        "unit": "ms/step",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if fp32 else "f64",
        "data": "synthetic",
        "config": {"workload": "wavy contact sheets, 2 objects x %d triangles (%dx%d quads), fp64, "
                               "26-DOP, 63-bit Morton, %d patches, leaf size %d" % (n, args.nq, args.nq, args.od, args.leaf),
                   "precision": "fp32 tree, exact fp64 re-check" if fp32 else "fp64",
                   "candidate_pairs": int(qstats["candidate_pairs"]),
                   "rechecked_pairs": int(qstats["rechecked_pairs"]),
                   "triangles_per_object": n, "od": args.od, "leaf_size": args.leaf, "pairs": int(npairs),
                   "patch_pairs": int(qstats["patch_pairs"]),
                   "node_visits_per_wave": round(qstats["node_visits"] / waves, 1),
                   "parallelism": "single GPU"},
        "stages_ms": {k: round(v, 4) for k, v in stages.items()},
        "step_hbm_frac": round(step_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "roofline": roofline,
    }
    world.close()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(a, b, args.od, args.leaf)
        cpu_pairs = out["cpu_baseline"].pop("pairs")
        out["config"]["cpu_pairs_equal_count"] = bool(cpu_pairs == npairs)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
