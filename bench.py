"""Benchmark: build + broadphase + narrowphase of two 2M-triangle objects.

Workload (one step): two deformed contact sheets, 1024 x 1024 quads =
2,097,152 triangles each, fp64 vertices resident in HBM.  Sheet A:
z = 0.01 sin(2 pi 8 x) cos(2 pi 8 y); sheet B = A + (0, 0, 0.005) + seeded
N(0, 0.002^2) vertex noise.  Per step both objects are rebuilt
(26-DOP bounds, 63-bit Morton codes, radix sort, cluster splits into 256
patches, bottom-up tree with refit; one element per leaf, the CPU oracle's
trees use the workload's leaf size 4) and A is collided with B
(patch broadphase + element narrowphase); the step ends when the pair count is
on the host.

Prints one JSON line: value = ms per step (lower is better), the roofline of
the dominant kernel from HIP events on the library's stream, and the CPU
oracle timed on this host's cores as a baseline.  --precision both adds the
fp32-outward form's time and whether its pairs equal the fp64 form's (count
and an order-independent digest).  --workload spheres-self times self
collision of a 16,777,216-sphere cloud instead.
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


def kernel_bytes(stage, n, npairs, n_other, fp32=False, elem_bytes=72):
    """Algorithmic (compulsory) HBM bytes of one launch of a stage for an
    object of n elements (a triangle: 3 fp64 vertices = 72 B).  A node's 13
    (min, max) slab pairs take 208 B in fp64 and 104 B in fp32; the tree is
    one record per internal node holding both children's slabs plus links
    (448 B with fp64 slabs, 256 B with fp32-outward slabs: the fp64 and fp32
    forms; `fp32` names the record format)."""
    slab = 104 if fp32 else 208
    rec = 256 if fp32 else 448
    if stage == "bounds":      # read elements, write centroid
        return n * (elem_bytes + 24)
    if stage == "morton":      # read centroid, write code
        return n * (24 + 8)
    if stage == "sort":        # 8 passes x (read + write 8 B key + 4 B index)
        return n * 8 * 24
    if stage == "cluster":     # depth keys (read codes, write u32), one 8-bit pass
        return n * (8 + 4 + 16)
    if stage == "delta":       # read codes, write patch id + split height
        return n * (8 + 4 + 1 + 4)
    if stage == "tree":        # gather elements + permutation + patch ids + split
        # heights, write every record (each node's slabs and links sit in its
        # parent's record)
        return n * (elem_bytes + 4 + 4 + 1) + max(n - 1, 0) * rec
    if stage == "narrowphase":  # query slabs (from their parents' records), the
        # other object's records once, pair keys
        return n * (slab + 4 + 2) + (n_other - 1) * rec + 8 * npairs
    if stage == "broadphase":
        return 0
    return 0


# stage -> the kernel whose measured HBM traffic the roofline reports
STAGE_KERNEL = {"tree": ("k_tree_leaves<0, 3, false>", "k_tree_leaves<0, 3, true>"),
                "sort": ("k_sort_pass<unsigned long>", "k_sort_pass<unsigned long>")}


def stage_kernel(stage, fp32, launch):
    """The stage's main kernel as rocprofv3 names it (`fp32`: fp32-outward
    records, the fp64 and fp32 forms; the walk over fp64 records depends on
    the launch batch: one record in scalar registers, or K staged in LDS)."""
    if stage == "narrowphase":
        if fp32:
            return "k_narrow32<false, false>"
        b = launch.get("batch", 1) if launch else 1
        return "k_narrow64s<false, false>" if b == 1 else "k_narrow64<false, %d, false>" % b
    return STAGE_KERNEL.get(stage, (None, None))[1 if fp32 else 0]


def measured_traffic(stage, workload, form, launch=None):
    """HBM bytes per launch of the stage's main kernel from the rocprofv3
    FETCH_SIZE / WRITE_SIZE table of the form under profiles/ (None if
    absent)."""
    name = stage_kernel(stage, form != "fp64_records", launch)
    path = os.path.join(ROOT, "profiles", "traffic_%s_%s.txt" % (workload, form))
    if name is None or not os.path.exists(path):
        return None
    for line in open(path):
        f = line.rstrip("\n").split("\t")
        if len(f) == 3 and f[0] == name:
            return int(f[1]) + int(f[2])
    return None


def cpu_baseline(a, b, od, leaf, runs=5, gpu_records=None):
    """The CPU oracle (the reference path restated in C) on the full
    workload, timed as the median of `runs` whole steps (element k-DOPs,
    bounds, Morton codes, 1-bit LSD radix sort, cluster splits, patches,
    patch tree, per-patch-pair narrowphase) on all the cores this process may
    use (at most 16, the host share of one GPU) and on one thread.  The
    records of the first run are compared with the GPU's, row for row."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import build as oracle_build
    import oracle as O

    if not os.path.exists(O.LIB_PATH):
        oracle_build.build()
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count() or 1
    nall = max(1, min(affinity, 16))

    def step(nthreads):
        t0 = time.perf_counter()
        ka, ca = O.elements_poly(a)
        kb, cb = O.elements_poly(b)
        oa = O.Object(ka, ca, od=od, morton_bits=63, leaf_size=leaf, nthreads=nthreads)
        ob = O.Object(kb, cb, od=od, morton_bits=63, leaf_size=leaf, nthreads=nthreads)
        recs = oa.collide(ob)
        return (time.perf_counter() - t0) * 1e3, recs

    times_all, times_one = [], []
    recs = None
    for r in range(runs):
        ms, rr = step(nall)
        times_all.append(ms)
        if recs is None:
            recs = rr
    for r in range(runs):
        times_one.append(step(1)[0])
    equal = None
    if gpu_records is not None:
        g = np.asarray(gpu_records, np.uint64).reshape(-1, 4)
        g = g[np.lexsort((g[:, 2], g[:, 0], g[:, 3], g[:, 1]))]
        c = recs[np.lexsort((recs[:, 2], recs[:, 0], recs[:, 3], recs[:, 1]))]
        equal = bool(g.shape == c.shape and np.array_equal(g, c))
    med_all = float(np.median(times_all))
    med_one = float(np.median(times_one))
    return {
        "value": round(med_all, 3),
        "unit": "ms/step",
        "cores": nall,
        "kind": "port",
        "sample": "the full step of the same workload (2 x 2,097,152 triangles), median of %d runs on %d "
                  "threads (element k-DOPs and the per-patch-pair narrowphase in parallel; bounds, Morton, "
                  "1-bit radix sort, cluster splits and patches serial as in the reference)" % (runs, nall),
        "single_thread_ms": round(med_one, 3),
        "all_core_ms": round(med_all, 3),
        "runs_ms": {"all_core": [round(t, 1) for t in times_all], "single_thread": [round(t, 1) for t in times_one]},
        "host_cpus": os.cpu_count(),
        "affinity_cpus": affinity,
        "records": int(len(recs)),
        "records_equal_gpu": equal,
    }


def serving_leg(args, check_oracle=True):
    """The serving form: an open-loop Poisson stream of mesh-pair requests
    (log-uniform 10k-2M triangles per object) served by continuous batching;
    p50 / p99 latency from arrival to pairs on the host.  Every distinct
    request (mesh pair, motion variant) of the stream is checked against the
    CPU oracle (the restated reference pipeline, fp64) by pair count and
    digest."""
    from distbvh_amd import serve as S

    prec = "fp64" if args.precision == "both" else args.precision
    r = S.serve(rate=args.rate, duration=args.duration, precision=prec, max_batch=args.max_batch, record=True)
    out = {k: r[k] for k in ("requests", "rate_req_s", "max_batch", "refits", "rebuilds", "evictions", "deferred",
                             "steps", "mean_step_requests", "max_step_requests", "mesh_triangles")}
    out.update({"p50_ms": round(r["p50_ms"], 3), "p99_ms": round(r["p99_ms"], 3), "mean_ms": round(r["mean_ms"], 3),
                "max_ms": round(r["max_ms"], 3), "pairs_per_s": round(r["pairs_per_s"], 1), "pairs": r["pairs"],
                "wall_s": round(r["wall_s"], 3), "precision": prec})
    if check_oracle:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O

        pool = S.mesh_pool(8, 1e4, 2e6, 0xC5)
        ref = {}
        equal = 0
        for i, (m, (cnt, dig)) in enumerate(zip(r["picks"], r["per_request"])):
            key = (m, i % 4)
            if key not in ref:
                a, b = pool[m]
                mb = S.moved(b, 0xC5 * 1000 + 97 * m + key[1])
                ka, ca = O.elements_poly(a)
                kb, cb = O.elements_poly(mb)
                recs = O.Object(ka, ca, od=64, nthreads=16).collide(O.Object(kb, cb, od=64, nthreads=16))
                ref[key] = (len(recs), S.pair_digest(recs[:, [1, 3]]))
            equal += int(ref[key] == (cnt, dig))
        out["requests_equal_fp64_oracle"] = equal
        out["distinct_requests_checked"] = len(ref)
    return out


def stream_copy_gbs(torch, mib=1024, reps=5):
    """Measured device-to-device copy bandwidth (read + write bytes per
    second), the practical HBM ceiling beside the 8 TB/s peak."""
    src = torch.empty(mib << 20, dtype=torch.uint8, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del src, dst
    return round(2 * (mib << 20) / (ms * 1e-3) / 1e9, 1)


def pick_launch(args, world, autotune):
    """The narrowphase launch shape a form is timed at: --launch pins one
    (profiling runs of an autotuned shape without the tuner's launches),
    --no-tune keeps the default, else the autotuner picks it."""
    if args.launch:
        batch, block, xcd = (int(x) for x in args.launch.split(","))
        world.set_launch(batch=batch, block_size=block, xcd_remap=bool(xcd))
        return world.launch()
    return world.launch() if args.no_tune else autotune()


def run_steps(world, step, steps, warmup, stream, torch):
    """Warmup, then `steps` timed steps bracketed by HIP events on the
    library's stream; returns (ms per step, per-stage ms per step, per-stage
    ms per launch, last step's return value)."""
    for _ in range(warmup):
        step()
        world.stage_times()
    torch.cuda.synchronize()
    stage_sum, stage_cnt = {}, {}
    total_ms = 0.0
    ret = None
    for _ in range(steps):
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        ret = step()
        e1.record(stream)
        torch.cuda.synchronize()
        total_ms += e0.elapsed_time(e1)
        for name, ms in world.stage_times():
            stage_sum[name] = stage_sum.get(name, 0.0) + ms
            stage_cnt[name] = stage_cnt.get(name, 0) + 1
    stages = {k: stage_sum[k] / steps for k in stage_sum}
    per_launch = {k: stage_sum[k] / stage_cnt[k] for k in stage_sum}
    counts = {k: stage_cnt[k] / steps for k in stage_cnt}
    return total_ms / steps, stages, per_launch, counts, ret


def roofline_of(stages, per_launch, counts, nbytes_fn, ms_per_step, traffic_fn=None, dom=None):
    """Roofline of the dominant kernel's stage (default: the stage with the
    most time per step).  A stage that overlapped another (the two objects'
    builds on their own streams) has its wall time under that overlap, not
    its kernels' time, so such callers name the stage."""
    if dom is None:
        dom = max(stages, key=lambda k: stages[k])
    nbytes = nbytes_fn(dom)
    t_launch = per_launch[dom] * 1e-3
    achieved = nbytes / t_launch / 1e9 if t_launch > 0 else 0.0
    traffic = traffic_fn(dom) if traffic_fn else None
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "algorithmic_bytes": int(nbytes), "launch_ms": round(per_launch[dom], 4)}
    step_bytes = sum(nbytes_fn(k) * counts[k] for k in stages)
    step_frac = step_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS
    return roofline, round(step_frac, 4)


def bench_sheets(args, torch, stream):
    """Two wavy sheets, both rebuilt per step, A collided with B."""
    from distbvh_amd import CollisionWorld
    from distbvh_amd import capi
    from distbvh_amd import workloads as W

    a, b = W.wavy_sheets(args.nq)
    n = a.shape[2]
    ta = torch.from_numpy(a).cuda()
    tb = torch.from_numpy(b).cuda()
    out = {}
    split = {"clustering": capi.BVH_SPLIT_CLUSTERING, "geom_axis": capi.BVH_SPLIT_GEOM_AXIS,
             "ml_geom_axis": capi.BVH_SPLIT_ML_GEOM_AXIS}[args.split]

    def make(prec, timing):
        world = CollisionWorld(args.od, morton_bits=63, leaf_size=args.leaf, timing=timing,
                               stream=stream.cuda_stream, precision=prec, serial_builds=args.serial_builds)
        oa = world.create_collision_object()
        ob = world.create_collision_object()

        def step_launch():
            oa.set_entity_data_poly(ta, split_algorithm=split)
            ob.set_entity_data_poly(tb, split_algorithm=split)
            oa.init_broadphase()
            ob.init_broadphase()
            oa.broadphase(ob)

        def step():
            step_launch()
            return oa.result_count()

        return world, oa, ob, step_launch, step

    for prec in args.precisions:
        fp32 = prec != "fp64_records"  # fp32-outward records
        # the timed steps run without per-stage events; the stage breakdown
        # (and the narrowphase launch time of the roofline) comes from the same
        # step on a second world that records them
        world, oa, ob, step_launch, step = make(prec, False)
        step()
        # each form timed at its fastest launch shape
        tune = pick_launch(args, world, lambda: oa.autotune(ob))
        ms, _, _, _, npairs = run_steps(world, step, args.steps, args.warmup, stream, torch)
        world_t, oa_t, ob_t, _, step_t = make(prec, True)
        world_t.set_launch(batch=tune["batch"], block_size=tune["block_size"], xcd_remap=bool(tune["xcd_remap"]))
        step_t()
        world_t.stage_times()
        _, stages, per_launch, counts, _ = run_steps(world_t, step_t, max(5, args.steps // 2), 2, stream, torch)
        world_t.close()
        qstats = oa.result_stats()
        digest = oa.digest()
        graph_ms = None
        if args.split == "clustering":
            # the batching form on the same step: one HIP graph per step (a
            # branch per build), replayed on the resident element buffers
            with world.capture() as g:
                step_launch()
            gms, _, _, _, gpairs = run_steps(world, lambda: (g.launch(), oa.result_count())[1], args.steps,
                                             args.warmup, stream, torch)
            graph_ok = int(gpairs) == int(npairs) and oa.digest() == digest
            g.close()
            graph_ms = {"ms_per_step": round(gms, 4), "pairs_equal_eager": bool(graph_ok)}
        wl = "wavy_sheets_2M" if args.nq == 1024 and args.od == 256 else None
        # the narrowphase: the kernel with the most device time per step (the
        # builds of A and B overlap, so their stage times are wall times)
        roof, step_frac = roofline_of(stages, per_launch, counts,
                                      lambda k: kernel_bytes(k, n, npairs, n, fp32), ms,
                                      (lambda k: measured_traffic(k, wl, prec, tune)) if wl else None,
                                      dom="narrowphase")
        roof["kernel_name"] = stage_kernel(roof["kernel"], fp32, tune)
        # after the timed steps: the eager step once more, then its 32-byte
        # records materialised and copied to the host (k_records + copy)
        step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        recs = oa.results()
        records_ms = (time.perf_counter() - t0) * 1e3
        out[prec] = {"ms_per_step": ms, "stages": stages, "npairs": int(npairs), "stats": qstats,
                     "roofline": roof, "step_hbm_frac": step_frac, "digest": digest, "graph": graph_ms,
                     "launch": {k: tune[k] for k in ("batch", "block_size", "xcd_remap")},
                     "records": recs, "records_ms": records_ms}
        world.close()
    return a, b, n, out


def bench_spheres_self(args, torch, stream):
    """One sphere cloud (Gaussian mixture, translated far from the origin),
    rebuilt per step, self collision (each unordered pair once)."""
    from distbvh_amd import CollisionWorld
    from distbvh_amd import workloads as W

    s = W.sphere_cloud(args.spheres)
    n = s.shape[1]
    ts = torch.from_numpy(s).cuda()
    del s
    out = {}
    for prec in args.precisions:
        fp32 = prec != "fp64_records"  # fp32-outward records
        world = CollisionWorld(args.od, morton_bits=63, leaf_size=args.leaf, timing=True,
                               stream=stream.cuda_stream, precision=prec)
        o = world.create_collision_object()

        def step():
            o.set_entity_data_sphere(ts)
            o.init_broadphase()
            o.self_collision()
            return o.result_count()

        step()
        # each form timed at its fastest launch shape
        tune = pick_launch(args, world, o.autotune)
        world.stage_times()
        ms, stages, per_launch, counts, npairs = run_steps(world, step, args.steps, args.warmup, stream, torch)
        qstats = o.result_stats()

        def nb(k):
            if k == "bounds":  # read x, y, z, r; write centroid
                return n * (32 + 24)
            return kernel_bytes(k, n, npairs, n, fp32, elem_bytes=32)

        roof, step_frac = roofline_of(stages, per_launch, counts, nb, ms)
        out[prec] = {"ms_per_step": ms, "stages": stages, "npairs": int(npairs), "stats": qstats,
                     "roofline": roof, "step_hbm_frac": step_frac, "digest": o.digest(),
                     "launch": {k: tune[k] for k in ("batch", "block_size", "xcd_remap")}}
        world.close()
    return n, out


def batch_oracle_sample(a, counts, pairs, lattice=16, nsample=8, seed=7):
    """FP64 oracle on a sample of the batch: for `nsample` meshes, the
    brute-force 26-DOP pairs between the mesh and every other mesh of its
    3x3x3 lattice neighbourhood (no other mesh can reach it), against the
    GPU's cross-mesh pairs that involve the mesh."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    starts = np.concatenate([[0], np.cumsum(counts)])
    mesh = np.repeat(np.arange(len(counts)), counts)
    rng = np.random.default_rng(seed)
    # meshes of at most 8,192 triangles, so the brute force stays within seconds
    small = np.flatnonzero(np.asarray(counts) <= 8192)
    pick = rng.choice(small, size=min(nsample, len(small)), replace=False)
    got_all = np.asarray(pairs, np.uint64)
    ok = 0
    npairs = 0
    for m in pick:
        ci, cj, cl = m % lattice, (m // lattice) % lattice, m // (lattice * lattice)
        nb = []
        for k in range(len(counts)):
            if k == m:
                continue
            ki, kj, kl = k % lattice, (k // lattice) % lattice, k // (lattice * lattice)
            if abs(ki - ci) <= 1 and abs(kj - cj) <= 1 and abs(kl - cl) <= 1:
                nb.append(np.arange(starts[k], starts[k + 1]))
        idx_m = np.arange(starts[m], starts[m + 1])
        idx_n = np.concatenate(nb) if nb else np.zeros(0, np.int64)
        km, _ = O.elements_poly(a[:, :, idx_m])
        kn, _ = O.elements_poly(a[:, :, idx_n])
        bp = O.brute_pairs(km, kn, nthreads=16) if len(idx_n) else np.zeros((0, 2), np.uint64)
        gi = idx_m[bp[:, 0].astype(np.int64)].astype(np.uint64)
        gj = idx_n[bp[:, 1].astype(np.int64)].astype(np.uint64)
        want = np.sort(np.minimum(gi, gj) << np.uint64(32) | np.maximum(gi, gj))
        sel = (mesh[(got_all >> np.uint64(32)).astype(np.int64)] == m) | \
              (mesh[(got_all & np.uint64(0xffffffff)).astype(np.int64)] == m)
        got = np.sort(got_all[sel])
        ok += int(np.array_equal(got, want))
        npairs += len(want)
    return {"meshes": int(len(pick)), "meshes_equal": ok, "pairs_checked": int(npairs)}


def bench_batch(args, torch, stream):
    """Batching form: a batch of 4,096 height-field meshes (lognormal sizes,
    ~19M triangles) in one object with one patch per mesh (the meshes are
    the workload's patches), rebuilt per step and collided patch against
    patch (pairs between different meshes); the step is captured once into a
    HIP graph and replayed, after the autotuner picked the launch shape."""
    from distbvh_amd import CollisionWorld, GIVEN
    from distbvh_amd import workloads as W

    a, counts = W.patch_batch()
    n = a.shape[2]
    ta = torch.from_numpy(a).cuda()
    offsets = np.cumsum(counts)[:-1]
    # fp32-outward k-DOPs unless a precision is named; the other forms run too
    prec = args.precisions[0] if args.precision_given and args.precision != "both" else "fp32"
    forms = {}
    for p in [prec] + [f for f in ("fp32", "fp64", "fp64_records") if f != prec]:
        world = CollisionWorld(4096, morton_bits=63, timing=False, stream=stream.cuda_stream, precision=p)
        o = world.create_collision_object()
        o.set_patch_offsets(offsets)

        def eager():
            o.set_entity_data_poly(ta, split_algorithm=GIVEN)
            o.init_broadphase()
            o.self_collision_cross_patch()

        eager()
        tune = pick_launch(args, world, o.autotune)
        eager()
        npairs_eager = o.result_count()
        digest_eager = o.digest()
        ms_eager, _, _, _, _ = run_steps(world, lambda: (eager(), o.result_count())[1], args.steps, args.warmup,
                                         stream, torch)
        with world.capture() as g:
            eager()

        def replay():
            g.launch()
            return o.result_count()

        ms_graph, _, _, _, npairs = run_steps(world, replay, args.steps, args.warmup, stream, torch)
        ok = npairs == npairs_eager and o.digest() == digest_eager
        st = o.result_stats()
        forms[p] = {"ms_graph": ms_graph, "ms_eager": ms_eager, "npairs": int(npairs), "digest": o.digest(),
                    "graph_ok": bool(ok), "stats": st, "tune": tune}
        if p == "fp64" and not args.no_cpu:
            pr = o.pairs()
            forms[p]["pairs_host"] = (pr[:, 0] << np.uint64(32)) | pr[:, 1]
        g.close()
        world.close()
    f = forms[prec]
    f64 = forms["fp64"]
    sample = batch_oracle_sample(a, counts, f64["pairs_host"]) if not args.no_cpu else None
    ms_graph = f["ms_graph"]
    st = f["stats"]
    tune = f["tune"]
    return {"metric": "ms/step batch build + patch-vs-patch collision, 4096 meshes (%d triangles), one HIP graph" % n,
            "value": round(ms_graph, 4), "unit": "ms/step", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": round(ms_graph, 4), "higher_is_better": False,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64" if prec == "fp64_records" else "f32+f64",
            "data": "synthetic",
            "config": {"workload": "4,096 height-field meshes on a 16^3 lattice, lognormal sizes (median "
                                   "2,048, clipped to [64, 65,536]), 10%% neighbour overlap; %d triangles; one "
                                   "patch per mesh (BVH_SPLIT_GIVEN); pairs between different meshes "
                                   "(bvh_self_collision_cross_patch)" % n,
                       "precision": prec, "od": 4096,
                       "autotuned": {"batch": tune["batch"], "block_size": tune["block_size"],
                                     "xcd_remap": tune["xcd_remap"],
                                     "query_ms": round(tune["ms"], 4) if "ms" in tune else None,
                                     "tried_ms": [round(t, 4) for t in tune.get("tried_ms", [])]},
                       "pairs": f["npairs"], "pairs_per_s": round(f["npairs"] / (ms_graph * 1e-3), 1),
                       "candidate_pairs": int(st["candidate_pairs"]), "rechecked_pairs": int(st["rechecked_pairs"]),
                       "node_visits_per_wave": round(st["node_visits"] / ((n + 63) // 64), 1),
                       "eager_ms_per_step": round(f["ms_eager"], 4), "graph_pairs_equal_eager": f["graph_ok"],
                       "parallelism": "single GPU"},
            "forms": {p: {"graph_ms": round(v["ms_graph"], 4), "eager_ms": round(v["ms_eager"], 4),
                          "pairs": v["npairs"]} for p, v in forms.items()},
            "fp32_pairs_equal_fp64": bool(forms["fp32"]["npairs"] == f64["npairs"] and
                                          forms["fp32"]["digest"] == f64["digest"]),
            "fp64_records_pairs_equal_fp64": bool(forms["fp64_records"]["npairs"] == f64["npairs"] and
                                                  forms["fp64_records"]["digest"] == f64["digest"]),
            "fp64_oracle_sample": sample,
            "cpu_baseline": None}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", choices=["sheets", "spheres-self", "batch", "serving"], default="sheets",
                    help="sheets: the 2 x 2,097,152-triangle contact step (default); spheres-self: "
                         "self collision of a translated Gaussian-mixture sphere cloud; batch: one HIP "
                         "graph per step over a batch of 4,096 meshes; serving: Poisson request stream")
    ap.add_argument("--rate", type=float, default=200.0, help="serving: requests per second")
    ap.add_argument("--duration", type=float, default=2.0, help="serving: seconds of arrivals")
    ap.add_argument("--max-batch", type=int, default=16,
                    help="serving: most requests coalesced into one step (1: each request alone)")
    ap.add_argument("--no-serving", action="store_true",
                    help="sheets: leave the serving leg out of the line")
    ap.add_argument("--nq", type=int, default=1024, help="quads per sheet side")
    ap.add_argument("--spheres", type=int, default=16777216, help="spheres of the spheres-self workload")
    ap.add_argument("--od", type=int, default=256)
    ap.add_argument("--leaf", type=int, default=4,
                    help="leaf size of the production workload (4), used by the CPU oracle's element "
                         "trees; the GPU trees hold one element per leaf and test leaves from their "
                         "parents' records, so it does not change their work")
    ap.add_argument("--precision", choices=["fp64", "fp32", "fp64_records", "both"], default="both",
                    help="fp64: the default form (fp32-outward bounds in the tree; a candidate is "
                         "kept when its fp32 margin proves the fp64 overlap, else decided in fp64; "
                         "fp64 element k-DOP exports); fp32: the fp32-outward form (the same walk, fp32 "
                         "slabs only); fp64_records: fp64 slabs in every tree record; both: all three "
                         "(the line reports fp64 and compares the others)")
    ap.add_argument("--split", choices=["clustering", "geom_axis", "ml_geom_axis"], default="clustering",
                    help="sheets: the split_algorithm of set_entity_data (patches)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-runs", type=int, default=5, help="runs of each CPU baseline variant (median)")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the default launch shape instead of autotuning it before timing "
                         "(profiling runs: no tuner launches in the trace)")
    ap.add_argument("--launch", default="", metavar="BATCH,BLOCK,XCD",
                    help="pin the narrowphase launch shape of every form (e.g. 4,256,1) instead of "
                         "autotuning it (profiling runs of a shape the autotuner picked)")
    ap.add_argument("--serial-builds", action="store_true",
                    help="sheets: build A and B one after another on the world stream "
                         "(BVH_FLAG_SERIAL_BUILDS) instead of concurrently on their own streams")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="first copy a MIB-sized device buffer once (the reference kernel for "
                         "scaling FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    args.precision_given = any(a == "--precision" or a.startswith("--precision=") for a in sys.argv[1:])
    args.precisions = ["fp64", "fp32", "fp64_records"] if args.precision == "both" else [args.precision]

    if args.gpus != 1:
        print("bench.py measures one GPU: this library has no multi-GPU path", file=sys.stderr)
        sys.exit(2)

    import torch

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    if args.calibrate_copy:
        src = torch.empty(args.calibrate_copy << 20, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    stream = torch.cuda.current_stream()
    if args.workload == "serving":
        r = serving_leg(args, check_oracle=not args.no_cpu)
        out = {"metric": "serving p99 latency (arrival -> pairs on host), Poisson %.0f req/s, mesh pairs "
                         "10k-2M triangles" % args.rate,
               "value": r["p99_ms"], "unit": "ms", "n_gpus": 1, "steps": r["requests"], "warmup": 0,
               "ms_per_step": r["mean_ms"], "higher_is_better": False, "scaling": "weak",
               "vs_baseline": None, "dtype": "f64" if r["precision"] == "fp64_records" else "f32+f64",
               "data": "synthetic",
               "config": {"workload": "open-loop Poisson stream; each request a wavy-sheet pair (log-uniform "
                                      "10k-2M triangles per object, 8 pairs, 4 motion variants) with B rigidly "
                                      "moved + 0.1%% jitter, uploaded from pinned host memory; refit B, rebuild "
                                      "when the root 26-DOP grew > 1.2x; continuous batching: every waiting "
                                      "request (up to %d, one per mesh pair, within one 200 GB budget) joins "
                                      "the next coalesced step (one narrowphase launch); pairs copied to the "
                                      "host" % args.max_batch,
                          "precision": r["precision"], "parallelism": "single GPU"},
               "serving": r, "cpu_baseline": None}
        print(json.dumps(out), flush=True)
        return
    if args.workload == "batch":
        print(json.dumps(bench_batch(args, torch, stream)), flush=True)
        return
    if args.workload == "sheets":
        a, b, n, res = bench_sheets(args, torch, stream)
        workload = ("wavy contact sheets, 2 objects x %d triangles (%dx%d quads), fp64 vertices, 26-DOP, "
                    "63-bit Morton, %d patches, leaf size %d" % (n, args.nq, args.nq, args.od, args.leaf))
        if args.split != "clustering":
            workload += ", %s patches" % args.split
        metric = "ms/step build+broad+narrow @2M tris/object, %HBM peak; batched p99 latency"
    else:
        n, res = bench_spheres_self(args, torch, stream)
        workload = ("sphere cloud self collision, %d spheres (r = 1e-3, 64-component Gaussian mixture, "
                    "translated by (1e4, -1e4, 1e4)), 26-DOP, 63-bit Morton, %d patches, leaf size %d"
                    % (n, args.od, args.leaf))
        metric = "ms/step build+self-collision @%d spheres" % n
    first = res[args.precisions[0]]
    form = args.precisions[0]
    waves = (n + 63) // 64
    st = first["stats"]
    out = {
        "metric": metric,
        "value": round(first["ms_per_step"], 4),
        "unit": "ms/step",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(first["ms_per_step"], 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64 (pair predicate) + f32 (outward bounds, margin proofs)",
                  "fp32": "f32 (outward slabs) + f64 (undecided pairs)", "fp64_records": "f64"}[form],
        "data": "synthetic",
        "config": {"workload": workload,
                   "precision": {"fp64": "fp64: the fp64 pair set; tree records hold fp32 bounds rounded "
                                         "outward, a candidate is kept when its fp32 margin proves the fp64 "
                                         "overlap, else decided on fp64 k-DOPs",
                                 "fp32": "fp32-outward slabs, exact fp64 re-check of what fp32 cannot prove",
                                 "fp64_records": "fp64 slabs in every tree record"}[form],
                   "elements_per_object": n, "od": args.od,
                   "leaf_size": 1,
                   "leaf_size_note": "the GPU trees hold one element per leaf and test both children of a "
                                     "record at once; the workload's leaf size %d is the CPU oracle's "
                                     "(cpu_oracle_leaf_size)" % args.leaf,
                   "cpu_oracle_leaf_size": args.leaf,
                   "timed_step": "both builds, patch broadphase and element narrowphase, ending when the pair "
                                 "count is on the host; the 32-byte records are materialised after the timed "
                                 "region (records_ms)",
                   "pairs": first["npairs"],
                   "patch_pairs": int(st["patch_pairs"]),
                   "candidate_pairs": int(st["candidate_pairs"]),
                   "rechecked_pairs": int(st["rechecked_pairs"]),
                   "node_visits_per_wave": round(st["node_visits"] / waves, 1),
                   "launch_autotuned": first["launch"],
                   "builds": ("one object" if args.workload != "sheets" else
                              "serial (world stream)" if args.serial_builds else
                              "concurrent (A and B each on its own stream; the build stages in stages_ms "
                              "are wall times under that overlap)"),
                   "parallelism": "single GPU"},
        "stages_ms": {k: round(v, 4) for k, v in first["stages"].items()},
        "stages_note": "per-stage device times from HIP events recorded around each stage, on a second world "
                       "running the same step (the timed steps record no stage events)",
        "stream_copy_gbs": stream_copy_gbs(torch),
        "graph_form": first.get("graph"),
        "records_ms": round(first["records_ms"], 3) if "records_ms" in first else None,
        "step_hbm_frac": first["step_hbm_frac"],
        "roofline": first["roofline"],
    }
    if len(args.precisions) == 3:
        r64, r32, rr = res["fp64"], res["fp32"], res["fp64_records"]
        out["precision_forms"] = {
            "fp64_ms": round(r64["ms_per_step"], 4), "fp32_ms": round(r32["ms_per_step"], 4),
            "fp64_records_ms": round(rr["ms_per_step"], 4),
            "fp32_speedup": round(r64["ms_per_step"] / r32["ms_per_step"], 3),
            "fp32_speedup_vs_fp64_records": round(rr["ms_per_step"] / r32["ms_per_step"], 3),
            "pairs_identical": bool(r64["npairs"] == r32["npairs"] == rr["npairs"] and
                                    r64["digest"] == r32["digest"] == rr["digest"]),
            "fp64_records_stages_ms": {k: round(v, 4) for k, v in rr["stages"].items()},
            "fp64_records_launch_autotuned": rr["launch"],
            "fp64_records_roofline": rr["roofline"],
            "fp32_stages_ms": {k: round(v, 4) for k, v in r32["stages"].items()},
            "fp32_graph_form": r32.get("graph"),
            "fp32_launch_autotuned": r32["launch"],
            "fp32_candidate_pairs": int(r32["stats"]["candidate_pairs"]),
            # leaf candidates of the fp32 walk that are not fp64 pairs, per fp64 pair
            "fp32_excess_false_positive_rate": round(
                (int(r32["stats"]["candidate_pairs"]) - int(r64["npairs"])) / max(1, int(r64["npairs"])), 6),
            "fp32_rechecked_pairs": int(r32["stats"]["rechecked_pairs"]),
            "fp32_max_slab_error_ulps": r32["stats"]["max_slab_error_ulps"],
        }
    if not args.no_cpu and args.workload == "sheets" and args.split == "clustering":
        out["cpu_baseline"] = cpu_baseline(a, b, args.od, args.leaf, runs=args.cpu_runs,
                                           gpu_records=res["fp64"]["records"] if "fp64" in res else None)
        out["config"]["records_equal_cpu_oracle"] = out["cpu_baseline"]["records_equal_gpu"]
        out["gpu_fraction_of_cpu_time"] = round(first["ms_per_step"] / out["cpu_baseline"]["value"], 5)
    else:
        out["cpu_baseline"] = None
    if args.workload == "sheets" and not args.no_serving:
        out["serving"] = serving_leg(args, check_oracle=not args.no_cpu)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
