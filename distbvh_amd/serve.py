"""Serving loop: collision queries arriving as an open-loop Poisson stream.

Each request names a mesh pair (A static, B moved by a seeded rigid motion plus
vertex jitter) and carries B's vertices in host memory.  The server keeps the
built objects of every mesh pair resident in HBM (least recently used pairs
are dropped beyond a device-memory budget).  Per request it uploads B, refits
B's tree on the new vertices, and rebuilds B instead when the refitted root
26-DOP grew by more than `growth_limit` (sum of the 13 slab widths against the
last build) or when the pair is not cached; then it runs broadphase +
narrowphase of A against B and copies the pairs to the host.  Latency is
measured from a request's arrival time to its pairs on the host, so queueing
behind earlier requests counts.  Several slots (host threads, each with its
own world and stream) serve requests at once, so requests that arrive while
others are on the GPU overlap with them on the device.

The workload's sizes and seeds are parameters of `serve()`; `bench.py
--workload serving` prints its p50 / p99 latency and pair throughput.
"""
import math
import time

import numpy as np

from .collision import CollisionWorld
from . import workloads as W


def _rotation(ax, ay, az):
    cx, sx = math.cos(ax), math.sin(ax)
    cy, sy = math.cos(ay), math.sin(ay)
    cz, sz = math.cos(az), math.sin(az)
    rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
    return rz @ ry @ rx


def moved(verts, seed, jitter=1e-3, angle=0.02, shift=0.002):
    """B moved rigidly about its centre (rotation up to `angle` rad per axis,
    translation up to `shift`) plus per-vertex uniform jitter of `jitter`
    times the mesh extent (0.1% by default), all from `seed`."""
    u = W.uniform(seed, 6)
    rot = _rotation(*(angle * (2 * u[:3] - 1)))
    t = shift * (2 * u[3:] - 1)
    c = verts.mean(axis=(0, 2))
    ext = float(np.max(verts.max(axis=(0, 2)) - verts.min(axis=(0, 2))))
    out = np.einsum("ij,vjn->vin", rot, verts - c[None, :, None]) + (c + t)[None, :, None]
    jit = W.uniform(seed ^ 0x77, out.size).reshape(out.shape)
    out += jitter * ext * (2 * jit - 1)
    return np.ascontiguousarray(out)


def mesh_pool(k, lo, hi, seed):
    """k mesh pairs with triangle counts log-uniform over [lo, hi]: wavy
    contact sheets of nq x nq quads (2 nq^2 triangles)."""
    u = W.uniform(seed, k)
    sizes = np.exp(np.log(lo) + u * (np.log(hi) - np.log(lo)))
    pool = []
    for i, s in enumerate(sizes):
        nq = max(8, int(round(math.sqrt(s / 2.0))))
        a, b = W.wavy_sheets(nq, seed=seed + 17 * i)
        pool.append((a, b))
    return pool


def _width(kd):
    return float(np.sum(kd[1::2] - kd[0::2]))


class Server:
    """Resident objects per mesh pair, refit-or-rebuild per request."""

    def __init__(self, od=64, precision="fp64", budget_bytes=200 << 30, growth_limit=1.2, leaf_size=1):
        self.world = CollisionWorld(od, leaf_size=leaf_size, precision=precision)
        self.budget = budget_bytes
        self.growth_limit = growth_limit
        self.cache = {}   # mesh id -> dict(a, b, width0, last_use)
        self.free = []    # released object pairs, reused for the next pair admitted
        self.refits = 0
        self.rebuilds = 0
        self.evictions = 0
        self.clock = 0

    def _resident_bytes(self):
        return sum(e["a"].device_bytes() + e["b"].device_bytes() for e in self.cache.values())

    def _evict(self):
        while self.cache and self._resident_bytes() > self.budget:
            victim = min(self.cache, key=lambda m: self.cache[m]["last_use"])
            e = self.cache.pop(victim)
            # the device memory goes back; the handles are kept for reuse
            e["a"].release()
            e["b"].release()
            self.free.append((e["a"], e["b"]))
            self.evictions += 1

    def handle(self, mesh_id, a_verts, b_verts):
        self.clock += 1
        e = self.cache.get(mesh_id)
        if e is None:
            if self.free:
                oa, ob = self.free.pop()
            else:
                oa = self.world.create_collision_object()
                ob = self.world.create_collision_object()
            oa.set_entity_data_poly(a_verts)
            oa.init_broadphase()
            e = {"a": oa, "b": ob, "width0": None}
            self.cache[mesh_id] = e
        e["last_use"] = self.clock
        ob = e["b"]
        rebuild = e["width0"] is None
        if not rebuild:
            ob.refit(b_verts)
            if _width(ob.root_kdop()) > self.growth_limit * e["width0"]:
                rebuild = True
            else:
                self.refits += 1
        if rebuild:
            ob.set_entity_data_poly(b_verts)
            e["width0"] = _width(ob.root_kdop())
            self.rebuilds += 1
        ob.init_broadphase()
        e["a"].broadphase(ob)
        pairs = e["a"].pairs()
        self._evict()
        return pairs

    def close(self):
        self.cache.clear()
        self.world.close()


def _pair_digest(pairs):
    """Order-independent digest of a pair list: SHA-256 of the sorted rows."""
    import hashlib

    p = np.ascontiguousarray(pairs, dtype=np.uint64)
    if len(p):
        p = p[np.lexsort((p[:, 1], p[:, 0]))]
    return hashlib.sha256(p.tobytes()).hexdigest()[:16]


def serve(rate=200.0, duration=2.0, k=8, lo=1e4, hi=2e6, variants=4, seed=0xC5, precision="fp64", od=64,
          budget_bytes=200 << 30, warm=True, pin=True, slots=4, record=False):
    """Runs the open-loop stream for `duration` seconds of arrivals and
    returns latency percentiles (ms), pairs/s and the refit / rebuild counts.

    Requests are served by `slots` workers, each a host thread driving its
    own world (its own HIP stream) and holding the resident objects of the
    mesh pairs it owns (pair m belongs to slot m % slots; its share of the
    device-memory budget bounds its cache).  A request that arrives while
    others are on the GPU is admitted as soon as its slot is free, and the
    requests of different slots run on the device at the same time.  With
    `record`, the result also lists every request's (pair count, digest) in
    arrival order, so a run in the mix can be compared with a run of the
    same requests alone (slots=1)."""
    import queue
    import threading

    import torch

    from . import capi

    capi.load()  # before the workers start, so that they share one handle
    pool = mesh_pool(k, lo, hi, seed)
    # request payloads: `variants` moved copies of each B, in pinned host memory
    payload = []
    for i, (a, b) in enumerate(pool):
        vs = []
        for v in range(variants):
            mv = moved(b, seed * 1000 + 97 * i + v)
            t = torch.from_numpy(mv)
            vs.append(t.pin_memory() if pin else t)
        at = torch.from_numpy(a)
        payload.append((at.pin_memory() if pin else at, vs))
    nreq = max(1, int(rate * duration))
    gaps = -np.log(1.0 - W.uniform(seed ^ 0x1, nreq)) / rate
    arrivals = np.cumsum(gaps)
    picks = np.minimum((W.uniform(seed ^ 0x2, nreq) * k).astype(np.int64), k - 1)
    slots = max(1, min(int(slots), k))
    lat = np.zeros(nreq)
    counts = np.zeros(nreq, np.int64)
    digests = [None] * nreq
    queues = [queue.Queue() for _ in range(slots)]
    warmed = threading.Barrier(slots + 1)
    start = {}
    stats = [None] * slots
    errors = []

    def worker(s):
        srv = None
        waited = False
        try:
            srv = Server(od=od, precision=precision, budget_bytes=budget_bytes // slots)
            if warm:  # first build of every owned pair (cold start is not part of the stream)
                for m in range(s, k, slots):
                    at, vs = payload[m]
                    srv.handle(m, at.numpy(), vs[0].numpy())
            waited = True
            warmed.wait()
            while True:
                item = queues[s].get()
                if item is None:
                    break
                r, m = item
                at, vs = payload[m]
                pairs = srv.handle(m, at.numpy(), vs[r % variants].numpy())
                lat[r] = (time.perf_counter() - start["t0"]) - arrivals[r]
                counts[r] = len(pairs)
                if record:
                    digests[r] = _pair_digest(pairs)
        except Exception as e:  # reported after the stream
            errors.append(e)
            if not waited:
                warmed.wait()
        finally:
            if srv is not None:
                stats[s] = (srv.refits, srv.rebuilds, srv.evictions)
                srv.close()

    threads = [threading.Thread(target=worker, args=(s,), daemon=True) for s in range(slots)]
    for th in threads:
        th.start()
    warmed.wait()
    start["t0"] = t0 = time.perf_counter()
    for r in range(nreq):
        now = time.perf_counter() - t0
        if now < arrivals[r]:
            time.sleep(arrivals[r] - now)
        m = int(picks[r])
        queues[m % slots].put((r, m))
    for q in queues:
        q.put(None)
    for th in threads:
        th.join()
    wall = time.perf_counter() - t0
    if errors:
        raise errors[0]
    npairs = int(counts.sum())
    out = {
        "requests": nreq, "rate_req_s": rate, "wall_s": wall, "slots": slots,
        "p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
        "mean_ms": float(lat.mean() * 1e3), "max_ms": float(lat.max() * 1e3),
        "pairs": npairs, "pairs_per_s": npairs / wall,
        "refits": sum(x[0] for x in stats if x), "rebuilds": sum(x[1] for x in stats if x),
        "evictions": sum(x[2] for x in stats if x),
        "mesh_triangles": [int(p[0].shape[2]) for p in pool],
    }
    if record:
        out["per_request"] = [(int(c), d) for c, d in zip(counts, digests)]
    return out
