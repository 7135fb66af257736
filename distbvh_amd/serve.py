"""Serving loop: collision queries arriving as an open-loop Poisson stream,
served by continuous batching.

Each request names a mesh pair (A static, B moved by a seeded rigid motion plus
vertex jitter) and carries B's vertices in pinned host memory.  One world holds
the built objects of the mesh pairs resident in HBM under one device-memory
budget shared by every request (least recently used pairs are released when a
new pair needs the room).  A dispatcher takes every request that has arrived
and is waiting -- up to `max_batch`, at most one per mesh pair, as far as the
budget admits -- into one coalesced step: each admitted pair's B is uploaded
and refit (rebuilt when the refitted root 26-DOP grew by more than
`growth_limit`, or when the pair was not resident), then the element
narrowphases of all of them run as one launch (bvh_broadphase_batch) and each
request's pairs are copied to the host.  Requests that arrive while a step is
on the GPU wait for the next step.  Latency is measured from a request's
arrival to its pairs on the host, so queueing counts.

The workload's sizes and seeds are parameters of `serve()`; `bench.py` prints
its p50 / p99 latency and pair throughput.
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


def mesh_pool(k, lo, hi, seed, sizes=None):
    """k mesh pairs with triangle counts log-uniform over [lo, hi] (or the
    given `sizes`): wavy contact sheets of nq x nq quads (2 nq^2 triangles)."""
    if sizes is None:
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
    """Resident objects per mesh pair in one world under one budget;
    refit-or-rebuild per request; requests served in coalesced steps."""

    def __init__(self, od=64, precision="fp64", budget_bytes=200 << 30, growth_limit=1.2, leaf_size=1):
        self.world = CollisionWorld(od, leaf_size=leaf_size, precision=precision)
        self.budget = budget_bytes
        self.growth_limit = growth_limit
        self.cache = {}   # mesh id -> dict(a, b, width0, last_use, bytes)
        self.free = []    # released object pairs, reused for the next pair admitted
        self.bytes_per_triangle = 1400.0  # device bytes of a resident pair per triangle of A + B (refined)
        self.refits = 0
        self.rebuilds = 0
        self.evictions = 0
        self.deferred = 0
        self.steps = 0
        self.clock = 0

    def resident_bytes(self):
        return sum(e["bytes"] for e in self.cache.values())

    def _release(self, m):
        e = self.cache.pop(m)
        # the device memory goes back; the handles are kept for reuse
        e["a"].release()
        e["b"].release()
        self.free.append((e["a"], e["b"]))
        self.evictions += 1

    def admit(self, mesh_id, ntri, busy):
        """Room for mesh pair `mesh_id` (ntri triangles per object): True if it
        is resident or fits the budget after releasing least recently used
        pairs that are not in `busy` (the step being formed)."""
        if mesh_id in self.cache:
            return True
        need = self.bytes_per_triangle * 2 * ntri
        while self.resident_bytes() + need > self.budget:
            idle = [m for m in self.cache if m not in busy]
            if not idle:
                return False
            self._release(min(idle, key=lambda m: self.cache[m]["last_use"]))
        return True

    def step(self, reqs, on_done=None, digests=False):
        """One coalesced step.  reqs: list of (mesh id, A vertices, B vertices)
        with distinct mesh ids, all admitted.  Returns each request's (k, 2)
        pairs; on_done(i, pairs) is called as request i's pairs reach the
        host.  With `digests`, returns (pairs, digests) instead: each
        request's order-independent device digest (bvh_result_digest),
        computed after every request of the step reached the host."""
        self.steps += 1
        entries = []
        refitted = []
        for m, a_verts, b_verts in reqs:
            self.clock += 1
            e = self.cache.get(m)
            if e is None:
                if self.free:
                    oa, ob = self.free.pop()
                else:
                    oa = self.world.create_collision_object()
                    ob = self.world.create_collision_object()
                oa.set_entity_data_poly(a_verts)
                oa.init_broadphase()
                ob.set_entity_data_poly(b_verts)
                e = {"a": oa, "b": ob, "width0": None, "bytes": 0}
                self.cache[m] = e
                self.rebuilds += 1
            else:
                e["b"].refit(b_verts)
                refitted.append((e, b_verts))
            e["last_use"] = self.clock
            entries.append(e)
        for e, b_verts in refitted:
            if _width(e["b"].root_kdop()) > self.growth_limit * e["width0"]:
                e["b"].set_entity_data_poly(b_verts)
                e["width0"] = None
                self.rebuilds += 1
            else:
                self.refits += 1
        for e in entries:
            if e["width0"] is None:
                e["width0"] = _width(e["b"].root_kdop())
                e["bytes"] = e["a"].device_bytes() + e["b"].device_bytes()
                ntri = e["a"].n + e["b"].n
                if ntri:
                    self.bytes_per_triangle = max(self.bytes_per_triangle, e["bytes"] / ntri)
            e["b"].init_broadphase()
        self.world.broadphase_batch([e["a"] for e in entries], [e["b"] for e in entries])
        out = []
        for i, e in enumerate(entries):
            pairs = e["a"].pairs()
            if on_done is not None:
                on_done(i, pairs)
            out.append(pairs)
        if digests:
            return out, [e["a"].digest() for e in entries]
        return out

    def close(self):
        self.cache.clear()
        self.world.close()


def _mix64(z):
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xbf58476d1ce4e5b9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94d049bb133111eb)
    return z ^ (z >> np.uint64(31))


def pair_digest(pairs):
    """Host form of the device digest bvh_result_digest of (k, 2) element
    pairs: the sum and the xor of two 64-bit mixes (splitmix64 finalizer) of
    every key a << 32 | b.  Order-independent, so a request's device digest
    can be compared with the digest of pairs computed elsewhere."""
    p = np.asarray(pairs, dtype=np.uint64).reshape(-1, 2)
    if len(p) == 0:
        return 0, 0
    keys = (p[:, 0] << np.uint64(32)) | p[:, 1]
    with np.errstate(over="ignore"):
        s = int(_mix64(keys).sum(dtype=np.uint64))
        x = int(np.bitwise_xor.reduce(_mix64(keys ^ np.uint64(0x9e3779b97f4a7c15))))
    return s, x


def serve(rate=200.0, duration=2.0, k=8, lo=1e4, hi=2e6, variants=4, seed=0xC5, precision="fp64", od=64,
          budget_bytes=200 << 30, warm=True, pin=True, max_batch=16, record=False, sizes=None):
    """Runs the open-loop stream for `duration` seconds of arrivals and
    returns latency percentiles (ms), pairs/s, the refit / rebuild /
    eviction counts and the coalesced step sizes.

    One dispatcher (this thread) serves the stream by continuous batching:
    whenever the GPU is free it forms a step of every waiting request (FIFO,
    up to `max_batch`, one per mesh pair, within the shared device-memory
    budget; a request that does not fit waits for room) and runs it as one
    coalesced step.  max_batch=1 serves each request alone, one after
    another.  With `record`, the result also lists every request's (pair
    count, digest) in arrival order, so a run in the mix can be compared
    with a run of the same requests alone."""
    import torch

    from . import capi

    capi.load()
    pool = mesh_pool(k, lo, hi, seed, sizes=sizes)
    k = len(pool)
    # request payloads: `variants` moved copies of each B, in pinned host memory
    payload = []
    for i, (a, b) in enumerate(pool):
        vs = []
        for v in range(variants):
            mv = moved(b, seed * 1000 + 97 * i + v)
            t = torch.from_numpy(mv)
            vs.append((t.pin_memory() if pin else t).numpy())
        at = torch.from_numpy(a)
        payload.append(((at.pin_memory() if pin else at).numpy(), vs))
    nreq = max(1, int(rate * duration))
    gaps = -np.log(1.0 - W.uniform(seed ^ 0x1, nreq)) / rate
    arrivals = np.cumsum(gaps)
    picks = np.minimum((W.uniform(seed ^ 0x2, nreq) * k).astype(np.int64), k - 1)
    lat = np.zeros(nreq)
    counts = np.zeros(nreq, np.int64)
    digests = [None] * nreq
    step_sizes = []
    srv = Server(od=od, precision=precision, budget_bytes=budget_bytes)
    try:
        if warm:  # first build of every pair that fits (cold start is not part of the stream)
            for m in range(k):
                a, vs = payload[m]
                if srv.admit(m, a.shape[2], set()):
                    srv.step([(m, a, vs[0])])
        srv.refits = srv.rebuilds = srv.evictions = srv.steps = 0
        pending = []
        nxt = 0
        served = 0
        t0 = time.perf_counter()
        while served < nreq:
            now = time.perf_counter() - t0
            while nxt < nreq and arrivals[nxt] <= now:
                pending.append(nxt)
                nxt += 1
            if not pending:
                time.sleep(max(0.0, arrivals[nxt] - now))
                continue
            batch, busy = [], set()
            for r in pending:
                if len(batch) >= max_batch:
                    break
                m = int(picks[r])
                if m in busy:
                    continue
                if not srv.admit(m, payload[m][0].shape[2], busy):
                    srv.deferred += 1
                    continue
                batch.append(r)
                busy.add(m)
            if not batch:  # nothing fits beside what is resident: the oldest request alone
                r = pending[0]
                srv.admit(int(picks[r]), payload[int(picks[r])][0].shape[2], set())
                batch = [r]
            taken = set(batch)
            pending = [r for r in pending if r not in taken]
            reqs = [(int(picks[r]), payload[int(picks[r])][0], payload[int(picks[r])][1][r % variants])
                    for r in batch]

            def done(i, pairs, batch=batch):
                r = batch[i]
                lat[r] = (time.perf_counter() - t0) - arrivals[r]
                counts[r] = len(pairs)

            res = srv.step(reqs, done, digests=record)
            if record:
                for r, dig in zip(batch, res[1]):
                    digests[r] = dig
            step_sizes.append(len(batch))
            served += len(batch)
        wall = time.perf_counter() - t0
        stats = (srv.refits, srv.rebuilds, srv.evictions, srv.deferred, srv.steps)
    finally:
        srv.close()
    npairs = int(counts.sum())
    out = {
        "requests": nreq, "rate_req_s": rate, "wall_s": wall, "max_batch": max_batch,
        "p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
        "mean_ms": float(lat.mean() * 1e3), "max_ms": float(lat.max() * 1e3),
        "pairs": npairs, "pairs_per_s": npairs / wall,
        "refits": stats[0], "rebuilds": stats[1], "evictions": stats[2], "deferred": stats[3],
        "steps": stats[4], "mean_step_requests": float(np.mean(step_sizes)), "max_step_requests": int(max(step_sizes)),
        "mesh_triangles": [int(p[0].shape[2]) for p in pool],
    }
    if record:
        out["per_request"] = [(int(c), d) for c, d in zip(counts, digests)]
        out["picks"] = [int(m) for m in picks]
    return out
