"""Seeded synthetic collision workloads (host-side, numpy).

All randomness comes from SplitMix64 counters, so every stream is a pure
function of (seed, index) and the same bytes are handed to the GPU and to the
CPU checker.  Uniform doubles are the top 53 bits / 2^53; normal deviates use
Box-Muller in fp64 on the host.

Element layout: polytope elements are SoA planes `verts[v, c, e]` (vertex v,
coordinate c, element e), spheres are planes `xyzr[4, e]`.
"""
import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed, count, offset=0):
    """The SplitMix64 sequence: outputs offset .. offset+count-1 of stream `seed`."""
    with np.errstate(over="ignore"):
        k = np.arange(offset + 1, offset + count + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * GOLDEN
        z = (z ^ (z >> np.uint64(30))) * M1
        z = (z ^ (z >> np.uint64(27))) * M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed, count, offset=0):
    return (splitmix64(seed, count, offset) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal(seed, count, offset=0):
    u = uniform(seed, 2 * count, 2 * offset)
    u1 = 1.0 - u[0::2]  # in (0, 1]
    u2 = u[1::2]
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


def grid_triangles(vx, vy, vz, nx, ny):
    """Triangulate a (ny+1, nx+1) vertex grid.  Quad (i, j) -> triangles
    (v00, v10, v11) and (v00, v11, v01); triangle id = 2*(j*nx + i) + {0, 1}.
    Returns SoA planes (3, 3, 2*nx*ny)."""
    i = np.arange(nx)
    j = np.arange(ny)
    ii, jj = np.meshgrid(i, j)  # (ny, nx), row j
    ii = ii.ravel()
    jj = jj.ravel()
    w = nx + 1
    v00 = jj * w + ii
    v10 = jj * w + ii + 1
    v11 = (jj + 1) * w + ii + 1
    v01 = (jj + 1) * w + ii
    t0 = np.stack([v00, v10, v11])  # (3, nq)
    t1 = np.stack([v00, v11, v01])
    tri = np.empty((3, 2 * nx * ny), dtype=np.int64)
    tri[:, 0::2] = t0
    tri[:, 1::2] = t1
    coords = np.stack([vx.ravel(), vy.ravel(), vz.ravel()])  # (3, nverts)
    out = np.empty((3, 3, tri.shape[1]), np.float64)
    for v in range(3):
        out[v] = coords[:, tri[v]]
    return np.ascontiguousarray(out)


def flat_sheets(nq=64, seed=0xC1):
    """Flat sheet pair: object A is an nq x nq quad grid on [0,1]^2 x {0};
    object B is A shifted by (0.5 cell, 0.5 cell, 0.02) plus seeded uniform
    vertex noise of +-0.1 cell on every axis (shared vertices move together).
    nq = 64 gives 8,192 triangles per object."""
    h = 1.0 / nq
    g = np.arange(nq + 1, dtype=np.float64) / nq
    vx, vy = np.meshgrid(g, g)
    vz = np.zeros_like(vx)
    a = grid_triangles(vx, vy, vz, nq, nq)
    nv = (nq + 1) * (nq + 1)
    u = uniform(seed, 3 * nv).reshape(nv, 3)
    noise = -0.1 * h + (0.2 * h) * u
    bx = (vx.ravel() + 0.5 * h) + noise[:, 0]
    by = (vy.ravel() + 0.5 * h) + noise[:, 1]
    bz = (vz.ravel() + 0.02) + noise[:, 2]
    b = grid_triangles(bx, by, bz, nq, nq)
    return a, b


def wavy_sheets(nq=1024, seed=0xC2):
    """Deformed contact sheets: A: z = 0.01 sin(2 pi 8 x) cos(2 pi 8 y) on an
    nq x nq quad grid over [0,1]^2; B = A + (0, 0, 0.005) + N(0, 0.002^2) on
    every axis of every vertex.  nq = 1024 gives 2,097,152 triangles/object."""
    g = np.arange(nq + 1, dtype=np.float64) / nq
    vx, vy = np.meshgrid(g, g)
    vz = 0.01 * np.sin(2.0 * np.pi * vx * 8.0) * np.cos(2.0 * np.pi * vy * 8.0)
    a = grid_triangles(vx, vy, vz, nq, nq)
    nv = (nq + 1) * (nq + 1)
    z = normal(seed, 3 * nv).reshape(nv, 3) * 0.002
    bx = vx.ravel() + z[:, 0]
    by = vy.ravel() + z[:, 1]
    bz = (vz.ravel() + 0.005) + z[:, 2]
    b = grid_triangles(bx, by, bz, nq, nq)
    return a, b


def sphere_cloud(n=16777216, radius=1e-3, ncomp=64, sigma=0.01,
                 shift=(1e4, -1e4, 1e4), seed=0xC3):
    """Isotropic Gaussian-mixture sphere cloud (equal weights), translated to
    stress fp32 cancellation.  Returns planes (4, n): x, y, z, r."""
    means = uniform(seed, 3 * ncomp).reshape(ncomp, 3)
    comp = np.minimum((uniform(seed ^ 0x1, n) * ncomp).astype(np.int64), ncomp - 1)
    z = normal(seed ^ 0x2, 3 * n).reshape(3, n)
    out = np.empty((4, n), np.float64)
    for c in range(3):
        out[c] = (means[comp, c] + sigma * z[c]) + shift[c]
    out[3] = radius
    return out


def hex_grid(nx, ny, nz, origin_shift=0.0):
    """8-vertex hexahedra on [0,1]^3 in x-fastest order (the reference test
    fixture's buildElementGrid, tests/TestCommon.hpp:148-181).  Returns SoA
    planes (8, 3, nx*ny*nz)."""
    dx, dy, dz = 1.0 / nx, 1.0 / ny, 1.0 / nz
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    x = origin_shift + i.ravel() * dx
    y = origin_shift + j.ravel() * dy
    z = origin_shift + k.ravel() * dz
    corners = [(x, y, z), (x + dx, y, z), (x + dx, y + dy, z), (x, y + dy, z),
               (x, y, z + dz), (x + dx, y, z + dz), (x + dx, y + dy, z + dz), (x, y + dy, z + dz)]
    out = np.empty((8, 3, x.shape[0]), np.float64)
    for v, (cx, cy, cz) in enumerate(corners):
        out[v, 0] = cx
        out[v, 1] = cy
        out[v, 2] = cz
    return out


def touching_spheres(n=4096, radius=1e-3, shift=(1e3, -1e3, 1e3), max_ulps=4, seed=0xC4):
    """Two sphere sets whose i-th spheres are within a few fp64 ulps of
    touching along x: B_i's centre is fl(A_i.x + 2r) stepped k ulps, k in
    [-max_ulps, max_ulps].  The pairs with a negative gap overlap strictly,
    the others do not; every pair is a case an fp32 test cannot decide.
    Sphere i sits on a lattice 4r apart in y and z, far from the origin.
    Returns (a, b) planes (4, n)."""
    side = int(np.ceil(np.sqrt(n)))
    i = np.arange(n)
    a = np.empty((4, n), np.float64)
    a[0] = shift[0] + 0.37 * radius * (uniform(seed, n) - 0.5)
    a[1] = shift[1] + (i % side) * (8.0 * radius)
    a[2] = shift[2] + (i // side) * (8.0 * radius)
    a[3] = radius
    b = a.copy()
    k = (i % (2 * max_ulps + 1)) - max_ulps
    x = a[0] + 2.0 * radius
    for s in range(max_ulps):
        x = np.where(k > s, np.nextafter(x, np.inf), x)
        x = np.where(k < -s, np.nextafter(x, -np.inf), x)
    b[0] = x
    return a, b


def patch_batch(npatch=4096, median=2048, lo=64, hi=65536, sigma=1.32, lattice=16, overlap=0.1,
                height=0.05, seed=0xC4):
    """A batch of independent height-field meshes: patch k sits in cell
    (k % L, (k // L) % L, k // L^2) of an L^3 lattice (cell 1/L), widened by
    overlap/2 of a cell on each side in x and y so that neighbours overlap by
    `overlap`; its vertex heights are cell_z + U[0, height].  Patch sizes are
    lognormal (median `median` triangles, log-sd `sigma`) clipped to [lo, hi];
    a patch of s triangles is an m x m quad grid, m = max(2, round(sqrt(s/2))).
    Returns (planes (3, 3, N) float64, per-patch triangle counts)."""
    cell = 1.0 / lattice
    z = normal(seed, npatch)
    sizes = np.clip(np.rint(median * np.exp(sigma * z)), lo, hi).astype(np.int64)
    ms = np.maximum(2, np.rint(np.sqrt(sizes / 2.0))).astype(np.int64)
    counts = 2 * ms * ms
    total = int(counts.sum())
    out = np.empty((3, 3, total), np.float64)
    off = 0
    hoff = 0
    hseed = seed ^ 0x5A5A
    for k in range(npatch):
        m = int(ms[k])
        ci, cj, cl = k % lattice, (k // lattice) % lattice, k // (lattice * lattice)
        x0 = (ci - overlap / 2) * cell
        y0 = (cj - overlap / 2) * cell
        w = (1.0 + overlap) * cell
        nvtx = (m + 1) * (m + 1)
        hz = cl * cell + height * uniform(hseed, nvtx, offset=hoff).reshape(m + 1, m + 1)
        hoff += nvtx
        g = np.arange(m + 1) / m
        vx = np.broadcast_to(x0 + w * g[None, :], (m + 1, m + 1))
        vy = np.broadcast_to(y0 + w * g[:, None], (m + 1, m + 1))
        t = grid_triangles(vx, vy, hz, m, m)
        out[:, :, off:off + t.shape[2]] = t
        off += t.shape[2]
    assert off == total
    return out, counts
