"""Host-side logic of the serving loop (no GPU): the request payload is a
rigid motion plus bounded jitter, the mesh pool spans the requested sizes,
and the arrival process is Poisson at the requested rate."""
import numpy as np

from distbvh_amd import serve as S
from distbvh_amd import workloads as W


def test_moved_is_rigid_up_to_jitter():
    _, b = W.wavy_sheets(16)
    m = S.moved(b, 1234, jitter=0.0)
    # pairwise distances between vertices of the first triangles are kept
    p0 = b[:, :, :50].transpose(2, 0, 1).reshape(-1, 3)
    p1 = m[:, :, :50].transpose(2, 0, 1).reshape(-1, 3)
    d0 = np.linalg.norm(p0[:, None] - p0[None], axis=2)
    d1 = np.linalg.norm(p1[:, None] - p1[None], axis=2)
    assert np.allclose(d0, d1, atol=1e-12)
    assert not np.allclose(m, b)
    ext = float(np.max(b.max(axis=(0, 2)) - b.min(axis=(0, 2))))
    j = S.moved(b, 1234, jitter=1e-3)
    assert np.max(np.abs(j - m)) <= 1e-3 * ext * (1 + 1e-12)


def test_mesh_pool_sizes_log_uniform_range():
    pool = S.mesh_pool(6, 1e4, 2e6, 0xC5)
    sizes = [a.shape[2] for a, _ in pool]
    assert all(0.8e4 <= s <= 2.2e6 for s in sizes)
    assert all(a.shape == b.shape for a, b in pool)


def test_poisson_arrival_rate():
    rate, n = 200.0, 20000
    gaps = -np.log(1.0 - W.uniform(0x1, n)) / rate
    assert abs(gaps.mean() - 1 / rate) < 0.05 / rate
    assert abs(gaps.std() - 1 / rate) < 0.05 / rate


def test_pair_digest_is_order_independent_host_form():
    """serve.pair_digest restates the device digest (sum and xor of two
    splitmix64 mixes of a << 32 | b): independent of the pair order, changed
    by any changed pair, and (0, 0) for no pairs."""
    rng = np.random.default_rng(5)
    p = rng.integers(0, 1 << 31, size=(1000, 2), dtype=np.uint64)
    d = S.pair_digest(p)
    assert d == S.pair_digest(p[rng.permutation(len(p))])
    q = p.copy()
    q[17, 1] += np.uint64(1)
    assert S.pair_digest(q) != d
    assert S.pair_digest(np.zeros((0, 2), np.uint64)) == (0, 0)
    # one key by hand: the splitmix64 finalizer of 1 << 32 | 2
    def mix(z):
        z = ((z ^ (z >> 30)) * 0xbf58476d1ce4e5b9) & (2**64 - 1)
        z = ((z ^ (z >> 27)) * 0x94d049bb133111eb) & (2**64 - 1)
        return z ^ (z >> 31)
    k = (1 << 32) | 2
    assert S.pair_digest(np.array([[1, 2]], np.uint64)) == (mix(k), mix(k ^ 0x9e3779b97f4a7c15))
