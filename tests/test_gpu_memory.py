"""Arena sizing of the compressor.

A first compression sizes its buffers from the plan: correction lists of 256
values per tile and an output bounded by Fixed packing at the widths the plan
fixes. The server's per-request estimate must bound what that run carves; a
tile with more exact corrections makes the call run again with room for every
value, and the containers stay the oracle's either way.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    from chaincompressor_amd import engine
    e = engine.Engine(0)
    e.set_autotune(False)
    e.set_tuning(64, 65536, 1)
    return e


def _points(kind, n, d, dtype, seed):
    from chaincompressor_amd import workloads
    rng = np.random.default_rng(seed)
    if kind == "clusters" and d == 3:
        return workloads.gaussian_clusters(n, seed=seed, clusters=16).astype(dtype)
    if kind == "wide":
        return rng.uniform(-1e6, 1e6, size=(n, d)).astype(dtype)
    return rng.random((n, d)).astype(dtype)


# (kind, n, d, dtype, eb, p, threshold, flags)
CASES = [
    ("clusters", 1_000_000, 3, np.float32, 1e-3, 16, 2, 0),
    ("clusters", 400_000, 3, np.float32, 1e-4, 16, 2, 0),
    ("clusters", 300_000, 3, np.float64, 1e-5, 8, 0, 0),
    ("uniform", 150_000, 2, np.float32, 1e-3, 300, 2, 0),      # offsets of 4 bytes
    ("uniform", 100_000, 1, np.float64, 1e-6, 64, 3, 1),       # NO_BL
    ("uniform", 120_000, 4, np.float32, 2e-3, 64, -2, 4),      # NO_L2
    ("uniform", 90_000, 3, np.float32, 1e-3, 16, 2, 2),        # NO_L1
    ("wide", 30_000, 3, np.float64, 1e-12 * 1e6 / 2 ** 10, 1, 2, 0),   # ~50-bit values, many corrections
    ("wide", 40_000, 3, np.float64, 1e-6, 1, 2, 0),                     # ~40-bit values
]


@pytest.mark.parametrize("kind,n,d,dtype,eb,p,t,flags", CASES)
def test_estimate_bounds_the_first_run(eng, kind, n, d, dtype, eb, p, t, flags):
    """A second run happens exactly when a tile of 4096 points has more than
    256 exact corrections (counted by the oracle); the output bound never
    overflows. Without a second run, the estimate bounds the carve."""
    from chaincompressor_amd import engine
    pts = _points(kind, n, d, dtype, 40 + n % 97)
    print("case", kind, n, d, np.dtype(dtype).name, eb, p, flags, flush=True)
    ref = O.compress_streams(pts, eb, p, t, bl=not (flags & 1), l1=not (flags & 2), l2=not (flags & 4))
    per_tile = np.bincount((ref.corr_index // d) // 4096) if len(ref.corr_index) else np.zeros(1, np.int64)
    _, retries0 = eng.memory_stats()
    b, _ = eng.compress(pts, eb, p, t, flags=flags)
    need, retries = eng.memory_stats()
    assert retries - retries0 == int(per_tile.max() > 256)
    if retries == retries0:
        est = engine.serve_estimate_compress(n, d, dtype, p, flags, 65536)
        assert 0 < need <= est, (need, est)
    assert b == ref.lch1(64, 65536)[0]


def test_estimate_per_point():
    """float32 xyz at p = 16: at most 250 bytes per particle for large requests."""
    from chaincompressor_amd import engine
    for n in (1_000_000, 20_000_000, 50_000_000):
        for chunk in (8192, 65536):
            assert engine.serve_estimate_compress(n, 3, np.float32, 16, 0, chunk) <= 250 * n


def test_many_corrections_run_again(eng):
    """~1,800 exact corrections per 4096-point tile (eb near the spacing of
    float64 at 1e6): the first run overflows its lists, the second writes the
    oracle's containers."""
    rng = np.random.default_rng(7)
    pts = rng.uniform(1e6, 2e6, size=(20_000, 3))
    eb = 1e-10
    ref = O.compress_streams(pts, eb, 16, 2)
    assert len(ref.corr_index) > 4 * 256
    _, r0 = eng.memory_stats()
    b, st = eng.compress(pts, eb, 16, 2)
    _, r1 = eng.memory_stats()
    assert r1 == r0 + 1
    assert st.num_corrections == len(ref.corr_index)
    assert b == ref.lch1(64, 65536)[0]
    b3, _ = eng.compress(pts, eb, 16, 2, fmt="lcp3")
    assert b3 == ref.lcp3()
    _, r1 = eng.memory_stats()
    dec = eng.decompress(b)
    assert np.array_equal(dec.view(np.uint64), O.decode(b).view(np.uint64))
    assert np.all(np.abs(dec - pts) <= eb)
    # the same frame inside a batch of ordinary ones: one second run for the
    # batch, every container equal to the frame's own
    from chaincompressor_amd import workloads
    frames = [workloads.gaussian_clusters(50_000, seed=3, clusters=4).astype(np.float64), pts,
              workloads.gaussian_clusters(7, seed=4, clusters=2).astype(np.float64)]
    outs, _ = eng.compress_batch(frames, eb, 16, 2)
    _, r2 = eng.memory_stats()
    assert r2 == r1 + 1
    for f, o in zip(frames, outs):
        assert o == eng.compress(f, eb, 16, 2)[0]
