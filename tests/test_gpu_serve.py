"""The request server (continuous batching on one GPU).

Every request's result must be the same when it runs alone and when it runs
in a mix of concurrent requests (containers byte-identical, coordinates
bit-identical), and the same as the engine's direct call. Admission by the
HBM budget, request validation and failure isolation inside a batch are
checked too.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    from chaincompressor_amd import engine
    return engine.Engine(0)


def _mix(eng):
    from chaincompressor_amd import workloads
    rng = np.random.default_rng(7)
    reqs = []
    for k in range(14):
        n = int(rng.integers(1, 300_000)) if k % 5 else int(rng.integers(1, 50))
        dt = np.float32 if k % 4 else np.float64
        eb = [1e-2, 1e-3, 1e-4][k % 3]
        pts = workloads.gaussian_clusters(n, seed=100 + k, clusters=8).astype(dt)
        reqs.append(("c", pts, eb))
    for k in range(6):
        n = int(rng.integers(1000, 200_000))
        pts = workloads.gaussian_clusters(n, seed=200 + k, clusters=4)
        blob, _ = eng.compress(pts, 1e-3, 16, 2)
        reqs.append(("d", blob, np.float64 if k % 2 else np.float32))
    return reqs


def _submit(srv, r):
    if r[0] == "c":
        return srv.submit_compress(r[1], r[2], 16, 2)
    return srv.submit_decompress(r[1], dtype=r[2])


def test_alone_equals_mix(eng):
    from chaincompressor_amd import engine
    reqs = _mix(eng)
    srv = engine.Server(0, hbm_budget_bytes=16 << 30, workers=3)
    try:
        tickets = [_submit(srv, r) for r in reqs]          # all at once: a mix
        mixed = [srv.poll(t) for t in tickets]
        alone = []
        for r in reqs:                                     # one at a time
            alone.append(srv.poll(_submit(srv, r)))
    finally:
        srv.close()
    for r, m, a in zip(reqs, mixed, alone):
        assert m.status == 0 and a.status == 0, (m.error, a.error)
        assert a.batch == 1
        if r[0] == "c":
            assert m.data == a.data
            direct, _ = eng.compress(r[1], r[2], 16, 2)
            assert m.data == direct
            assert np.all(np.abs(eng.decompress(m.data) - r[1].astype(np.float64)) <= r[2])
        else:
            assert m.data.dtype == np.dtype(r[2])
            assert np.array_equal(m.data.view(np.uint8), a.data.view(np.uint8))
            direct = eng.decompress(r[1], dtype=r[2])
            assert np.array_equal(m.data.view(np.uint8), direct.view(np.uint8))
        assert m.t_submit <= m.t_start <= m.t_end


def test_requests_are_batched():
    from chaincompressor_amd import engine, workloads
    srv = engine.Server(0, hbm_budget_bytes=8 << 30, workers=1)
    try:
        frames = [workloads.gaussian_clusters(20_000 + 1000 * k, seed=k) for k in range(24)]
        tickets = [srv.submit_compress(f, 1e-3) for f in frames]
        res = [srv.poll(t) for t in tickets]
    finally:
        srv.close()
    assert all(r.status == 0 for r in res)
    assert max(r.batch for r in res) > 1


def test_admission_and_validation():
    from chaincompressor_amd import engine
    srv = engine.Server(0, hbm_budget_bytes=256 << 20, workers=2)   # 128 MB per worker
    try:
        big = np.zeros((1_000_000, 3), dtype=np.float64)
        with pytest.raises(engine.ChainError, match="budget"):
            srv.submit_compress(big, 1e-3)
        with pytest.raises(engine.ChainValueError):
            srv.submit_compress(np.zeros((10, 3), np.float32), 0.0)
        with pytest.raises(engine.ChainValueError):
            srv.submit_compress(np.zeros((10, 5), np.float32), 1e-3)
        ok = srv.poll(srv.submit_compress(np.random.default_rng(1).random((50_000, 3)), 1e-3))
        assert ok.status == 0 and ok.data[:4] == b"LCH1"
    finally:
        srv.close()


def test_failing_request_is_isolated():
    from chaincompressor_amd import engine, workloads
    srv = engine.Server(0, hbm_budget_bytes=8 << 30, workers=1)
    try:
        frames = [workloads.gaussian_clusters(30_000, seed=50 + k) for k in range(8)]
        frames[3] = frames[3].copy()
        frames[3][17, 1] = np.nan
        tickets = [srv.submit_compress(f, 1e-3) for f in frames]
        res = [srv.poll(t) for t in tickets]
    finally:
        srv.close()
    assert res[3].status != 0 and "finite" in res[3].error
    for k, r in enumerate(res):
        if k != 3:
            assert r.status == 0, r.error


def test_large_requests_stay_in_the_arenas(eng):
    """A decompress request of several million points, sized close to one
    worker's decompression arena, and a million-point compress request run
    through the server without any arena being reallocated (nothing is
    allocated while serving), and give the engine's direct results."""
    from chaincompressor_amd import engine, workloads
    n = 4_000_000
    pts = workloads.gaussian_clusters(n, seed=77)
    blob, _ = eng.compress(pts, 1e-3, 16, 2)
    srv = engine.Server(0, hbm_budget_bytes=5 << 29, workers=1)   # 2.5 GB: 640 MB decode arena
    try:
        a0 = srv.arena_allocations()
        assert a0 == 2
        small = pts[:1_000_000]
        tickets = [srv.submit_decompress(blob), srv.submit_compress(small, 1e-3, 16, 2), srv.submit_decompress(blob)]
        res = [srv.poll(t) for t in tickets]
        assert srv.arena_allocations() == a0
    finally:
        srv.close()
    assert all(r.status == 0 for r in res), [r.error for r in res]
    direct = eng.decompress(blob)
    assert np.array_equal(res[0].data.view(np.uint64), direct.view(np.uint64))
    assert np.array_equal(res[2].data.view(np.uint64), direct.view(np.uint64))
    assert res[1].data == eng.compress(small, 1e-3, 16, 2)[0]
