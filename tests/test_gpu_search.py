"""(p, threshold) search on the device: every configuration's container size
and chain count equal the oracle's for that configuration, host and device
inputs agree, and a sweep larger than one batch is split correctly."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def test_search_matches_oracle_per_configuration():
    from chaincompressor_amd import engine, workloads
    eng = engine.Engine(0)
    pts = workloads.dambreak_like()
    ps, ts = [1, 4, 16, 64, 256], [-2, 0, 2, 5]
    rows, best = eng.search(pts, 1e-3, ps, ts)
    assert len(rows) == len(ps) * len(ts)
    for p, t, nbytes, chained in rows:
        s = O.compress_streams(pts, 1e-3, p, t)
        assert nbytes == len(s.lch1()[0]), (p, t)
        assert chained == pts.shape[0] - s.blocks.shape[0], (p, t)
    assert best[2] == min(r[2] for r in rows)
    # device-resident input gives the same table
    ptr = eng.device_malloc(pts.nbytes)
    eng.to_device(ptr, pts)
    rows_dev, _ = eng.search(None, 1e-3, ps, ts, device=(ptr, pts.shape[0], 3, np.float32))
    eng.device_free(ptr)
    assert rows_dev == rows


def test_search_splits_large_sweeps():
    from chaincompressor_amd import engine, workloads
    eng = engine.Engine(0)
    pts = workloads.gaussian_clusters(3_000_000, seed=61)   # 2^27 / 3e6 -> 44 pairs per batch
    ps = [2, 4, 8, 16, 32, 64, 128]
    ts = list(range(-3, 4))
    rows, best = eng.search(pts, 1e-3, ps, ts)
    assert len(rows) == 49
    for p, t, nbytes, _ in (rows[0], rows[44], rows[48]):
        b, _ = eng.compress(pts, 1e-3, p, t)
        assert len(b) == nbytes
