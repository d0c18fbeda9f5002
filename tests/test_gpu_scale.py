"""Parity at scale: frames past 2^20 explicit records and 2^23 points, where
tiles, rANS chunks and look-back chains are numerous. The LCH1 container must
equal the oracle's and decode within eb, bit-exact with the oracle decoder."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    from chaincompressor_amd import engine
    return engine.Engine(0)


@pytest.mark.parametrize("n,eb,p,t", [(10_000_000, 1e-2, 16, 2), (10_000_000, 1e-3, 16, 2),
                                      (6_000_000, 1e-4, 4, 0)])
def test_large_frame_parity(eng, n, eb, p, t):
    from chaincompressor_amd import workloads
    pts = workloads.gaussian_clusters(n, seed=workloads.BASE_SEED + 2)
    print("case", n, eb, flush=True)
    b, st = eng.compress(pts, eb, p, t, fmt="lch1")
    s = O.compress_streams(pts, eb, p, t)
    ref, kinds = s.lch1()
    assert b == ref, kinds
    dec = eng.decompress(b)
    assert np.array_equal(dec.view(np.uint64), O.decode(b).view(np.uint64))
    assert np.all(np.abs(dec - pts) <= eb)
