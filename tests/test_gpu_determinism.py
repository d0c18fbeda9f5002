"""Run-to-run determinism of the compressor at a size with more than 2^20
explicit records (one-byte choice2 and rel components, many tiles, look-back
chains and rANS chunks). The same frame is compressed several times by one
engine with differently sized frames compressed in between, so the arena
holds other data from earlier calls each time, and once by a fresh engine.
Every container must be byte-identical, and equal to the oracle's."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def test_repeated_compression_is_byte_identical():
    from chaincompressor_amd import engine, workloads
    pts = workloads.gaussian_clusters(10_000_000, seed=workloads.BASE_SEED + 2)
    other = workloads.gaussian_clusters(12_000_000, seed=workloads.BASE_SEED + 7)
    small = workloads.gaussian_clusters(3_000_000, seed=workloads.BASE_SEED + 8)
    eng = engine.Engine(0)
    outs = []
    for k, filler in enumerate((None, other, small, other)):
        if filler is not None:
            eng.compress(filler, 1e-4, 4, 0, fmt="lch1")
        b, _ = eng.compress(pts, 1e-3, 16, 2, fmt="lch1")
        print("run", k, len(b), "bytes", flush=True)
        outs.append(b)
    eng.close()
    fresh = engine.Engine(0)
    outs.append(fresh.compress(pts, 1e-3, 16, 2, fmt="lch1")[0])
    fresh.close()
    assert all(o == outs[0] for o in outs[1:]), [len(o) for o in outs]
    ref, kinds = O.compress_streams(pts, 1e-3, 16, 2).lch1()
    assert outs[0] == ref, kinds
