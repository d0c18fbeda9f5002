"""The entropy-stage autotuner: it picks one of its candidate chunk sizes,
the containers written with the tuned setting equal the oracle's LCH1
writer with the same setting, and they decode exactly."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def test_autotune_picks_a_candidate_and_stays_exact():
    from chaincompressor_amd import engine, workloads
    eng = engine.Engine(0)
    frames = workloads.uneven_batch(frames=6, lo=2e4, hi=6e5, seed=77)
    big = np.ascontiguousarray(np.concatenate(frames))
    ptr = eng.device_malloc(big.nbytes)
    eng.to_device(ptr, big)
    ns = [f.shape[0] for f in frames]
    offs = np.concatenate([[0], np.cumsum(ns)[:-1]])
    ptrs = [ptr + int(o) * 12 for o in offs]
    (lanes, chunk), trials = eng.autotune(ptrs, ns, np.float32, 3, 1e-3, 16, 2)
    assert lanes == 64 and chunk in (8192, 16384, 32768, 65536)
    assert all(ms > 0 and b > 0 for _, ms, b in trials)
    best = min(b for _, _, b in trials)
    assert dict((c, b) for c, _, b in trials)[chunk] <= best * 1.02
    outs, _ = eng.compress_batch(frames, 1e-3, 16, 2)
    for f, b in zip(frames, outs):
        ref, _ = O.compress_streams(f, 1e-3, 16, 2).lch1(lanes_max=lanes, chunk_symbols=chunk)
        assert b == ref
        assert np.array_equal(eng.decompress(b).view(np.uint64), O.decode(b).view(np.uint64))
    eng.device_free(ptr)


def test_automatic_mode_round_trip():
    from chaincompressor_amd import engine, workloads
    eng = engine.Engine(0)
    eng.set_autotune(True)
    frames = [workloads.gaussian_clusters(n, seed=90 + n % 7) for n in (30_000, 250_000, 31_000)]
    for f in frames:
        b, _ = eng.compress(f, 1e-3, 16, 2)
        d = eng.decompress(b)
        assert np.array_equal(d.view(np.uint64), O.decode(b).view(np.uint64))
        assert np.all(np.abs(d - f) <= 1e-3)
