"""The autotuner: it searches chunk sizes, lane counts and the L1 strategy,
picks one of its trials, the containers written with the tuned setting equal
the oracle's LCH1 writer with the same RVC1 setting, and they decode
exactly. The L1 strategy changes no byte of the output."""
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
    best, trials = eng.autotune(ptrs, ns, np.float32, 3, 1e-3, 16, 2)
    lanes, chunk = best["lanes_max"], best["chunk_symbols"]
    assert best in [t for t, _, _ in trials]
    assert {t["chunk_symbols"] for t, _, _ in trials} == {8192, 16384, 32768, 65536}
    assert {t["lanes_max"] for t, _, _ in trials} == {32, 64} and {t["l1_tiles"] for t, _, _ in trials} == {0, 1}
    assert all(ms > 0 and b > 0 for _, ms, b in trials)
    smallest = min(b for _, _, b in trials)
    assert [b for t, _, b in trials if t == best][0] <= smallest * 1.02
    outs, _ = eng.compress_batch(frames, 1e-3, 16, 2)
    for f, b in zip(frames, outs):
        ref, _ = O.compress_streams(f, 1e-3, 16, 2).lch1(lanes_max=lanes, chunk_symbols=chunk)
        assert b == ref
        assert np.array_equal(eng.decompress(b).view(np.uint64), O.decode(b).view(np.uint64))
    # every knob setting writes the oracle's bytes for its RVC1 layout
    for t, _, _ in trials:
        eng.set_tuning(**t)
        outs, _ = eng.compress_batch(frames[:2], 1e-3, 16, 2)
        for f, b in zip(frames[:2], outs):
            assert b == O.compress_streams(f, 1e-3, 16, 2).lch1(t["lanes_max"], t["chunk_symbols"])[0], t
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
