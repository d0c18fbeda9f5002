"""Float64 MD-like frames (fcc lattice + thermal noise in a periodic box,
atoms in ID order) at eb = 1e-4: batch compression and batch decompression
on the GPU, bit-exact against the CPU decoder (the oracle's restatement of
cpp/decompress/lcps_cpu_decompress_serial.cpp:1182-1338) on the same bytes,
and the LCP3 streams byte-identical to the CPU compressor's."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    from chaincompressor_amd import engine
    return engine.Engine(0)


def md_frames(sizes, seed):
    from chaincompressor_amd import workloads
    return [workloads.md_fcc_frame(n, np.random.default_rng([seed, k])) for k, n in enumerate(sizes)]


def test_md_batch_bit_exact_against_cpu_decoder(eng):
    frames = md_frames([100_000, 137_000, 250_001, 4_000, 180_000], seed=41)
    eb, p, t = 1e-4, 16, 2
    lch1, _ = eng.compress_batch(frames, eb, p, t, fmt="lch1")
    lcp3, _ = eng.compress_batch(frames, eb, p, t, fmt="lcp3")
    dec = eng.decompress_batch(lch1)
    dec3 = eng.decompress_batch(lcp3)
    for f, b, b3, d, d3 in zip(frames, lch1, lcp3, dec, dec3):
        print("frame", f.shape[0], flush=True)
        assert b3 == O.compress_streams(f, eb, p, t).lcp3()
        ref = O.decode(b)
        assert np.array_equal(d.view(np.uint64), ref.view(np.uint64))
        assert np.array_equal(d3.view(np.uint64), ref.view(np.uint64))
        assert np.all(np.abs(d - f) <= eb)


def test_md_batch_full_size_round_trip(eng):
    """Sixteen frames of up to 2M atoms: batch results equal per-frame
    results and every coordinate is within eb."""
    rng = np.random.default_rng(5)
    frames = md_frames(list(rng.integers(100_000, 2_000_001, size=16)), seed=43)
    outs, _ = eng.compress_batch(frames, 1e-4, 16, 2, fmt="lch1")
    decs = eng.decompress_batch(outs)
    for k in (0, 7, 15):
        single, _ = eng.compress(frames[k], 1e-4, 16, 2, fmt="lch1")
        assert single == outs[k]
        assert np.array_equal(eng.decompress(single).view(np.uint64), decs[k].view(np.uint64))
    for f, d in zip(frames, decs):
        assert np.all(np.abs(d - f) <= 1e-4)
