"""GPU entropy stage parity: the LCH1 container written on the device (RVN1
tables built on the GPU, chunked lane-interleaved rANS, Fixed/Tagged
packing, chooser, layout) must be byte-identical to the oracle's LCH1 writer
for the same RVC1 parameters, through the C ABI."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    from chaincompressor_amd import engine
    return engine.Engine(0)


def flags_of(bl, l1, l2):
    from chaincompressor_amd import engine
    return (0 if bl else engine.NO_BL) | (0 if l1 else engine.NO_L1) | (0 if l2 else engine.NO_L2)


def test_golden_inputs_lch1(eng, golden_small):
    g = golden_small
    eng.set_rans(64, 65536)
    bad = []
    for name in [str(x) for x in g["names"]]:
        eb, p, t, bl, l1, l2 = g[f"{name}__params"]
        pts = g[f"{name}__points"]
        print("case", name, flush=True)
        try:
            b, st = eng.compress(pts, float(eb), int(p), int(t), flags=flags_of(bl, l1, l2), fmt="lch1")
        except Exception as exc:
            bad.append((name, str(exc)))
            continue
        ref, kinds = O.compress_streams(pts, float(eb), int(p), int(t), bl, l1, l2).lch1(64, 65536)
        if b != ref:
            bad.append((name, "bytes", kinds))
    assert not bad, bad


@pytest.mark.parametrize("lanes,chunk", [(64, 65536), (16, 4096), (1, 1 << 20), (32, 1000), (4, 64)])
def test_rans_parameters(eng, lanes, chunk):
    from chaincompressor_amd import workloads
    pts = workloads.gaussian_clusters(150_000, seed=21, clusters=12)
    eng.set_rans(lanes, chunk)
    try:
        b, st = eng.compress(pts, 1e-3, 16, 2, fmt="lch1")
    finally:
        eng.set_rans(64, 65536)
    ref, kinds = O.compress_streams(pts, 1e-3, 16, 2).lch1(lanes, chunk)
    assert b == ref, kinds


@pytest.mark.parametrize("seed,n,d,dtype,eb,p,t", [
    (1, 400_000, 3, np.float32, 1e-3, 16, 2),
    (2, 250_000, 3, np.float64, 1e-4, 8, 0),
    (3, 100_000, 2, np.float32, 1e-3, 1, 5),
    (4, 60_000, 4, np.float32, 2e-3, 256, -2),
    (5, 90_000, 1, np.float64, 1e-5, 64, 3),
    # p not a power of two: quantised values below 2^32, then above it
    (6, 120_000, 3, np.float32, 1e-3, 10, 2),
    (7, 80_000, 2, np.float64, 1e-9, 12, 1),
    (8, 50_000, 3, np.float64, 1e-11, 7, 0),
])
def test_random_lch1(eng, seed, n, d, dtype, eb, p, t):
    from chaincompressor_amd import workloads
    rng = np.random.default_rng(seed)
    pts = workloads.gaussian_clusters(n, seed=seed, clusters=16).astype(dtype) if d == 3 else \
        rng.random((n, d)).astype(dtype)
    print("case", seed, flush=True)
    b, st = eng.compress(pts, eb, p, t, fmt="lch1")
    ref, kinds = O.compress_streams(pts, eb, p, t).lch1()
    assert b == ref, kinds
    # the container decodes (oracle decoder) to coordinates within eb
    dec = O.decode(b)
    assert np.all(np.abs(dec - pts.astype(np.float64)) <= eb)


def test_batch_lch1_matches_single(eng):
    from chaincompressor_amd import workloads
    frames = [workloads.gaussian_clusters(n, seed=60 + k, clusters=4) for k, n in
              enumerate([3, 9000, 120_000, 1, 4096, 70_001])]
    outs, _ = eng.compress_batch(frames, 1e-3, 16, 2, fmt="lch1")
    for f, o in zip(frames, outs):
        single, _ = eng.compress(f, 1e-3, 16, 2, fmt="lch1")
        assert o == single
        assert o == O.compress_streams(f, 1e-3, 16, 2).lch1()[0]


@pytest.mark.parametrize("seed,d,chunk,mixed", [(70, 4, 65536, False), (71, 3, 8192, False), (72, 1, 16384, False),
                                                 (73, 2, 4096, False), (70, 4, 8192, True), (71, 3, 16384, True),
                                                 (72, 1, 65536, True), (73, 2, 8192, True)])
def test_wide_values_lch1(eng, seed, d, chunk, mixed):
    """Quantised values of about 50 bits (p = 1, eb ~ 1e-9 over a range of 2e6).
    Uniform points make the Fixed payloads 45-50 bits wide; with most points
    packed next to the minimum (mixed) the kept blocks go Tagged with widths
    up to 50, for D up to 4 (the widest Tagged tiles)."""
    rng = np.random.default_rng(seed)
    n = 30_000
    pts = rng.uniform(-1e6, 1e6, size=(n, d))
    pts[: n // 3] = np.round(pts[: n // 3], 3)       # a share of short differences
    if mixed:
        far = rng.random(n) < 0.1
        pts[~far] = -1e6 + rng.uniform(0, 1e-3, size=(int((~far).sum()), d))
    eb = 1e-12 * 1e6 / 2 ** 10
    print("case", seed, d, chunk, mixed, flush=True)
    eng.set_rans(64, chunk)
    try:
        b, st = eng.compress(pts, eb, 1, 2, fmt="lch1")
    finally:
        eng.set_rans(64, 65536)
    ref, kinds = O.compress_streams(pts, eb, 1, 2).lch1(64, chunk)
    assert b == ref, kinds
    want = O.decode(b)
    got = eng.decompress(b)
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64))
    assert np.all(np.abs(got - pts) <= eb)
