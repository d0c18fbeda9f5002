"""GPU decoder parity through the C ABI.

Decoded float64 coordinates must be bit-identical to the reference decoder
(lcps_cpu_decompress_serial, cpp/decompress/lcps_cpu_decompress_serial.cpp:1182-1338)
on the same integer streams: the golden digests of the compiled reference's
output, and the oracle's decoder on the same bytes. The GPU also decodes the
reference's own LCP3 streams from the golden set.
"""
import hashlib

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def sha(b):
    return hashlib.sha256(bytes(b)).hexdigest()


@pytest.fixture(scope="module")
def eng():
    from chaincompressor_amd import engine
    return engine.Engine(0)


def flags_of(bl, l1, l2):
    from chaincompressor_amd import engine
    return (0 if bl else engine.NO_BL) | (0 if l1 else engine.NO_L1) | (0 if l2 else engine.NO_L2)


def test_golden_roundtrip_bit_exact(eng, golden_small):
    g = golden_small
    bad = []
    for name in [str(x) for x in g["names"]]:
        eb, p, t, bl, l1, l2 = g[f"{name}__params"]
        pts = g[f"{name}__points"]
        print("case", name, flush=True)
        try:
            b, _ = eng.compress(pts, float(eb), int(p), int(t), flags=flags_of(bl, l1, l2), fmt="lch1")
            dec = eng.decompress(b)
        except Exception as exc:
            bad.append((name, str(exc)))
            continue
        if sha(dec.tobytes()) != str(g[f"{name}__decoded_sha"]):
            bad.append((name, "decoded"))
    assert not bad, bad


def test_golden_reference_lcp3_streams(eng, golden_small):
    """The reference's own LCP3 bytes, decoded on the GPU."""
    g = golden_small
    n = 0
    for name in [str(x) for x in g["names"]]:
        if f"{name}__lcp3" not in g:
            continue
        print("case", name, flush=True)
        dec = eng.decompress(g[f"{name}__lcp3"].tobytes())
        assert sha(dec.tobytes()) == str(g[f"{name}__decoded_sha"]), name
        n += 1
    assert n >= 8


def test_golden_reference_lcp2_streams(golden_small, golden_large):
    """The reference CPU compressor's LCP2 bytes through chain_caller's decoder
    (host LCP2 -> LCP3, then the GPU): bit-identical to lcps_decompress_serial."""
    pytest.importorskip("chaincompressor_amd.lcp2")
    from chaincompressor_amd import chain_caller
    for g in (golden_small, golden_large):
        for name in [str(x) for x in g["names"]]:
            print("case", name, flush=True)
            dec = chain_caller.chain_decompress(g[f"{name}__lcp2"].tobytes())
            assert sha(np.ascontiguousarray(dec).tobytes()) == str(g[f"{name}__decoded_sha"]), name


def test_golden_large_roundtrip(eng, golden_large):
    g = golden_large
    db = g["dambreak_qb11__points"]
    for name in [str(x) for x in g["names"]]:
        if name.startswith("dambreak"):
            pts = db
        else:
            from chaincompressor_amd import workloads
            pts = workloads.dambreak_like()
        eb, p, t = g[f"{name}__params"]
        print("case", name, flush=True)
        b, _ = eng.compress(pts, float(eb), int(p), int(t), fmt="lch1")
        dec = eng.decompress(b)
        assert sha(dec.tobytes()) == str(g[f"{name}__decoded_sha"]), name


@pytest.mark.parametrize("seed,n,d,dtype,eb,p,t", [
    (11, 1_000_000, 3, np.float32, 1e-3, 16, 2),
    (12, 300_000, 3, np.float64, 1e-4, 8, 0),
    (13, 200_000, 2, np.float32, 1e-3, 1, 5),
    (14, 90_000, 4, np.float32, 2e-3, 256, -2),
    (15, 120_000, 1, np.float64, 1e-5, 64, 3),
])
def test_random_roundtrip(eng, seed, n, d, dtype, eb, p, t):
    from chaincompressor_amd import workloads
    rng = np.random.default_rng(seed)
    pts = workloads.gaussian_clusters(n, seed=seed, clusters=16).astype(dtype) if d == 3 else \
        rng.random((n, d)).astype(dtype)
    print("case", seed, flush=True)
    b, _ = eng.compress(pts, eb, p, t, fmt="lch1")
    dec = eng.decompress(b)
    ref = O.decode(b)
    assert np.array_equal(dec.view(np.uint64), ref.view(np.uint64))
    assert np.all(np.abs(dec - pts.astype(np.float64)) <= eb)
    # float32 output is the float64 reconstruction rounded once
    dec32 = eng.decompress(b, dtype=np.float32)
    assert np.array_equal(dec32, dec.astype(np.float32))
    # SoA output is the transpose
    soa = eng.decompress(b, layout="soa")
    assert np.array_equal(soa, dec.T)
    # the LCP3 container of the same frame decodes to the same coordinates
    b3, _ = eng.compress(pts, eb, p, t, fmt="lcp3")
    assert np.array_equal(eng.decompress(b3).view(np.uint64), dec.view(np.uint64))


def test_batch_decode_matches_single(eng):
    from chaincompressor_amd import workloads
    frames = [workloads.gaussian_clusters(n, seed=80 + k, clusters=4) for k, n in
              enumerate([7, 50_000, 1, 4097, 200_000])]
    outs, _ = eng.compress_batch(frames, 1e-3, 16, 2, fmt="lch1")
    decs = eng.decompress_batch(outs)
    for o, dcode, f in zip(outs, decs, frames):
        single = eng.decompress(o)
        assert np.array_equal(single.view(np.uint64), dcode.view(np.uint64))
        assert np.all(np.abs(dcode - f.astype(np.float64)) <= 1e-3)


def test_malformed_streams_raise(eng):
    """Truncation, trailing bytes and corrupt headers are errors, never crashes."""
    from chaincompressor_amd import engine, workloads
    pts = workloads.gaussian_clusters(20_000, seed=3, clusters=4)
    b, _ = eng.compress(pts, 1e-3, 16, 2, fmt="lch1")
    for bad in (b[:-1], b + b"\0", b[:40], b"LCP2" + b[4:], b"XXXX" + b[4:]):
        with pytest.raises((engine.ChainError, engine.ChainValueError)):
            eng.decompress(bad)
    # flipped payload bytes: an error or coordinates, never a hang or fault
    arr = bytearray(b)
    for pos in (len(b) // 3, len(b) // 2, len(b) - 40):
        c = bytearray(arr)
        c[pos] ^= 0x5A
        try:
            eng.decompress(bytes(c))
        except (engine.ChainError, engine.ChainValueError):
            pass
    # the engine still works afterwards
    assert np.all(np.abs(eng.decompress(b) - pts) <= 1e-3)


def test_pybind_modules_drop_in(eng):
    """The reference's call signatures (chain_caller.py:33-40)."""
    from chaincompressor_amd import chain_caller, workloads
    pts = workloads.dambreak_like()
    enc = chain_caller.chain_gpu(pts.T, 1e-3, 16, 2, 0)
    dec = chain_caller.gpu_decompress(enc)
    assert dec.shape == pts.shape and dec.dtype == np.float64
    assert np.all(np.abs(dec - pts) <= 1e-3)
    ret = chain_caller.chain(pts, 1e-3, 16, 2, 0, ec=False)
    streams = O.compress_streams(pts, 1e-3, 16, 2)
    assert ret[0] == streams.lcp3()
    assert np.array_equal(chain_caller.chain_decompress(ret[0]), dec)
    # ec=True (the default) returns the reference's own container: LCP2 bytes
    # and stats of lcps_compress2
    ret2 = chain_caller.chain(pts, 1e-3, 16, 2, 0)
    b2, st2 = streams.lcp2()
    assert ret2[0] == b2 and ret2[1:] == st2
    # this engine's entropy-coded container on request
    ret3 = chain_caller.chain(pts, 1e-3, 16, 2, 0, container="lch1")
    assert ret3[0][:4] == b"LCH1"
    assert np.array_equal(chain_caller.chain_decompress(ret3[0]), dec)
