"""LCG2 (the reference GPU compressor's container) on the CPU: the oracle's
decoder against containers built by hand from the format description
(cu/chain/lcps_gpu_cpp.cu:5-35) and against a pure-Python restatement with
exact float32 arithmetic; the host inflation of chaincompressor_amd.lcp2.
"""
from fractions import Fraction

import numpy as np
import pytest

import oracle as O
import lcg2_util as U


def _hand_built(codec=2):
    """Six points in two dimensions, p = 4, eb = 0.125: points 0 and 2 follow
    successors (1 and 4), points 1, 3, 4, 5 are explicit; records 0 and 2 are
    kept blocks, records 1 and 3 wrapped differences."""
    z = O.zstd_frame
    idrel = [1, 0, 2, 0, 0, 0]
    blocks = [[3, 7], [5, 6], [5, 9], [2, 8]]
    choice2 = [0, 1, 0, 1]
    wmin, wmax = 2, 9          # range 8: corrections in [-4, 3]
    # symbol = ZigZag(previous - current)
    diff = [[3, 2], [6, 2]]    # (3,7)->(5,6): (-2, 1); (5,9)->(2,8): (3, 1)
    keep = [blocks[0], blocks[2]]
    rel = [[1, 2], [0, 3], [2, 2], [3, 0], [1, 1], [0, 0]]
    comps = [U.component(idrel, 1, codec, z), U.component(diff, 2, codec, z), U.component(keep, 2, codec, z),
             U.component(choice2, 1, codec, z), U.component(rel, 2, codec, z)]
    mins = [np.float32(-1.5), np.float32(10.0)]
    b = U.container(2, 6, 4, 0.125, mins, wmin, wmax, comps)
    # expected: q = block(root) * 4 + rel, x = (q + 0.5) * 0.25 + min (exact here)
    root_block = [blocks[0], blocks[0], blocks[2], blocks[1], blocks[2], blocks[3]]
    want = np.array([[(bk[d] * 4 + rel[i][d] + 0.5) * 0.25 + float(mins[d]) for d in range(2)]
                     for i, bk in enumerate(root_block)], dtype=np.float32)
    return b, want, (idrel, diff, keep, choice2, rel, mins, wmin, wmax)


def test_f32_rounding_helper():
    """The exact helper against float32 arithmetic where the result is known."""
    assert U.f32_round(Fraction(1, 3)) == np.float32(1.0) / np.float32(3.0)
    assert U.f32_round(Fraction(16777217)) == np.float32(16777216)      # tie to even
    assert U.f32_round(Fraction(16777219)) == np.float32(16777220)      # tie to even (up)
    assert U.f32_round(Fraction(2) ** -149) == np.float32(2.0 ** -149)  # smallest subnormal
    # (1 + 2^-12)^2 = 1 + 2^-11 + 2^-24, a tie in float32 that rounds to even:
    # fmaf keeps the 2^-24 that the product rounded first loses
    a, c = np.float32(1 + 2 ** -12), np.float32(-(1 + 2 ** -11))
    assert U.f32_fma(a, a, c) == np.float32(2.0 ** -24)
    assert np.float32(a * a) + c == np.float32(0.0)


@pytest.mark.parametrize("codec", [2, 0])
def test_hand_built_container(codec):
    b, want, _ = _hand_built(codec)
    got = O.decode_lcg2(b)
    assert got.dtype == np.float32 and got.shape == (6, 2)
    assert np.array_equal(got, want)


def test_python_restatement_matches_oracle():
    """Streams of the oracle's compressor, stored as LCG2 with float32 eb and
    minima: the oracle's decoder equals the Python restatement bit for bit
    (q from the same integer streams, then fmaf in round-to-nearest)."""
    from chaincompressor_amd import workloads
    for seed, n, d, eb, p in ((1, 1500, 3, 1e-3, 16), (2, 900, 2, 7e-4, 5), (3, 700, 1, 3e-3, 64),
                              (4, 600, 4, 2e-3, 8)):
        rng = np.random.default_rng(seed)
        pts = workloads.gaussian_clusters(n, seed=seed, clusters=4) if d == 3 else rng.random((n, d))
        s = O.compress_streams(pts, eb, p, 2)
        b = s.lcg2()
        got = O.decode_lcg2(b)
        mins = np.asarray(s.mins, dtype=np.float32)
        want = U.decode(s.idrel, s.diff, s.keep, s.choice2, s.rel, d, s.p_eff, np.float32(eb), mins,
                        s.wrap_min, s.wrap_max)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), seed


def test_roundtrip_error():
    """float32 reconstruction: within eb plus float32 rounding of the coordinate."""
    from chaincompressor_amd import workloads
    pts = workloads.gaussian_clusters(200_000, seed=9, clusters=8)
    s = O.compress_streams(pts, 1e-3, 16, 2)
    x = O.decode_lcg2(s.lcg2()).astype(np.float64)
    assert np.all(np.abs(x - pts) <= 1e-3 + 4 * np.spacing(np.float32(1.0)))


def test_inflate_stores_every_section():
    from chaincompressor_amd import lcp2
    from chaincompressor_amd import workloads
    pts = workloads.gaussian_clusters(150_000, seed=4, clusters=8)
    b = O.compress_streams(pts, 1e-3, 16, 2).lcg2()
    inf = lcp2.inflate_lcg2(b)
    assert inf[:4] == b"LCG2" and inf != b
    assert lcp2.inflate_lcg2(inf) == inf                  # already inflated: unchanged
    assert np.array_equal(O.decode_lcg2(inf), O.decode_lcg2(b))
    hb, _, _ = _hand_built(2)
    assert lcp2.inflate_lcg2(hb) == _hand_built(0)[0]      # Zstd sections -> stored ones, all else equal


def test_malformed_and_ans_refused():
    from chaincompressor_amd import lcp2
    b, _, (idrel, diff, keep, choice2, rel, mins, wmin, wmax) = _hand_built(2)
    bad = [b[:-1], b + b"\0", b"LCG3" + b[4:], b[:4] + b"\x02" + b[5:], b[:40]]
    for x in bad:
        with pytest.raises(RuntimeError):
            O.decode_lcg2(x)
        with pytest.raises(lcp2.Lcp2Error):
            lcp2.inflate_lcg2(x)
    # a Tagged component: ANS tags (codec 1) then Zstd values -> unsupported
    import struct
    tagged = struct.pack("<BIQQB", 1, 2, 6, 12, 2) + struct.pack("<BQI", 1, 6, 1) + struct.pack("<QQ", 6, 4) + b"\0" * 4
    comps = [U.component(idrel, 1, 2, O.zstd_frame), U.component(diff, 2, 2, O.zstd_frame),
             U.component(keep, 2, 2, O.zstd_frame), U.component(choice2, 1, 2, O.zstd_frame), tagged]
    t = U.container(2, 6, 4, 0.125, mins, wmin, wmax, comps)
    with pytest.raises(lcp2.Lcg2Unsupported):
        lcp2.inflate_lcg2(t)
    with pytest.raises(RuntimeError, match="ANS"):
        O.decode_lcg2(t)
