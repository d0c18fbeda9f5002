"""Host-side LCP2 export against the compiled reference (CPU).

LCP3 streams -- the reference's own golden LCP3 bytes, and the oracle's
LCP3 of the large frames (pinned by their SHA-256) -- are converted and must
equal lcps_compress2(..., ec=True)'s LCP2 bytes and stats exactly. The golden
LCP2 bytes were written with libzstd 1.4.9; with another libzstd only the
decoded coordinates can be compared.
"""
import hashlib

import numpy as np
import pytest

import oracle as O

lcp2 = pytest.importorskip("chaincompressor_amd.lcp2")


def sha(b):
    return hashlib.sha256(bytes(b)).hexdigest()


@pytest.fixture(scope="module")
def same_zstd():
    try:
        return lcp2.zstd_version() == "1.4.9"
    except lcp2.Lcp2Error as e:
        pytest.skip(str(e))


def test_golden_small_lcp3_to_lcp2(golden_small, same_zstd):
    g = golden_small
    n = 0
    for name in [str(x) for x in g["names"]]:
        if f"{name}__lcp3" in g:
            lcp3 = g[f"{name}__lcp3"].tobytes()
        else:
            eb, p, t, bl, l1, l2 = g[f"{name}__params"]
            lcp3 = O.compress_streams(g[f"{name}__points"], float(eb), int(p), int(t), bool(bl), bool(l1),
                                      bool(l2)).lcp3()
            assert sha(lcp3) == str(g[f"{name}__lcp3_sha"]), name
        out, st = lcp2.to_lcp2(lcp3)
        if same_zstd:
            assert out == g[f"{name}__lcp2"].tobytes(), name
            assert st == [int(x) for x in g[f"{name}__stats"]], name
        assert sha(O.decode(out).tobytes()) == str(g[f"{name}__decoded_sha"]), name
        n += 1
    assert n == len(g["names"]) and n >= 20


@pytest.mark.parametrize("name", ["dambreak_qb11", "cfg1_dambreak_like"])
def test_golden_large_lcp3_to_lcp2(golden_large, same_zstd, name):
    g = golden_large
    if name.startswith("dambreak"):
        pts = g["dambreak_qb11__points"]
    else:
        from chaincompressor_amd import workloads
        pts = workloads.dambreak_like()
    eb, p, t = g[f"{name}__params"]
    lcp3 = O.compress_streams(pts, float(eb), int(p), int(t)).lcp3()
    assert sha(lcp3) == str(g[f"{name}__lcp3_sha"])
    out, st = lcp2.to_lcp2(lcp3)
    if same_zstd:
        assert out == g[f"{name}__lcp2"].tobytes()
        assert st == [int(x) for x in g[f"{name}__stats"]]
    assert sha(O.decode(out).tobytes()) == str(g[f"{name}__decoded_sha"])


def test_rejects_non_lcp3():
    with pytest.raises(lcp2.Lcp2Error):
        lcp2.to_lcp2(b"LCH1" + bytes(60))
    with pytest.raises(lcp2.Lcp2Error):
        lcp2.to_lcp2(b"LCP3")


def test_reference_lcp2_imports_to_lcp3(golden_small, golden_large):
    """The reference's own LCP2 bytes -> LCP3 of the same streams (compared
    with the reference's LCP3 bytes / digests) -> the reference's decoded output."""
    g = golden_small
    for name in [str(x) for x in g["names"]]:
        lcp3 = lcp2.to_lcp3(g[f"{name}__lcp2"].tobytes())
        assert sha(lcp3) == str(g[f"{name}__lcp3_sha"]), name
        assert sha(O.decode(lcp3).tobytes()) == str(g[f"{name}__decoded_sha"]), name
    gl = golden_large
    for name in [str(x) for x in gl["names"]]:
        lcp3 = lcp2.to_lcp3(gl[f"{name}__lcp2"].tobytes())
        assert sha(lcp3) == str(gl[f"{name}__lcp3_sha"]), name


def test_import_rejects_damaged_lcp2(golden_small):
    b = golden_small[f"{str(golden_small['names'][0])}__lcp2"].tobytes()
    for bad in (b[:-1], b + b"\0", b[:20], b"LCP3" + b[4:]):
        with pytest.raises(lcp2.Lcp2Error):
            lcp2.to_lcp3(bad)
