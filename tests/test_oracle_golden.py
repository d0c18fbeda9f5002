"""The CPU oracle against the golden vectors written by the compiled reference.

Pins: lcps_compress2 (cpp/chain/lcps_cpp.cpp:2017-2338) for ec=False (LCP3,
exact integer streams) and ec=True (LCP2, Zstd-22 + RVN1 chooser), its stats
list (:2331-2337), and lcps_decompress_serial
(cpp/decompress/lcps_cpu_decompress_serial.cpp:1182-1338).
"""
import hashlib

import numpy as np
import pytest

import oracle as O


def sha(b):
    return hashlib.sha256(bytes(b)).hexdigest()


def small_names(g):
    return [str(x) for x in g["names"]]


def params(g, name):
    eb, p, t, bl, l1, l2 = g[f"{name}__params"]
    return float(eb), int(p), int(t), bool(bl), bool(l1), bool(l2)


def test_small_cases_match_reference(golden_small):
    g = golden_small
    for name in small_names(g):
        eb, p, t, bl, l1, l2 = params(g, name)
        pts = g[f"{name}__points"]
        s = O.compress_streams(pts, eb, p, t, bl, l1, l2)
        lcp3 = s.lcp3()
        assert sha(lcp3) == str(g[f"{name}__lcp3_sha"]), name
        if f"{name}__lcp3" in g:
            assert lcp3 == g[f"{name}__lcp3"].tobytes(), name
        lcp2, st = s.lcp2()
        assert lcp2 == g[f"{name}__lcp2"].tobytes(), name
        assert st == [int(x) for x in g[f"{name}__stats"]], name
        dec = O.decode(lcp2)
        assert sha(dec.tobytes()) == str(g[f"{name}__decoded_sha"]), name
        dec3 = O.decode(lcp3)
        assert np.array_equal(dec3.view(np.uint64), dec.view(np.uint64)), name
        # decoding the reference's own LCP2 bytes
        dref = O.decode(g[f"{name}__lcp2"].tobytes())
        assert np.array_equal(dref.view(np.uint64), dec.view(np.uint64)), name


@pytest.mark.parametrize("name", ["dambreak_qb11", "dambreak_qb12", "dambreak_qb13", "dambreak_qb14",
                                  "cfg1_dambreak_like"])
def test_large_cases_match_reference(golden_large, name):
    g = golden_large
    if name.startswith("dambreak"):
        pts = g["dambreak_qb11__points"]
    else:
        from chaincompressor_amd import workloads
        pts = workloads.dambreak_like()
    assert sha(np.ascontiguousarray(pts).tobytes()) == str(g[f"{name}__points_sha"])
    eb, p, t = g[f"{name}__params"]
    s = O.compress_streams(pts, float(eb), int(p), int(t))
    lcp3 = s.lcp3()
    assert len(lcp3) == int(g[f"{name}__lcp3_len"]) and sha(lcp3) == str(g[f"{name}__lcp3_sha"])
    lcp2, st = s.lcp2()
    assert lcp2 == g[f"{name}__lcp2"].tobytes()
    assert st == [int(x) for x in g[f"{name}__stats"]]
    assert sha(O.decode(lcp2).tobytes()) == str(g[f"{name}__decoded_sha"])
