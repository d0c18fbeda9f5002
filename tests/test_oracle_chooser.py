"""The LCH1 chooser rule, pinned on its own.

For every component of a container the three candidates are sized here
independently of the chooser: Fixed from the per-dimension minimum and width
(closed form), Tagged from the per-record widths (closed form) plus the size of
their RVC1 section, Raw-rANS from the RVC1 section of the values (eligible only
when every value is < 4096).

The rule: each candidate is first estimated (a RVC1 section as
sum(count * (pb - log2 f)) bits in 1/65536, log2 taken from the exponent and a
table of round(65536 log2(1 + k/256)) over the 8 bits below the leading one,
plus its exact header, table, chunk size words and lane states; Fixed exact).
The candidates within 1/256 + 64 bytes of the smallest estimate are the
contenders, and the container must hold the smallest contender, ties going to
Tagged, then Fixed, then Raw-rANS (the tie order of
cpp/chain/lcps_cpp.cpp:1140-1174), at that contender's exact size. The table
and the estimate are recomputed here from math.log2, so a mistake in the
oracle's (or the GPU's, which must equal the oracle's bytes) fails here.
"""
import math
import struct

import numpy as np
import pytest

import oracle as O

LANES, CHUNK = 64, 65536


def bitlen(a):
    a = np.asarray(a, dtype=np.uint64)
    out = np.zeros(a.shape, dtype=np.int64)
    v = a.copy()
    while np.any(v):
        nz = v != 0
        out[nz] += 1
        v[nz] >>= np.uint64(1)
    return out


LOG2_FRAC = [round(65536 * math.log2(1 + k / 256)) for k in range(256)]


def log2_q16(f):
    e = int(f).bit_length() - 1
    k = (f >> (e - 8)) & 255 if e >= 8 else (f << (8 - e)) & 255
    return (e << 16) + LOG2_FRAC[k]


def section_estimate(values, lanes_max):
    v = np.asarray(values, dtype=np.uint64).reshape(-1)
    n = v.size
    if n == 0:
        return 30
    probs, pb = O.rvn1_table(v)
    sec = O.rvc1_encode(v, 1, CHUNK)            # for the table length only
    table_len = struct.unpack_from("<Q", sec, 13)[0]
    counts = np.bincount(v.astype(np.int64), minlength=len(probs))
    bits = sum(int(c) * ((pb << 16) - log2_q16(int(probs[s]))) for s, c in enumerate(counts) if c)
    nchunks = -(-n // CHUNK)
    lanes = O.rvc1_lanes_for(n, lanes_max)
    return 30 + table_len + 4 * nchunks + 4 * lanes * nchunks + ((bits + (8 << 16) - 1) >> 19)


def candidate_estimates(values):
    v = np.asarray(values, dtype=np.int64)
    n, dims = v.shape
    tags = np.maximum(1, bitlen(np.bitwise_or.reduce(v, axis=1) if n else np.zeros(0, np.int64)))
    tagr = 4 + 4 + 8 + 1 + 8 + section_estimate(tags, LANES) + 8 + (int(tags.sum()) * dims + 7) // 8
    rans = None
    if n == 0 or int(v.max()) < 4096:
        rans = 4 + 4 + 8 + 8 + section_estimate(v.reshape(-1), LANES)
    return {"TAGR": tagr, "RANS": rans}


def contenders(est, fixp):
    vals = [est["TAGR"], fixp] + ([est["RANS"]] if est["RANS"] is not None else [])
    best = min(vals)
    lim = best + (best >> 8) + 64
    out = [k for k, e in (("TAGR", est["TAGR"]), ("FIXP", fixp), ("RANS", est["RANS"])) if e is not None and e <= lim]
    return out


def candidate_sizes(values):
    v = np.asarray(values, dtype=np.int64)
    n, dims = v.shape
    # Fixed: magic, dims, n, then per dimension (i64 minimum, u8 width, u64 bytes) and the payloads
    fixp = 16 + 17 * dims
    for d in range(dims):
        if n:
            w = max(1, int(bitlen(int(v[:, d].max() - v[:, d].min()))))
            fixp += (n * w + 7) // 8
    # Tagged: magic, dims, n, u8 max width, u64 + RVC1 section of widths, u64 + payload
    tags = np.maximum(1, bitlen(np.bitwise_or.reduce(v, axis=1) if n else np.zeros(0, np.int64)))
    sec = O.rvc1_encode(tags, O.rvc1_lanes_for(n, LANES), CHUNK)
    tagr = 4 + 4 + 8 + 1 + 8 + len(sec) + 8 + (int(tags.sum()) * dims + 7) // 8
    # Raw-rANS: magic, dims, n, u64 + RVC1 section of the point-major values
    rans = None
    if n == 0 or int(v.max()) < 4096:
        flat = v.reshape(-1)
        rans = 4 + 4 + 8 + 8 + len(O.rvc1_encode(flat, O.rvc1_lanes_for(flat.size, LANES), CHUNK))
    return {"TAGR": tagr, "FIXP": fixp, "RANS": rans}


def components_of(blob):
    dims = struct.unpack_from("<I", blob, 4)[0]
    pos = 48 + 8 * dims
    out = []
    for _ in range(5):
        (ln,) = struct.unpack_from("<Q", blob, pos)
        out.append((blob[pos + 8:pos + 12].decode(), ln))
        pos += 8 + ln
    return out


def matrices(s):
    return [s.idrel.reshape(-1, 1), s.diff, s.keep, s.choice2.reshape(-1, 1), s.rel]


def check(s):
    blob, kinds = s.lch1(LANES, CHUNK)
    for comp, (m, (magic, ln)) in enumerate(zip(matrices(s), components_of(blob))):
        sizes = candidate_sizes(m)
        order = contenders(candidate_estimates(m), sizes["FIXP"])
        best = min(order, key=lambda k: (sizes[k], ("TAGR", "FIXP", "RANS").index(k)))
        assert magic == best, (comp, magic, order, sizes)
        assert ln == sizes[best], (comp, ln, sizes)
    return kinds


def test_chooser_on_golden_inputs(golden_small):
    g = golden_small
    seen = set()
    for name in [str(x) for x in g["names"]]:
        eb, p, t, bl, l1, l2 = g[f"{name}__params"]
        s = O.compress_streams(g[f"{name}__points"], float(eb), int(p), int(t), bl, l1, l2)
        seen.update(check(s))
    assert seen == {0, 1, 2}   # every kind wins somewhere


@pytest.mark.parametrize("eb,p,t", [(1e-2, 16, 2), (1e-3, 16, 2), (1e-4, 4, 0), (1e-3, 300, -2)])
def test_chooser_on_cluster_frames(eb, p, t):
    from chaincompressor_amd import workloads
    check(O.compress_streams(workloads.gaussian_clusters(120_000, seed=31, clusters=10), eb, p, t))
