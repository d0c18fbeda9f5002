"""LCG2 test helpers: containers built by hand from the format of
cu/chain/lcps_gpu_cpp.cu:5-35, and a pure-Python restatement of its decoder
(cu/decompress/lcps_gpu_decompress_ans.cu:2345-2452, 2700-2900) with exact
float32 arithmetic, for small streams.
"""
import struct
from fractions import Fraction

import numpy as np

NAMES = (b"IDR0", b"DIF0", b"KEP0", b"CH20", b"REL0")


def f32_round(x: Fraction) -> np.float32:
    """x rounded to float32, to nearest with ties to even (normal and subnormal)."""
    if x == 0:
        return np.float32(0.0)
    sign = -1 if x < 0 else 1
    a = abs(x)
    e = a.numerator.bit_length() - a.denominator.bit_length()
    if Fraction(2) ** e > a:
        e -= 1
    e = max(e, -126)
    quantum = Fraction(2) ** (e - 23)
    m = a / quantum
    fl = m.numerator // m.denominator
    r = m - fl
    if r > Fraction(1, 2) or (r == Fraction(1, 2) and fl % 2 == 1):
        fl += 1
    return np.float32(sign * float(fl * quantum))


def f32_fma(a, b, c) -> np.float32:
    """fmaf(a, b, c): the exact a*b + c, rounded once."""
    return f32_round(Fraction(float(a)) * Fraction(float(b)) + Fraction(float(c)))


def coordinate(q: int, step: np.float32, minimum: np.float32) -> np.float32:
    """fmaf(float(q) + 0.5f, step, minimum) in round-to-nearest (:2605-2625)."""
    mid = f32_round(Fraction(float(f32_round(Fraction(q)))) + Fraction(1, 2))
    return f32_fma(mid, step, minimum)


def pack_msb(values, width: int) -> bytes:
    bits = "".join(format(int(v), "0%db" % width) for v in values)
    bits += "0" * (-len(bits) % 8)
    return bytes(int(bits[k:k + 8], 2) for k in range(0, len(bits), 8))


def _section(payload: bytes, codec: int, zstd) -> bytes:
    if codec == 0:
        return struct.pack("<BQ", 0, len(payload)) + payload
    limit = 65536
    chunks = [payload[k:k + limit] for k in range(0, len(payload), limit)]
    frames = [zstd(c) for c in chunks]
    head = struct.pack("<BQI", codec, len(payload), len(chunks))
    table = b"".join(struct.pack("<QQ", len(c), len(f)) for c, f in zip(chunks, frames))
    return head + table + b"".join(frames)


def component(values, dims: int, codec: int = 2, zstd=None) -> bytes:
    """Fixed component of an (n, dims) integer matrix (empty when n == 0)."""
    v = np.asarray(values, dtype=np.int64).reshape(-1, dims)
    n = v.shape[0]
    if n == 0:
        return struct.pack("<BIQB", 0, dims, 0, 0)
    head = struct.pack("<BIQ", 2, dims, n)
    payload = b""
    for d in range(dims):
        mn = int(v[:, d].min())
        width = max(1, int(int(v[:, d].max()) - mn).bit_length())
        part = pack_msb(v[:, d] - mn, width)
        head += struct.pack("<iBQ", mn, width, len(part))
        payload += part
    return head + struct.pack("<B", 1) + _section(payload, codec, zstd)


def container(dims, n, p, eb, mins, wrap_min, wrap_max, comps) -> bytes:
    """comps: five component bodies in the order IDR0, DIF0, KEP0, CH20, REL0."""
    out = b"LCG2" + struct.pack("<IIQifii", 1, dims, n, p, float(np.float32(eb)), wrap_min, wrap_max)
    out += b"".join(struct.pack("<f", float(np.float32(m))) for m in mins)
    out += struct.pack("<I", 5)
    for name, body in zip(NAMES, comps):
        out += name + struct.pack("<Q", len(body)) + body
    return out


def decode(idrel, diff, keep, choice2, rel, dims, p, eb, mins, wrap_min, wrap_max):
    """The reference decoder in Python: blocks from keep/diff (modular reset
    scan, :2560-2590), roots by following successors (:2592-2603), float32
    coordinates. Returns float32 (n, dims)."""
    n = len(idrel)
    rng = wrap_max - wrap_min + 1
    blocks = []
    di = ki = 0
    for i, ch in enumerate(choice2):
        if ch == 0:
            blocks.append([int(x) for x in keep[ki]])
            ki += 1
        else:
            assert i > 0
            cur = []
            for d in range(dims):
                s = int(diff[di][d])
                assert 0 <= s < rng
                corr = -(s // 2) - 1 if s & 1 else s // 2
                step = rng - corr if corr > 0 else -corr
                cur.append(wrap_min + ((blocks[-1][d] - wrap_min) + step) % rng)
            blocks.append(cur)
            di += 1
    pos = np.cumsum([1 if v == 0 else 0 for v in idrel]) - 1
    step32 = np.float32(2.0) * np.float32(eb)
    out = np.zeros((n, dims), dtype=np.float32)
    for i in range(n):
        r = i
        while idrel[r] != 0:
            r = r + int(idrel[r])
        blk = blocks[pos[r]]
        for d in range(dims):
            q = blk[d] * p + int(rel[i][d])
            assert 0 <= int(rel[i][d]) < p and q <= 2 ** 31 - 1
            out[i, d] = coordinate(q, step32, np.float32(mins[d]))
    return out
