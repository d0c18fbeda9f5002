"""Read the layout of LCH1 / LCP3 containers without decoding them.

``describe(blob)`` returns the header fields and, per component (idrel, diff,
keep, choice2, rel), the encoding the compressor chose and the parameters of
its rANS sections: precision, lanes, chunk size and chunk count. It is a
host-side inspection helper; decoding runs on the GPU engine.

LCH1 layout (little-endian):
  "LCH1" u32 D u64 N i64 p f64 eb i64 wrap_min i64 wrap_max f64 mins[D]
  5 x (u64 length, component)      idrel, diff, keep, choice2, rel
  u64 corrections, corrections x (u64 index, f64 value)
component:
  "FIXP" u32 D u64 n  D x (i64 min, u8 width, u64 bytes)  payloads
  "TAGR" u32 D u64 n u8 maxw  u64 len RVC1(widths)  u64 len packed values
  "RANS" u32 D u64 n  u64 len RVC1(values)
RVC1 section:
  "RVC1" u8 pb u64 count u64 table_len table u8 lanes u32 chunk_symbols
  u32 nchunks  nchunks x u32 chunk bytes  chunks (lanes x u32 states + bytes)
LCP3 components are "RAW1" matrices of the reference (cpp/chain/lcps_cpp.cpp:1184-1211).
"""
from __future__ import annotations

import struct

COMPONENTS = ("idrel", "diff", "keep", "choice2", "rel")


class _R:
    def __init__(self, b: bytes, pos: int = 0, end: int | None = None):
        self.b, self.pos, self.end = b, pos, len(b) if end is None else end

    def take(self, n):
        if n > self.end - self.pos:
            raise ValueError("truncated container")
        v = self.b[self.pos:self.pos + n]
        self.pos += n
        return v

    def u(self, fmt):
        return struct.unpack("<" + fmt, self.take(struct.calcsize("<" + fmt)))[0]

    def sub(self, n):
        r = _R(self.b, self.pos, self.pos + n)
        self.take(n)
        return r


def _rvc1(r: _R) -> dict:
    if r.take(4) != b"RVC1":
        raise ValueError("invalid RVC1 magic")
    pb = r.u("B")
    count = r.u("Q")
    table_len = r.u("Q")
    r.take(table_len)
    lanes = r.u("B")
    chunk = r.u("I")
    nch = r.u("I")
    sizes = [r.u("I") for _ in range(nch)]
    return {"precision_bits": pb, "symbols": count, "table_bytes": table_len, "lanes": lanes,
            "chunk_symbols": chunk, "chunks": nch, "payload_bytes": sum(sizes),
            "state_bytes": 4 * lanes * nch}


def _component(r: _R) -> dict:
    magic = r.take(4)
    if magic == b"FIXP":
        d = r.u("I")
        n = r.u("Q")
        dims = [{"min": r.u("q"), "width": r.u("B"), "bytes": r.u("Q")} for _ in range(d)]
        return {"kind": "FIXP", "dims": d, "records": n, "widths": [x["width"] for x in dims],
                "payload_bytes": sum(x["bytes"] for x in dims)}
    if magic == b"TAGR":
        d = r.u("I")
        n = r.u("Q")
        maxw = r.u("B")
        tags = _rvc1(r.sub(r.u("Q")))
        packed = r.u("Q")
        return {"kind": "TAGR", "dims": d, "records": n, "max_width": maxw, "widths": tags,
                "packed_bytes": packed}
    if magic == b"RANS":
        d = r.u("I")
        n = r.u("Q")
        return {"kind": "RANS", "dims": d, "records": n, "values": _rvc1(r.sub(r.u("Q")))}
    if magic == b"RAW1":
        return {"kind": "RAW1"}
    raise ValueError(f"unknown component magic {magic!r}")


def describe(blob: bytes) -> dict:
    r = _R(bytes(blob))
    magic = r.take(4)
    if magic not in (b"LCH1", b"LCP3", b"LCP2"):
        raise ValueError("not a Chain container")
    d = r.u("I")
    n = r.u("Q")
    p = r.u("q")
    eb = r.u("d")
    wmin = r.u("q")
    wmax = r.u("q")
    mins = [r.u("d") for _ in range(d)]
    out = {"magic": magic.decode(), "dims": d, "points": n, "p": p, "eb": eb, "wrap": (wmin, wmax),
           "mins": mins, "bytes": len(blob), "components": {}}
    if magic != b"LCH1":
        return out
    for name in COMPONENTS:
        ln = r.u("Q")
        c = _component(r.sub(ln))
        c["size"] = ln
        out["components"][name] = c
    out["corrections"] = r.u("Q")
    return out
