"""ctypes wrapper of the CPU oracle (oracle/liboracle.so) -- TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import
this module, and only as the checker. The product path
(chaincompressor_amd) never imports it.

The oracle restates the reference CPU path:
  cpp/chain/lcps_cpp.cpp:2017-2338  (lcps_compress2: streams, LCP3/LCP2 writers)
  cpp/decompress/lcps_cpu_decompress_serial.cpp:1182-1338 (LCP2 decoder)
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class _Streams(C.Structure):
    _fields_ = [
        ("dims", C.c_uint32), ("n", C.c_uint64), ("p_eff", C.c_int64), ("eb", C.c_double),
        ("mins", C.POINTER(C.c_double)),
        ("idrel", C.POINTER(C.c_int64)),
        ("nblocks", C.c_uint64), ("blocks", C.POINTER(C.c_int64)),
        ("nchoice", C.c_uint64), ("choice2", C.POINTER(C.c_int64)),
        ("ndiff", C.c_uint64), ("diff", C.POINTER(C.c_int64)),
        ("nkeep", C.c_uint64), ("keep", C.POINTER(C.c_int64)),
        ("rel", C.POINTER(C.c_int64)),
        ("wrap_min", C.c_int64), ("wrap_max", C.c_int64),
        ("ncorr", C.c_uint64), ("corr_index", C.POINTER(C.c_uint64)), ("corr_value", C.POINTER(C.c_double)),
        ("max_bid", C.c_uint64),
    ]


class _RansParams(C.Structure):
    _fields_ = [("lanes_max", C.c_uint32), ("chunk_symbols", C.c_uint32)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run make -C oracle)")
        L = C.CDLL(path)
        L.oc_last_error.restype = C.c_char_p
        P = C.POINTER
        u8p = P(C.c_uint8)
        L.oc_compress_streams.argtypes = [P(C.c_double), C.c_uint64, C.c_uint32, C.c_double, C.c_int64,
                                          C.c_int64, C.c_int, C.c_int, C.c_int, P(_Streams)]
        L.oc_streams_free.argtypes = [P(_Streams)]
        L.oc_write_lcp3.argtypes = [P(_Streams), P(u8p), P(C.c_size_t)]
        L.oc_write_lcp2.argtypes = [P(_Streams), P(u8p), P(C.c_size_t), P(C.c_uint64)]
        L.oc_write_lch1.argtypes = [P(_Streams), P(_RansParams), P(u8p), P(C.c_size_t), P(C.c_uint32)]
        L.oc_decode.argtypes = [u8p, C.c_size_t, P(P(C.c_double)), P(C.c_uint64), P(C.c_uint32)]
        L.oc_write_lcg2.argtypes = [P(_Streams), C.c_float, P(C.c_float), P(u8p), P(C.c_size_t)]
        L.oc_decode_lcg2.argtypes = [u8p, C.c_size_t, P(P(C.c_float)), P(C.c_uint64), P(C.c_uint32)]
        L.oc_zstd_frame.argtypes = [u8p, C.c_size_t, P(u8p), P(C.c_size_t)]
        L.oc_pack_fixed.argtypes = [P(C.c_int64), C.c_uint64, C.c_uint32, P(C.c_int64), u8p, P(u8p), P(C.c_size_t)]
        L.oc_pack_tagged.argtypes = [P(C.c_int64), C.c_uint64, C.c_uint32, P(u8p), P(u8p), P(C.c_size_t),
                                     P(C.c_uint32)]
        L.oc_rvn1_encode.argtypes = [P(C.c_uint64), C.c_uint64, P(u8p), P(C.c_size_t)]
        L.oc_rvn1_decode.argtypes = [u8p, C.c_size_t, C.c_uint64, P(P(C.c_uint64))]
        L.oc_rvn1_table.argtypes = [P(C.c_uint64), C.c_uint64, P(P(C.c_uint32)), P(C.c_uint32), P(C.c_int)]
        L.oc_rvc1_encode.argtypes = [P(C.c_uint64), C.c_uint64, C.c_uint32, C.c_uint32, P(u8p), P(C.c_size_t)]
        L.oc_rvc1_decode.argtypes = [u8p, C.c_size_t, C.c_uint64, P(P(C.c_uint64))]
        L.oc_rvc1_lanes_for.argtypes = [C.c_uint64, C.c_uint32]
        L.oc_rvc1_lanes_for.restype = C.c_uint32
        L.oc_free.argtypes = [C.c_void_p]
        _LIB = L
    return _LIB


class OracleError(RuntimeError):
    pass


def _check(rc):
    if rc != 0:
        raise OracleError(lib().oc_last_error().decode())


def _take_bytes(ptr, n) -> bytes:
    try:
        return C.string_at(ptr, n) if n else b""
    finally:
        lib().oc_free(ptr)


def _arr(ptr, n, dtype, shape=None):
    if n == 0:
        return np.zeros(shape if shape is not None else (0,), dtype=dtype)
    a = np.ctypeslib.as_array(ptr, shape=(n,)).copy().astype(dtype, copy=False)
    return a.reshape(shape) if shape is not None else a


def _u8(b: bytes):
    buf = (C.c_uint8 * max(1, len(b))).from_buffer_copy(b if b else b"\0")
    return buf


@dataclass
class Streams:
    dims: int
    n: int
    p_eff: int
    eb: float
    mins: np.ndarray
    idrel: np.ndarray
    blocks: np.ndarray
    choice2: np.ndarray
    diff: np.ndarray
    keep: np.ndarray
    rel: np.ndarray
    wrap_min: int
    wrap_max: int
    corr_index: np.ndarray
    corr_value: np.ndarray
    max_bid: int
    _raw: object = None

    def __del__(self):
        if self._raw is not None:
            lib().oc_streams_free(C.byref(self._raw))
            self._raw = None

    def lcp3(self) -> bytes:
        out = C.POINTER(C.c_uint8)(); n = C.c_size_t()
        _check(lib().oc_write_lcp3(C.byref(self._raw), C.byref(out), C.byref(n)))
        return _take_bytes(out, n.value)

    def lcp2(self):
        out = C.POINTER(C.c_uint8)(); n = C.c_size_t(); st = (C.c_uint64 * 5)()
        _check(lib().oc_write_lcp2(C.byref(self._raw), C.byref(out), C.byref(n), st))
        b = _take_bytes(out, n.value)
        return b, [int(st[1]), int(st[2]), int(st[3]), int(st[4])]

    def lcg2(self, eb=None, mins=None) -> bytes:
        """LCG2 with Fixed components (the reference GPU compressor's container):
        these streams under a float32 error bound and float32 minima (by default
        the frame's, rounded to float32)."""
        eb32 = np.float32(self.eb if eb is None else eb)
        m = np.asarray(self.mins if mins is None else mins, dtype=np.float32)
        mp = (C.c_float * self.dims)(*[float(x) for x in m])
        out = C.POINTER(C.c_uint8)(); n = C.c_size_t()
        _check(lib().oc_write_lcg2(C.byref(self._raw), C.c_float(float(eb32)), mp, C.byref(out), C.byref(n)))
        return _take_bytes(out, n.value)

    def lch1(self, lanes_max=64, chunk_symbols=65536):
        out = C.POINTER(C.c_uint8)(); n = C.c_size_t(); kinds = (C.c_uint32 * 5)()
        rp = _RansParams(lanes_max, chunk_symbols)
        _check(lib().oc_write_lch1(C.byref(self._raw), C.byref(rp), C.byref(out), C.byref(n), kinds))
        return _take_bytes(out, n.value), [int(k) for k in kinds]


def compress_streams(points, eb, p, threshold, bl=True, l1=True, l2=True) -> Streams:
    pts = np.ascontiguousarray(np.asarray(points), dtype=np.float64)
    if pts.ndim != 2:
        raise ValueError("points must be (N, D)")
    n, d = pts.shape
    raw = _Streams()
    _check(lib().oc_compress_streams(pts.ctypes.data_as(C.POINTER(C.c_double)), n, d, float(eb), int(p),
                                     int(threshold), int(bl), int(l1), int(l2), C.byref(raw)))
    s = raw
    return Streams(
        dims=s.dims, n=s.n, p_eff=s.p_eff, eb=s.eb,
        mins=_arr(s.mins, d, np.float64),
        idrel=_arr(s.idrel, n, np.int64),
        blocks=_arr(s.blocks, s.nblocks * d, np.int64, (s.nblocks, d)),
        choice2=_arr(s.choice2, s.nchoice, np.int64),
        diff=_arr(s.diff, s.ndiff * d, np.int64, (s.ndiff, d)),
        keep=_arr(s.keep, s.nkeep * d, np.int64, (s.nkeep, d)),
        rel=_arr(s.rel, n * d, np.int64, (n, d)),
        wrap_min=s.wrap_min, wrap_max=s.wrap_max,
        corr_index=_arr(s.corr_index, s.ncorr, np.uint64),
        corr_value=_arr(s.corr_value, s.ncorr, np.float64),
        max_bid=s.max_bid, _raw=raw)


def compress_lcp2(points, eb, p, threshold, bl=True, l1=True, l2=True):
    """Equivalent of lcps_cpp.lcps_compress2(..., ec=True) -> [bytes, stats...]."""
    s = compress_streams(points, eb, p, threshold, bl, l1, l2)
    b, st = s.lcp2()
    return [b] + st


def decode(data: bytes) -> np.ndarray:
    out = C.POINTER(C.c_double)(); n = C.c_uint64(); d = C.c_uint32()
    buf = _u8(data)
    _check(lib().oc_decode(buf, len(data), C.byref(out), C.byref(n), C.byref(d)))
    try:
        return np.ctypeslib.as_array(out, shape=(n.value * d.value,)).copy().reshape(n.value, d.value)
    finally:
        lib().oc_free(out)


def decode_lcg2(data: bytes) -> np.ndarray:
    """lcg2::decode of the reference GPU decoder: float32 (N, D)."""
    out = C.POINTER(C.c_float)(); n = C.c_uint64(); d = C.c_uint32()
    buf = _u8(data)
    _check(lib().oc_decode_lcg2(buf, len(data), C.byref(out), C.byref(n), C.byref(d)))
    try:
        return np.ctypeslib.as_array(out, shape=(n.value * d.value,)).copy().reshape(n.value, d.value)
    finally:
        lib().oc_free(out)


def zstd_frame(data: bytes) -> bytes:
    """One Zstd frame (level 22) of data."""
    out = C.POINTER(C.c_uint8)(); n = C.c_size_t()
    _check(lib().oc_zstd_frame(_u8(data), len(data), C.byref(out), C.byref(n)))
    return _take_bytes(out, n.value)


def pack_fixed(values: np.ndarray):
    v = np.ascontiguousarray(values, dtype=np.int64)
    n, d = v.shape
    mins = (C.c_int64 * max(1, d))(); widths = (C.c_uint8 * max(1, d))()
    out = C.POINTER(C.c_uint8)(); ln = C.c_size_t()
    _check(lib().oc_pack_fixed(v.ctypes.data_as(C.POINTER(C.c_int64)), n, d, mins, widths, C.byref(out),
                               C.byref(ln)))
    return list(mins)[:d], list(widths)[:d], _take_bytes(out, ln.value)


def pack_tagged(values: np.ndarray):
    v = np.ascontiguousarray(values, dtype=np.int64)
    n, d = v.shape
    tags = C.POINTER(C.c_uint8)(); out = C.POINTER(C.c_uint8)(); ln = C.c_size_t(); mw = C.c_uint32()
    _check(lib().oc_pack_tagged(v.ctypes.data_as(C.POINTER(C.c_int64)), n, d, C.byref(tags), C.byref(out),
                                C.byref(ln), C.byref(mw)))
    t = np.frombuffer(_take_bytes(tags, n), dtype=np.uint8).copy()
    return t, _take_bytes(out, ln.value), int(mw.value)


def rvn1_encode(values) -> bytes:
    v = np.ascontiguousarray(values, dtype=np.uint64)
    out = C.POINTER(C.c_uint8)(); ln = C.c_size_t()
    _check(lib().oc_rvn1_encode(v.ctypes.data_as(C.POINTER(C.c_uint64)), v.size, C.byref(out), C.byref(ln)))
    return _take_bytes(out, ln.value)


def rvn1_decode(data: bytes, n: int) -> np.ndarray:
    out = C.POINTER(C.c_uint64)()
    _check(lib().oc_rvn1_decode(_u8(data), len(data), n, C.byref(out)))
    try:
        return np.ctypeslib.as_array(out, shape=(max(1, n),)).copy()[:n]
    finally:
        lib().oc_free(out)


def rvn1_table(values):
    v = np.ascontiguousarray(values, dtype=np.uint64)
    probs = C.POINTER(C.c_uint32)(); A = C.c_uint32(); pb = C.c_int()
    _check(lib().oc_rvn1_table(v.ctypes.data_as(C.POINTER(C.c_uint64)), v.size, C.byref(probs), C.byref(A),
                               C.byref(pb)))
    try:
        return np.ctypeslib.as_array(probs, shape=(A.value,)).copy(), int(pb.value)
    finally:
        lib().oc_free(probs)


def rvc1_encode(values, lanes, chunk_symbols) -> bytes:
    v = np.ascontiguousarray(values, dtype=np.uint64)
    out = C.POINTER(C.c_uint8)(); ln = C.c_size_t()
    _check(lib().oc_rvc1_encode(v.ctypes.data_as(C.POINTER(C.c_uint64)), v.size, lanes, chunk_symbols,
                                C.byref(out), C.byref(ln)))
    return _take_bytes(out, ln.value)


def rvc1_decode(data: bytes, n: int) -> np.ndarray:
    out = C.POINTER(C.c_uint64)()
    _check(lib().oc_rvc1_decode(_u8(data), len(data), n, C.byref(out)))
    try:
        return np.ctypeslib.as_array(out, shape=(max(1, n),)).copy()[:n]
    finally:
        lib().oc_free(out)


def rvc1_lanes_for(n: int, lanes_max: int) -> int:
    return int(lib().oc_rvc1_lanes_for(n, lanes_max))
