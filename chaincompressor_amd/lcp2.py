"""Host-side LCP2 export (libchain_lcp2.so, include/chain_lcp2.h).

``to_lcp2(lcp3)`` turns the raw container that the GPU engine writes with
``fmt="lcp3"`` (byte-identical to lcps_compress2(..., ec=False)) into the
reference's entropy-coded LCP2 (lcps_compress2(..., ec=True)): the Zstd-22 /
RVN1 chooser over every component (cpp/chain/lcps_cpp.cpp:1140-1174). The
result is readable by the reference's CPU decoder; with the reference's
libzstd the bytes are identical to its own. ``to_lcp3(lcp2)`` goes the other
way, so LCP2 files written by the reference's CPU compressor decode on the GPU.
``inflate_lcg2(lcg2)`` prepares streams of the reference's GPU compressor
(LCG2): their Zstd sections are inflated on the host into stored sections,
which the GPU engine decodes; sections coded with nvCOMP ANS are refused.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libchain_lcp2.so")
_LIB = None


class Lcp2Error(RuntimeError):
    pass


class Lcg2Unsupported(Lcp2Error):
    """An LCG2 component coded with nvCOMP ANS (Tagged or Raw encoding)."""


def _lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise Lcp2Error(f"{LIB_PATH} is missing: build it with __graft_entry__.build() (needs libzstd)")
        L = C.CDLL(LIB_PATH)
        L.chain_lcp2_from_lcp3.argtypes = [C.c_char_p, C.c_size_t, C.c_int, C.POINTER(C.POINTER(C.c_uint8)),
                                           C.POINTER(C.c_size_t), C.POINTER(C.c_uint64), C.c_char_p, C.c_size_t]
        L.chain_lcp3_from_lcp2.argtypes = [C.c_char_p, C.c_size_t, C.POINTER(C.POINTER(C.c_uint8)),
                                           C.POINTER(C.c_size_t), C.c_char_p, C.c_size_t]
        L.chain_lcg2_inflate.argtypes = [C.c_char_p, C.c_size_t, C.POINTER(C.POINTER(C.c_uint8)),
                                         C.POINTER(C.c_size_t), C.c_char_p, C.c_size_t]
        L.chain_lcp2_free.argtypes = [C.c_void_p]
        L.chain_lcp2_zstd_version.restype = C.c_char_p
        _LIB = L
    return _LIB


def zstd_version() -> str:
    return _lib().chain_lcp2_zstd_version().decode()


def to_lcp2(lcp3: bytes, zstd_level: int = 22):
    """LCP3 bytes -> (LCP2 bytes, [block_replacement_size, idrel_size, rel_size, num_chained])."""
    L = _lib()
    data = bytes(lcp3)
    out = C.POINTER(C.c_uint8)()
    n = C.c_size_t()
    st = (C.c_uint64 * 4)()
    err = C.create_string_buffer(256)
    rc = L.chain_lcp2_from_lcp3(data, len(data), zstd_level, C.byref(out), C.byref(n), st, err, 256)
    if rc != 0:
        raise Lcp2Error(err.value.decode())
    try:
        return C.string_at(out, n.value), [int(x) for x in st]
    finally:
        L.chain_lcp2_free(out)


def to_lcp3(lcp2: bytes) -> bytes:
    """LCP2 bytes (as lcps_compress2(..., ec=True) writes them) -> the LCP3 of
    the same integer streams, which the GPU engine decodes."""
    L = _lib()
    data = bytes(lcp2)
    out = C.POINTER(C.c_uint8)()
    n = C.c_size_t()
    err = C.create_string_buffer(256)
    rc = L.chain_lcp3_from_lcp2(data, len(data), C.byref(out), C.byref(n), err, 256)
    if rc != 0:
        raise Lcp2Error(err.value.decode())
    try:
        return C.string_at(out, n.value)
    finally:
        L.chain_lcp2_free(out)


def inflate_lcg2(lcg2: bytes) -> bytes:
    """LCG2 bytes of the reference GPU compressor -> the same container with
    every Zstd section stored (chain_lcg2_inflate, include/chain_lcp2.h)."""
    L = _lib()
    data = bytes(lcg2)
    out = C.POINTER(C.c_uint8)()
    n = C.c_size_t()
    err = C.create_string_buffer(256)
    rc = L.chain_lcg2_inflate(data, len(data), C.byref(out), C.byref(n), err, 256)
    if rc == 5:
        raise Lcg2Unsupported(err.value.decode())
    if rc != 0:
        raise Lcp2Error(err.value.decode())
    try:
        return C.string_at(out, n.value)
    finally:
        L.chain_lcp2_free(out)
