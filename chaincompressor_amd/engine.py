"""ctypes binding of the C ABI in include/chain_engine.h (libchain_engine.so).

This is the product path: every call runs the HIP kernels of
chaincompressor_amd/csrc on the GPU. There is no CPU fallback; a missing
library or GPU raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libchain_engine.so")

CHAIN_F32, CHAIN_F64 = 0, 1
CHAIN_AOS, CHAIN_SOA = 0, 1
FMT_LCH1, FMT_LCP3 = 0, 1
NO_BL, NO_L1, NO_L2 = 1, 2, 4

_FORMATS = {"lch1": FMT_LCH1, "lcp3": FMT_LCP3}


class ChainError(RuntimeError):
    pass


class ChainValueError(ValueError):
    pass


class _Buf(C.Structure):
    _fields_ = [("data", C.POINTER(C.c_uint8)), ("size", C.c_size_t)]


class _Frame(C.Structure):
    _fields_ = [("points", C.c_void_p), ("n", C.c_uint64), ("eb", C.c_double), ("p", C.c_int64),
                ("threshold", C.c_int64), ("flags", C.c_uint32)]


class _Stats(C.Structure):
    _fields_ = [("bytes", C.c_uint64), ("block_replacement_size", C.c_uint64), ("idrel_size", C.c_uint64),
                ("rel_size", C.c_uint64), ("num_chained", C.c_uint64), ("num_corrections", C.c_uint64)]

    def as_list(self):
        """The tail of lcps_compress2's return list (cpp/chain/lcps_cpp.cpp:2331-2337)."""
        return [int(self.block_replacement_size), int(self.idrel_size), int(self.rel_size), int(self.num_chained)]


EXPORTS = ["chain_engine_create", "chain_engine_destroy", "chain_last_error", "chain_engine_stream",
           "chain_buf_free", "chain_compress", "chain_compress_batch", "chain_compress_batch_device",
           "chain_engine_set_rans", "chain_decompress", "chain_decompress_batch", "chain_decompress_batch_device",
           "chain_device_malloc", "chain_device_free", "chain_copy", "chain_last_stage_times",
           "chain_last_stage_bytes", "chain_engine_set_timing", "chain_server_create", "chain_server_destroy",
           "chain_server_last_error", "chain_serve_submit", "chain_serve_poll", "chain_server_clock",
           "chain_engine_set_graphs", "chain_engine_graph_stats", "chain_engine_autotune",
           "chain_engine_set_autotune", "chain_compress_search", "chain_server_stats", "chain_engine_set_tuning",
           "chain_engine_get_tuning",
           "chain_decompress_batch_noalloc", "chain_serve_estimate_compress", "chain_engine_memory_stats"]

CHAIN_PENDING = 5
REQ_COMPRESS, REQ_DECOMPRESS = 0, 1

class _Tuning(C.Structure):
    _fields_ = [("lanes_max", C.c_uint32), ("chunk_symbols", C.c_uint32), ("l1_tiles", C.c_uint32),
                ("reserved", C.c_uint32)]

    def as_dict(self):
        return {"lanes_max": int(self.lanes_max), "chunk_symbols": int(self.chunk_symbols),
                "l1_tiles": int(self.l1_tiles)}


class _Request(C.Structure):
    _fields_ = [("kind", C.c_int), ("points", C.c_void_p), ("dtype", C.c_int), ("layout", C.c_int),
                ("dims", C.c_uint32), ("n", C.c_uint64), ("eb", C.c_double), ("p", C.c_int64),
                ("threshold", C.c_int64), ("flags", C.c_uint32), ("format", C.c_int),
                ("bytes", C.c_void_p), ("len", C.c_size_t), ("out_dtype", C.c_int), ("out_layout", C.c_int)]


class _Result(C.Structure):
    _fields_ = [("status", C.c_int), ("out", _Buf), ("n", C.c_uint64), ("dims", C.c_uint32), ("batch", C.c_uint32),
                ("t_submit", C.c_double), ("t_start", C.c_double), ("t_end", C.c_double),
                ("error", C.c_char * 256)]


_LIB = None


def load_library():
    """Load libchain_engine.so; raises if it was not built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise ChainError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    P = C.POINTER
    L.chain_engine_create.argtypes = [C.c_int, C.c_uint64, P(C.c_void_p)]
    L.chain_engine_destroy.argtypes = [C.c_void_p]
    L.chain_last_error.argtypes = [C.c_void_p]
    L.chain_last_error.restype = C.c_char_p
    L.chain_engine_stream.argtypes = [C.c_void_p]
    L.chain_engine_stream.restype = C.c_void_p
    L.chain_buf_free.argtypes = [P(_Buf)]
    L.chain_compress.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_uint32, C.c_uint64, C.c_double,
                                 C.c_int64, C.c_int64, C.c_uint32, C.c_int, P(_Buf), P(_Stats)]
    L.chain_compress_batch.argtypes = [C.c_void_p, P(_Frame), C.c_uint32, C.c_int, C.c_int, C.c_uint32, C.c_int,
                                       C.c_int, P(_Buf), P(_Stats)]
    L.chain_compress_batch_device.argtypes = [C.c_void_p, P(_Frame), C.c_uint32, C.c_int, C.c_int, C.c_uint32,
                                              P(C.c_void_p), P(C.c_uint64), P(C.c_uint64), P(_Stats)]
    L.chain_engine_set_rans.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32]
    L.chain_decompress.argtypes = [C.c_void_p, P(C.c_uint8), C.c_size_t, C.c_int, C.c_int, P(_Buf),
                                   P(C.c_uint64), P(C.c_uint32)]
    L.chain_decompress_batch.argtypes = [C.c_void_p, P(P(C.c_uint8)), P(C.c_size_t), C.c_uint32, C.c_int, C.c_int,
                                         P(_Buf), P(C.c_uint64), P(C.c_uint32)]
    L.chain_decompress_batch_device.argtypes = [C.c_void_p, P(P(C.c_uint8)), P(C.c_size_t), C.c_void_p, C.c_uint32,
                                                C.c_int, C.c_int, P(C.c_void_p), P(C.c_uint64), P(C.c_uint64),
                                                P(C.c_uint32)]
    L.chain_decompress_batch_noalloc.argtypes = [C.c_void_p, P(P(C.c_uint8)), P(C.c_size_t), C.c_void_p, C.c_uint32,
                                                 C.c_int, C.c_int, C.c_void_p, C.c_uint64, P(C.c_uint64),
                                                 P(C.c_uint64), P(C.c_uint32)]
    L.chain_device_malloc.argtypes = [C.c_void_p, C.c_uint64, P(C.c_void_p)]
    L.chain_device_free.argtypes = [C.c_void_p, C.c_void_p]
    L.chain_copy.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_int]
    L.chain_last_stage_times.argtypes = [C.c_void_p, P(C.c_float), C.c_int, C.c_char_p, C.c_size_t]
    L.chain_last_stage_bytes.argtypes = [C.c_void_p, P(C.c_uint64), C.c_int]
    L.chain_engine_set_timing.argtypes = [C.c_void_p, C.c_int]
    L.chain_engine_set_graphs.argtypes = [C.c_void_p, C.c_int]
    L.chain_engine_set_autotune.argtypes = [C.c_void_p, C.c_int]
    L.chain_compress_search.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_uint32, C.c_uint64, C.c_double,
                                        P(C.c_int64), C.c_uint32, P(C.c_int64), C.c_uint32, C.c_uint32, C.c_int,
                                        P(_Stats)]
    L.chain_engine_autotune.argtypes = [C.c_void_p, P(_Frame), C.c_uint32, C.c_int, C.c_int, C.c_uint32, C.c_int,
                                        P(_Tuning), P(_Tuning), P(C.c_float), P(C.c_uint64), C.c_uint32,
                                        P(C.c_uint32)]
    L.chain_engine_set_tuning.argtypes = [C.c_void_p, P(_Tuning)]
    L.chain_engine_get_tuning.argtypes = [C.c_void_p, P(_Tuning)]
    L.chain_engine_graph_stats.argtypes = [C.c_void_p, P(C.c_uint64), P(C.c_uint64)]
    L.chain_server_create.argtypes = [C.c_int, C.c_uint64, C.c_uint32, C.c_uint32, P(C.c_void_p)]
    L.chain_server_destroy.argtypes = [C.c_void_p]
    L.chain_server_last_error.argtypes = [C.c_void_p]
    L.chain_server_last_error.restype = C.c_char_p
    L.chain_serve_submit.argtypes = [C.c_void_p, P(_Request), P(C.c_uint64)]
    L.chain_serve_poll.argtypes = [C.c_void_p, C.c_uint64, C.c_int, P(_Result)]
    L.chain_server_clock.argtypes = [C.c_void_p]
    L.chain_server_clock.restype = C.c_double
    L.chain_server_stats.argtypes = [C.c_void_p, P(C.c_uint64)]
    L.chain_serve_estimate_compress.argtypes = [C.c_uint64, C.c_uint32, C.c_int, C.c_int64, C.c_uint32, C.c_uint32]
    L.chain_serve_estimate_compress.restype = C.c_uint64
    L.chain_engine_memory_stats.argtypes = [C.c_void_p, P(C.c_uint64), P(C.c_uint64)]
    _LIB = L
    return L


def serve_estimate_compress(n, dims, dtype, p, flags=0, chunk_symbols=65536) -> int:
    """Device bytes the server reserves for one compress request (an upper
    bound of the first run's arena carve)."""
    return int(load_library().chain_serve_estimate_compress(int(n), int(dims), _dtype_code(dtype), int(p),
                                                            int(flags), int(chunk_symbols)))


def _take(buf: _Buf) -> bytes:
    try:
        return C.string_at(buf.data, buf.size) if buf.size else b""
    finally:
        load_library().chain_buf_free(C.byref(buf))


class _Owned:
    """Releases a library-allocated buffer when the last array viewing it dies."""

    def __init__(self, buf: _Buf):
        self.buf = buf

    def __del__(self):
        load_library().chain_buf_free(C.byref(self.buf))


def _take_array(buf: _Buf, dtype, shape) -> np.ndarray:
    """A numpy view of a library-allocated coordinate buffer (no copy); the
    buffer goes back to the library when the array is garbage-collected."""
    dt = np.dtype(dtype)
    if buf.size == 0:
        load_library().chain_buf_free(C.byref(buf))
        return np.empty(shape, dtype=dt)
    owner = _Owned(buf)
    raw = (C.c_uint8 * buf.size).from_address(C.addressof(buf.data.contents))
    raw._owner = owner
    return np.frombuffer(raw, dtype=dt).reshape(shape)


def _dtype_code(dt) -> int:
    dt = np.dtype(dt)
    if dt == np.float32:
        return CHAIN_F32
    if dt == np.float64:
        return CHAIN_F64
    raise ChainValueError("points must be float32 or float64")


class Engine:
    """One GPU, one HIP stream, one reusable device arena."""

    def __init__(self, device: int = 0, hbm_budget_bytes: int = 0):
        L = load_library()
        h = C.c_void_p()
        rc = L.chain_engine_create(device, hbm_budget_bytes, C.byref(h))
        if rc != 0:
            raise ChainError(L.chain_last_error(None).decode())
        self._h = h
        self._L = L

    def close(self):
        if getattr(self, "_h", None):
            self._L.chain_engine_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def _check(self, rc):
        if rc == 0:
            return
        msg = self._L.chain_last_error(self._h).decode()
        if rc == 1:
            raise ChainValueError(msg)
        raise ChainError(msg)

    @property
    def stream(self) -> int:
        return int(self._L.chain_engine_stream(self._h) or 0)

    def set_rans(self, lanes_max=64, chunk_symbols=65536):
        """RVC1 lane count and chunk size of the containers this engine writes."""
        self._check(self._L.chain_engine_set_rans(self._h, lanes_max, chunk_symbols))

    def set_timing(self, on: bool):
        """Per-stage HIP event timing (on by default)."""
        self._check(self._L.chain_engine_set_timing(self._h, 1 if on else 0))

    def search(self, points, eb, ps, thresholds, flags=0, layout="aos", device=None):
        """Container size for every (p, threshold) pair of one frame, on the
        device. points: host array, or device=(ptr, n, dims, dtype).
        Returns [(p, threshold, bytes, num_chained)] and the smallest entry."""
        ps = [int(x) for x in ps]
        ts = [int(x) for x in thresholds]
        pa = (C.c_int64 * len(ps))(*ps)
        ta = (C.c_int64 * len(ts))(*ts)
        st = (_Stats * (len(ps) * len(ts)))()
        if device is None:
            pts = np.ascontiguousarray(points)
            n, d = pts.shape if layout == "aos" else pts.shape[::-1]
            ptr, dt, on_dev = pts.ctypes.data, pts.dtype, 0
        else:
            ptr, n, d, dt = device
            on_dev = 1
        self._check(self._L.chain_compress_search(self._h, C.c_void_p(ptr), _dtype_code(dt),
                                                  CHAIN_AOS if layout == "aos" else CHAIN_SOA, d, n, float(eb), pa,
                                                  len(ps), ta, len(ts), int(flags), on_dev, st))
        rows = [(ps[k // len(ts)], ts[k % len(ts)], int(st[k].bytes), int(st[k].num_chained))
                for k in range(len(ps) * len(ts))]
        return rows, min(rows, key=lambda r: r[2])

    def set_autotune(self, on: bool):
        """Tune the entropy stage's chunking once per batch size bucket."""
        self._check(self._L.chain_engine_set_autotune(self._h, 1 if on else 0))

    def autotune(self, ptrs, ns, dtype, dims, eb, p, threshold, on_device=True, layout="aos"):
        """Tune on a batch (device pointers, or host arrays with on_device=False);
        returns (best tuning dict, [(tuning dict, ms, bytes)] of every trial)."""
        F = len(ptrs)
        fr = (_Frame * F)()
        for k in range(F):
            fr[k] = _Frame(ptrs[k], int(ns[k]), float(eb), int(p), int(threshold), 0)
        best = _Tuning()
        cap = 16
        tr = (_Tuning * cap)()
        ms = (C.c_float * cap)()
        by = (C.c_uint64 * cap)()
        nt = C.c_uint32()
        self._check(self._L.chain_engine_autotune(self._h, fr, F, _dtype_code(dtype),
                                                  CHAIN_AOS if layout == "aos" else CHAIN_SOA, dims,
                                                  1 if on_device else 0, C.byref(best), tr, ms, by, cap,
                                                  C.byref(nt)))
        trials = [(tr[k].as_dict(), float(ms[k]), int(by[k])) for k in range(nt.value)]
        return best.as_dict(), trials

    def set_tuning(self, lanes_max=64, chunk_symbols=65536, l1_tiles=1):
        """RVC1 layout and L1 strategy of the compressions that follow."""
        t = _Tuning(lanes_max, chunk_symbols, 1 if l1_tiles else 0, 0)
        self._check(self._L.chain_engine_set_tuning(self._h, C.byref(t)))

    def tuning(self) -> dict:
        t = _Tuning()
        self._check(self._L.chain_engine_get_tuning(self._h, C.byref(t)))
        return t.as_dict()

    def set_graphs(self, on: bool):
        """Capture compression launches in a HIP graph and replay it for
        batches of the same launch shape."""
        self._check(self._L.chain_engine_set_graphs(self._h, 1 if on else 0))

    def memory_stats(self):
        """(arena bytes the last compression carved, compressions run again with worst-case bounds)."""
        need, retries = C.c_uint64(), C.c_uint64()
        self._check(self._L.chain_engine_memory_stats(self._h, C.byref(need), C.byref(retries)))
        return int(need.value), int(retries.value)

    def graph_stats(self):
        c, r = C.c_uint64(), C.c_uint64()
        self._check(self._L.chain_engine_graph_stats(self._h, C.byref(c), C.byref(r)))
        return int(c.value), int(r.value)

    def compress_batch_device(self, ptrs, ns, dtype, dims, eb, p, threshold, flags=0, layout="aos"):
        """Device-resident LCH1 batch: returns (device pointer, offsets, sizes, stats)."""
        F = len(ptrs)
        ebs = eb if isinstance(eb, (list, tuple, np.ndarray)) else [eb] * F
        ps = p if isinstance(p, (list, tuple, np.ndarray)) else [p] * F
        ts = threshold if isinstance(threshold, (list, tuple, np.ndarray)) else [threshold] * F
        fl = flags if isinstance(flags, (list, tuple, np.ndarray)) else [flags] * F
        fr = (_Frame * F)()
        for k in range(F):
            fr[k] = _Frame(ptrs[k], int(ns[k]), float(ebs[k]), int(ps[k]), int(ts[k]), int(fl[k]))
        dptr = C.c_void_p()
        offs = (C.c_uint64 * F)()
        sizes = (C.c_uint64 * F)()
        sts = (_Stats * F)()
        rc = self._L.chain_compress_batch_device(self._h, fr, F, _dtype_code(dtype),
                                                 CHAIN_AOS if layout == "aos" else CHAIN_SOA, dims, C.byref(dptr),
                                                 offs, sizes, sts)
        self._check(rc)
        return int(dptr.value or 0), list(offs), list(sizes), [sts[k] for k in range(F)]

    def compress(self, points, eb, p, threshold, flags=0, fmt="lch1", layout="aos"):
        """Compress one frame (host array, (N, D) for AoS or (D, N) for SoA)."""
        pts = np.asarray(points)
        dt = _dtype_code(pts.dtype)
        pts = np.ascontiguousarray(pts)
        if pts.ndim != 2:
            raise ChainValueError("points must be a 2-D array")
        if layout == "aos":
            n, d = pts.shape
        else:
            d, n = pts.shape
        out = _Buf()
        st = _Stats()
        rc = self._L.chain_compress(self._h, pts.ctypes.data_as(C.c_void_p), dt,
                                    CHAIN_AOS if layout == "aos" else CHAIN_SOA, d, n, float(eb), int(p),
                                    int(threshold), int(flags), _FORMATS[fmt], C.byref(out), C.byref(st))
        self._check(rc)
        return _take(out), st

    def compress_batch(self, frames, eb, p, threshold, flags=0, fmt="lch1", layout="aos", device_ptrs=None):
        """Compress a list of host arrays (or device pointers with device_ptrs=(ptrs, ns, dtype, dims))."""
        if device_ptrs is None:
            arrs = [np.ascontiguousarray(np.asarray(f)) for f in frames]
            if not arrs:
                raise ChainValueError("empty batch")
            dt = _dtype_code(arrs[0].dtype)
            dims = arrs[0].shape[1] if layout == "aos" else arrs[0].shape[0]
            ns = [a.shape[0] if layout == "aos" else a.shape[1] for a in arrs]
            ptrs = [a.ctypes.data for a in arrs]
            on_dev = 0
        else:
            ptrs, ns, dtype, dims = device_ptrs
            dt = _dtype_code(dtype)
            arrs = None
            on_dev = 1
        F = len(ptrs)
        ebs = eb if isinstance(eb, (list, tuple, np.ndarray)) else [eb] * F
        ps = p if isinstance(p, (list, tuple, np.ndarray)) else [p] * F
        ts = threshold if isinstance(threshold, (list, tuple, np.ndarray)) else [threshold] * F
        fl = flags if isinstance(flags, (list, tuple, np.ndarray)) else [flags] * F
        fr = (_Frame * F)()
        for k in range(F):
            fr[k] = _Frame(ptrs[k], int(ns[k]), float(ebs[k]), int(ps[k]), int(ts[k]), int(fl[k]))
        outs = (_Buf * F)()
        sts = (_Stats * F)()
        rc = self._L.chain_compress_batch(self._h, fr, F, dt, CHAIN_AOS if layout == "aos" else CHAIN_SOA, dims,
                                          on_dev, _FORMATS[fmt], outs, sts)
        self._check(rc)
        del arrs
        return [_take(outs[k]) for k in range(F)], [sts[k] for k in range(F)]

    def decompress(self, data: bytes, dtype=np.float64, layout="aos"):
        data = bytes(data)
        buf = C.cast(C.c_char_p(data), C.POINTER(C.c_uint8))   # no copy; data stays alive below
        out = _Buf()
        n = C.c_uint64()
        d = C.c_uint32()
        rc = self._L.chain_decompress(self._h, buf, len(data), _dtype_code(dtype),
                                      CHAIN_AOS if layout == "aos" else CHAIN_SOA, C.byref(out), C.byref(n),
                                      C.byref(d))
        self._check(rc)
        return _take_array(out, dtype, (n.value, d.value) if layout == "aos" else (d.value, n.value))

    def decompress_batch(self, streams, dtype=np.float64, layout="aos"):
        k = len(streams)
        bufs = [bytes(s) for s in streams]   # kept alive for the call; pointers need no copy
        ptrs = (C.POINTER(C.c_uint8) * k)(*[C.cast(C.c_char_p(b), C.POINTER(C.c_uint8)) for b in bufs])
        lens = (C.c_size_t * k)(*[len(s) for s in streams])
        outs = (_Buf * k)()
        ns = (C.c_uint64 * k)()
        ds = (C.c_uint32 * k)()
        rc = self._L.chain_decompress_batch(self._h, ptrs, lens, k, _dtype_code(dtype),
                                            CHAIN_AOS if layout == "aos" else CHAIN_SOA, outs, ns, ds)
        self._check(rc)
        return [_take_array(outs[j], dtype, (ns[j], ds[j]) if layout == "aos" else (ds[j], ns[j]))
                for j in range(k)]

    def decompress_batch_device(self, streams, d_streams=None, dtype=np.float64, layout="aos"):
        """Decode streams (host copies plan the work; d_streams is a device
        pointer to the same bytes concatenated). Returns (device pointer,
        element offsets, ns, dims); the output stays on the device."""
        k = len(streams)
        bufs = [bytes(s) for s in streams]   # kept alive for the call; pointers need no copy
        ptrs = (C.POINTER(C.c_uint8) * k)(*[C.cast(C.c_char_p(b), C.POINTER(C.c_uint8)) for b in bufs])
        lens = (C.c_size_t * k)(*[len(s) for s in streams])
        dout = C.c_void_p()
        offs = (C.c_uint64 * k)()
        ns = (C.c_uint64 * k)()
        ds = (C.c_uint32 * k)()
        rc = self._L.chain_decompress_batch_device(self._h, ptrs, lens, C.c_void_p(d_streams or 0), k,
                                                   _dtype_code(dtype), CHAIN_AOS if layout == "aos" else CHAIN_SOA,
                                                   C.byref(dout), offs, ns, ds)
        self._check(rc)
        return int(dout.value or 0), list(offs), list(ns), int(ds[0]) if k else 0

    def decompress_batch_noalloc(self, streams, d_streams, d_out, out_capacity, dtype=np.float64, layout="aos"):
        """Decode into the caller's device buffer d_out without any device
        allocation (intermediates in the library's fixed scratch). Returns
        (element offsets, ns, dims)."""
        k = len(streams)
        bufs = [bytes(s) for s in streams]
        ptrs = (C.POINTER(C.c_uint8) * k)(*[C.cast(C.c_char_p(b), C.POINTER(C.c_uint8)) for b in bufs])
        lens = (C.c_size_t * k)(*[len(s) for s in streams])
        offs = (C.c_uint64 * k)()
        ns = (C.c_uint64 * k)()
        ds = (C.c_uint32 * k)()
        rc = self._L.chain_decompress_batch_noalloc(self._h, ptrs, lens, C.c_void_p(d_streams), k, _dtype_code(dtype),
                                                    CHAIN_AOS if layout == "aos" else CHAIN_SOA, C.c_void_p(d_out),
                                                    int(out_capacity), offs, ns, ds)
        self._check(rc)
        return list(offs), list(ns), int(ds[0]) if k else 0

    def device_malloc(self, nbytes: int) -> int:
        p = C.c_void_p()
        self._check(self._L.chain_device_malloc(self._h, nbytes, C.byref(p)))
        return int(p.value)

    def device_free(self, ptr: int):
        self._check(self._L.chain_device_free(self._h, C.c_void_p(ptr)))

    def to_device(self, ptr: int, data) -> None:
        buf = np.ascontiguousarray(data)
        self._check(self._L.chain_copy(self._h, C.c_void_p(ptr), buf.ctypes.data_as(C.c_void_p), buf.nbytes, 1))

    def copy_device(self, dst: int, src: int, nbytes: int) -> None:
        self._check(self._L.chain_copy(self._h, C.c_void_p(dst), C.c_void_p(src), nbytes, 2))

    def to_host(self, ptr: int, nbytes: int) -> bytes:
        out = (C.c_uint8 * max(1, nbytes))()
        self._check(self._L.chain_copy(self._h, out, C.c_void_p(ptr), nbytes, 0))
        return bytes(out)[:nbytes]

    def to_host_into(self, ptr: int, out: np.ndarray) -> np.ndarray:
        """Copy out.nbytes bytes from the device straight into a contiguous host array."""
        if not out.flags.c_contiguous:
            raise ChainValueError("out must be C-contiguous")
        self._check(self._L.chain_copy(self._h, out.ctypes.data_as(C.c_void_p), C.c_void_p(ptr), out.nbytes, 0))
        return out

    def stage_times(self, with_bytes=False):
        """[(stage, ms)] of the last call, or [(stage, ms, algorithmic bytes)]."""
        ms = (C.c_float * 128)()
        names = C.create_string_buffer(8192)
        k = self._L.chain_last_stage_times(self._h, ms, 128, names, 8192)
        labels = names.value.decode().split(";") if k else []
        if not with_bytes:
            return [(labels[i], float(ms[i])) for i in range(k)]
        b = (C.c_uint64 * 128)()
        self._L.chain_last_stage_bytes(self._h, b, 128)
        return [(labels[i], float(ms[i]), int(b[i])) for i in range(k)]


class Served:
    """Result of one served request."""
    __slots__ = ("kind", "status", "error", "data", "n", "dims", "batch", "t_submit", "t_start", "t_end")

    @property
    def latency(self) -> float:
        return self.t_end - self.t_submit


class Server:
    """Request server over one GPU (continuous batching, HBM-budget admission).

    Input arrays / byte strings are kept referenced until their result is
    polled. Compress results are container bytes; decompress results are
    coordinate arrays."""

    def __init__(self, device: int = 0, hbm_budget_bytes: int = 64 << 30, workers: int = 4, max_batch: int = 64):
        L = load_library()
        h = C.c_void_p()
        rc = L.chain_server_create(device, hbm_budget_bytes, workers, max_batch, C.byref(h))
        if rc != 0:
            raise ChainError(L.chain_server_last_error(None).decode())
        self._h, self._L = h, L
        self._keep = {}

    def close(self):
        if getattr(self, "_h", None):
            self._L.chain_server_destroy(self._h)
            self._h = None
            self._keep.clear()

    def __del__(self):
        self.close()

    def clock(self) -> float:
        return float(self._L.chain_server_clock(self._h))

    def arena_allocations(self) -> int:
        """Device allocations made for the workers' arenas since creation."""
        a = C.c_uint64()
        if self._L.chain_server_stats(self._h, C.byref(a)) != 0:
            raise ChainError("invalid server")
        return int(a.value)

    def _submit(self, req, keep):
        t = C.c_uint64()
        rc = self._L.chain_serve_submit(self._h, C.byref(req), C.byref(t))
        if rc != 0:
            msg = self._L.chain_server_last_error(self._h).decode()
            raise (ChainValueError if rc == 1 else ChainError)(msg)
        self._keep[t.value] = keep
        return t.value

    def submit_compress(self, points, eb, p=16, threshold=2, flags=0, fmt="lch1", layout="aos") -> int:
        pts = np.ascontiguousarray(points)
        if pts.ndim != 2:
            raise ChainValueError("points must be a 2-D array")
        n, d = pts.shape if layout == "aos" else pts.shape[::-1]
        req = _Request(kind=REQ_COMPRESS, points=pts.ctypes.data, dtype=_dtype_code(pts.dtype),
                       layout=CHAIN_AOS if layout == "aos" else CHAIN_SOA, dims=d, n=n, eb=float(eb), p=int(p),
                       threshold=int(threshold), flags=int(flags), format=_FORMATS[fmt])
        return self._submit(req, ("c", pts))

    def submit_decompress(self, data: bytes, dtype=np.float64, layout="aos") -> int:
        b = bytes(data)
        req = _Request(kind=REQ_DECOMPRESS, bytes=C.cast(C.c_char_p(b), C.c_void_p).value, len=len(b),
                       out_dtype=_dtype_code(dtype), out_layout=CHAIN_AOS if layout == "aos" else CHAIN_SOA)
        return self._submit(req, ("d", b, np.dtype(dtype), layout))

    def poll(self, ticket: int, wait_ms: int = -1):
        """The request's Served result, or None if it has not finished."""
        r = _Result()
        rc = self._L.chain_serve_poll(self._h, ticket, int(wait_ms), C.byref(r))
        if rc == CHAIN_PENDING:
            return None
        if rc != 0:
            raise ChainError("invalid poll")
        keep = self._keep.pop(ticket, None)
        s = Served()
        s.kind = "compress" if keep and keep[0] == "c" else "decompress"
        s.status, s.error = r.status, r.error.decode(errors="replace")
        s.n, s.dims, s.batch = int(r.n), int(r.dims), int(r.batch)
        s.t_submit, s.t_start, s.t_end = r.t_submit, r.t_start, r.t_end
        if r.status != 0:
            s.data = None
            load_library().chain_buf_free(C.byref(r.out))
        elif s.kind == "compress":
            s.data = _take(r.out)
        else:
            dt, layout = keep[2], keep[3]
            s.data = _take_array(r.out, dt, (s.n, s.dims) if layout == "aos" else (s.dims, s.n))
        return s
