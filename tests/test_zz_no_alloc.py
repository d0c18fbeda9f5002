"""The no-allocation decompression entry point, through the C ABI.

Its output must equal the allocating device entry point's
(chain_decompress_batch_device) bit for bit, on two calls in a row, the second
with a smaller input (the shared scratch keeps the first call's contents).
Output buffers that are too small and batches that need more than the fixed
scratch are refused with an error, and the engine keeps working."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _upload(eng, data):
    arr = np.frombuffer(data, dtype=np.uint8) if isinstance(data, bytes) else np.ascontiguousarray(data)
    ptr = eng.device_malloc(max(1, arr.nbytes))
    eng.to_device(ptr, arr)
    return ptr


def test_noalloc_matches_allocating_entry_point():
    from chaincompressor_amd import engine, workloads
    eng = engine.Engine(0)
    big = [workloads.gaussian_clusters(1_200_000, seed=61), workloads.gaussian_clusters(70_001, seed=62)]
    small = [workloads.gaussian_clusters(150_000, seed=63, clusters=5)]
    out_cap = 1_300_000 * 3 * 8
    d_out = eng.device_malloc(out_cap)
    try:
        for frames, eb, dt in ((big, 1e-3, np.float64), (small, 1e-4, np.float32)):
            streams, _ = eng.compress_batch(frames, eb, 16, 2)
            d_streams = _upload(eng, b"".join(streams))
            try:
                print("noalloc case", [f.shape[0] for f in frames], eb, np.dtype(dt).name, flush=True)
                ref_ptr, ref_offs, ref_ns, ref_dims = eng.decompress_batch_device(streams, d_streams, dtype=dt)
                total = sum(ref_ns) * ref_dims * np.dtype(dt).itemsize
                ref = eng.to_host(ref_ptr, total)
                offs, ns, dims = eng.decompress_batch_noalloc(streams, d_streams, d_out, out_cap, dtype=dt)
                assert offs == ref_offs and ns == ref_ns and dims == ref_dims
                got = eng.to_host(d_out, total)
                assert got == ref
                dec = np.frombuffer(got, dtype=dt).reshape(-1, 3)
                for k, f in enumerate(frames):
                    o = offs[k] // 3
                    assert np.all(np.abs(dec[o:o + ns[k]].astype(np.float64) - f) <= eb + 1e-7)
                # an output buffer one element short is refused
                with pytest.raises(engine.ChainValueError, match="output buffer"):
                    eng.decompress_batch_noalloc(streams, d_streams, d_out, total - 1, dtype=dt)
            finally:
                eng.device_free(d_streams)
        # a batch needing more than the fixed scratch is refused with an error
        huge = workloads.gaussian_clusters(9_000_000, seed=64)
        blob, _ = eng.compress(huge, 1e-4, 4, 0)
        d_huge = _upload(eng, blob)
        try:
            with pytest.raises(engine.ChainError, match="scratch"):
                eng.decompress_batch_noalloc([blob], d_huge, d_out, out_cap)
        finally:
            eng.device_free(d_huge)
        # the engine still works afterwards
        streams, _ = eng.compress_batch(small, 1e-3, 16, 2)
        d_streams = _upload(eng, streams[0])
        try:
            offs, ns, _ = eng.decompress_batch_noalloc(streams, d_streams, d_out, out_cap)
            dec = np.frombuffer(eng.to_host(d_out, ns[0] * 24), dtype=np.float64).reshape(-1, 3)
            assert np.all(np.abs(dec - small[0]) <= 1e-3)
        finally:
            eng.device_free(d_streams)
    finally:
        eng.device_free(d_out)
