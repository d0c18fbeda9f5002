"""The benchmark workloads at their full sizes.

* One frame of 50M float32 Gaussian-cluster points at eb = 1e-2, 1e-3 and
  1e-4: the GPU's LCH1 container equals the oracle's byte for byte, the GPU
  decoder's float64 output equals the oracle decoder's bit for bit on the same
  bytes, and every coordinate is within eb.
* The 256-frame uneven float32 batch (about 1.6e8 points): the batched call's
  containers equal one call per frame, two frames equal the oracle's, and every
  decoded coordinate of the batch is within eb.
* The 128-frame float64 MD batch (fcc + thermal noise, eb = 1e-4): batch
  decompression is bit-identical to per-frame decompression and, on sampled
  frames, to the oracle decoder; every coordinate is within eb.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(600)]


@pytest.fixture(scope="module")
def eng():
    from chaincompressor_amd import engine
    return engine.Engine(0)


@pytest.fixture(scope="module")
def frame50m():
    from chaincompressor_amd import workloads
    return workloads.gaussian_clusters(50_000_000, seed=workloads.BASE_SEED + 2)


def _upload(eng, arr):
    ptr = eng.device_malloc(arr.nbytes)
    eng.to_device(ptr, arr)
    return ptr


@pytest.mark.parametrize("eb", [1e-2, 1e-3, 1e-4])
def test_plain_frame_50m(eng, frame50m, eb):
    pts = frame50m
    n = pts.shape[0]
    print("50M frame, eb", eb, flush=True)
    dptr = _upload(eng, pts)
    try:
        d, offs, sizes, _ = eng.compress_batch_device([dptr], [n], np.float32, 3, eb, 16, 2)
        blob = eng.to_host(d + offs[0], sizes[0])
    finally:
        eng.device_free(dptr)
    ref, kinds = O.compress_streams(pts, eb, 16, 2).lch1()
    assert blob == ref, kinds
    dec = eng.decompress(blob)
    assert np.array_equal(dec.view(np.uint64), O.decode(blob).view(np.uint64))
    assert float(np.max(np.abs(dec - pts))) <= eb


def test_uneven_batch_256(eng):
    from chaincompressor_amd import workloads
    frames = workloads.uneven_batch(frames=256)
    ns = [f.shape[0] for f in frames]
    big = np.concatenate(frames)
    print("uneven batch:", len(frames), "frames,", big.shape[0], "points", flush=True)
    dptr = _upload(eng, big)
    offs = np.concatenate([[0], np.cumsum(ns)[:-1]])
    ptrs = [dptr + int(o) * 12 for o in offs]
    try:
        d, co, cs, _ = eng.compress_batch_device(ptrs, ns, np.float32, 3, 1e-3, 16, 2)
        streams = [eng.to_host(d + o, s) for o, s in zip(co, cs)]
        for k in range(len(ns)):
            d1, o1, s1, _ = eng.compress_batch_device([ptrs[k]], [ns[k]], np.float32, 3, 1e-3, 16, 2)
            assert eng.to_host(d1 + o1[0], s1[0]) == streams[k], k
    finally:
        eng.device_free(dptr)
    for k in (int(np.argmin(ns)), int(np.argmax(ns))):
        assert streams[k] == O.compress_streams(frames[k], 1e-3, 16, 2).lch1()[0], k
    blob = b"".join(streams)
    dblob = _upload(eng, np.frombuffer(blob, dtype=np.uint8))
    try:
        out, ooffs, dns, dims = eng.decompress_batch_device(streams, dblob)
        assert dims == 3 and list(dns) == ns
        dec = np.frombuffer(eng.to_host(out, big.shape[0] * 3 * 8), dtype=np.float64).reshape(-1, 3)
    finally:
        eng.device_free(dblob)
    err = 0.0
    for k, f in enumerate(frames):
        o = int(ooffs[k]) // 3
        err = max(err, float(np.max(np.abs(dec[o:o + ns[k]] - f))))
    assert err <= 1e-3


def test_md_batch_128(eng):
    from chaincompressor_amd import workloads
    frames = workloads.md_fcc_batch()
    print("md batch:", len(frames), "frames,", sum(f.shape[0] for f in frames), "atoms", flush=True)
    outs, _ = eng.compress_batch(frames, 1e-4, 16, 2, fmt="lch1")
    decs = eng.decompress_batch(outs)
    for k, (f, d) in enumerate(zip(frames, decs)):
        assert np.all(np.abs(d - f) <= 1e-4), k
    rng = np.random.default_rng(3)
    for k in sorted(set(rng.integers(0, len(frames), size=4).tolist())):
        single = eng.decompress(outs[k])
        assert np.array_equal(single.view(np.uint64), decs[k].view(np.uint64)), k
        assert np.array_equal(O.decode(outs[k]).view(np.uint64), decs[k].view(np.uint64)), k
