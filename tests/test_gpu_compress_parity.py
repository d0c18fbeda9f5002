"""GPU compressor parity: the engine's pre-entropy container (LCP3, the
reference's ec=False output of cpp/chain/lcps_cpp.cpp:2306-2329) must be
byte-identical to the golden vectors of the compiled reference and to the
oracle, called through the C ABI (libchain_engine.so). Every test runs under
both L1 strategies (successors inside 4096-point tiles plus an item sort, and
one radix sort of every point)."""
import hashlib

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def sha(b):
    return hashlib.sha256(bytes(b)).hexdigest()


@pytest.fixture(scope="module", params=[1, 0], ids=["l1_tiles", "l1_global"])
def eng(request):
    from chaincompressor_amd import engine
    e = engine.Engine(0)
    e.set_tuning(64, 65536, request.param)
    return e


def flags_of(bl, l1, l2):
    from chaincompressor_amd import engine
    return (0 if bl else engine.NO_BL) | (0 if l1 else engine.NO_L1) | (0 if l2 else engine.NO_L2)


def test_golden_small_lcp3(eng, golden_small):
    g = golden_small
    bad = []
    for name in [str(x) for x in g["names"]]:
        eb, p, t, bl, l1, l2 = g[f"{name}__params"]
        pts = g[f"{name}__points"]
        print("case", name, pts.shape, flush=True)
        try:
            b, st = eng.compress(pts, float(eb), int(p), int(t), flags=flags_of(bl, l1, l2), fmt="lcp3")
        except Exception as exc:  # report every case, not only the first
            bad.append((name, str(exc)))
            continue
        if sha(b) != str(g[f"{name}__lcp3_sha"]):
            bad.append((name, "bytes"))
        elif st.as_list() != [int(x) for x in g[f"{name}__stats3"]]:
            bad.append((name, "stats"))
    assert not bad, bad


def test_golden_large_lcp3(eng, golden_large):
    g = golden_large
    db = g["dambreak_qb11__points"]
    for name in [str(x) for x in g["names"]]:
        if name.startswith("dambreak"):
            pts = db
        else:
            from chaincompressor_amd import workloads
            pts = workloads.dambreak_like()
        eb, p, t = g[f"{name}__params"]
        print("case", name, flush=True)
        b, st = eng.compress(pts, float(eb), int(p), int(t), fmt="lcp3")
        assert len(b) == int(g[f"{name}__lcp3_len"]) and sha(b) == str(g[f"{name}__lcp3_sha"]), name


@pytest.mark.parametrize("seed,n,d,dtype,eb,p,t", [
    (1, 300_000, 3, np.float32, 1e-3, 16, 2),
    (2, 200_000, 3, np.float64, 1e-4, 8, 0),
    (3, 100_000, 2, np.float32, 1e-3, 1, 5),
    (4, 50_000, 4, np.float32, 2e-3, 4, -2),
    (5, 70_000, 1, np.float64, 1e-5, 64, 3),
])
def test_random_vs_oracle(eng, seed, n, d, dtype, eb, p, t):
    from chaincompressor_amd import workloads
    rng = np.random.default_rng(seed)
    if d == 3:
        pts = workloads.gaussian_clusters(n, seed=seed, clusters=16).astype(dtype)
    else:
        pts = rng.random((n, d)).astype(dtype)
    print("case", seed, n, d, flush=True)
    b, st = eng.compress(pts, eb, p, t, fmt="lcp3")
    ref = O.compress_streams(pts, eb, p, t).lcp3()
    assert b == ref


def test_soa_layout_matches_aos(eng):
    rng = np.random.default_rng(9)
    pts = (rng.random((40_000, 3)) * 0.2).astype(np.float32)
    a, _ = eng.compress(pts, 1e-3, 16, 2, fmt="lcp3")
    b, _ = eng.compress(np.ascontiguousarray(pts.T), 1e-3, 16, 2, fmt="lcp3", layout="soa")
    assert a == b


def test_batch_matches_single(eng):
    from chaincompressor_amd import workloads
    frames = [workloads.gaussian_clusters(n, seed=40 + k, clusters=4) for k, n in enumerate([1, 5000, 70_000, 123, 4096, 4097])]
    outs, _ = eng.compress_batch(frames, 1e-3, 16, 2, fmt="lcp3")
    for f, o in zip(frames, outs):
        single, _ = eng.compress(f, 1e-3, 16, 2, fmt="lcp3")
        assert o == single
        assert o == O.compress_streams(f, 1e-3, 16, 2).lcp3()


def test_gpu_lcp3_exports_to_reference_lcp2(eng, golden_small):
    """GPU LCP3 -> host LCP2 export == lcps_compress2(..., ec=True) bytes
    (with the reference's libzstd 1.4.9; otherwise the decoded coordinates)."""
    lcp2 = pytest.importorskip("chaincompressor_amd.lcp2")
    try:
        same = lcp2.zstd_version() == "1.4.9"
    except lcp2.Lcp2Error as e:
        pytest.skip(str(e))
    g = golden_small
    for name in [str(x) for x in g["names"]]:
        eb, p, t, bl, l1, l2 = g[f"{name}__params"]
        flags = (0 if bl else 1) | (0 if l1 else 2) | (0 if l2 else 4)
        print("case", name, flush=True)
        b3, _ = eng.compress(g[f"{name}__points"], float(eb), int(p), int(t), flags=flags, fmt="lcp3")
        out, st = lcp2.to_lcp2(b3)
        if same:
            assert out == g[f"{name}__lcp2"].tobytes(), name
            assert st == [int(x) for x in g[f"{name}__stats"]], name
        assert sha(O.decode(out).tobytes()) == str(g[f"{name}__decoded_sha"]), name
