#!/usr/bin/env python3
"""Chain compressor benchmark on one MI355X.

Default workload: one frame of 50,000,000 float32 xyz particles (64 Gaussian
clusters, sigma log-uniform in [0.005, 0.05], plus 10 % uniform background,
cluster-major order), eb = 1e-3, p = 16, threshold = 2. A step is one
compression of the frame (points resident in HBM -> LCH1 container in HBM)
followed by one decompression of that container (-> float64 coordinates in
HBM). Throughput = uncompressed input bytes N*3*4 per wall-clock second of
the timed steps; the per-kernel device times come from HIP events the engine
records on its own stream.

    python bench.py [--steps K] [--warmup W] [--workload plain|batched|serving]

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import contextlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


@contextlib.contextmanager
def quiet_stdout():
    """The reference CPU modules printf stage times; keep stdout for the JSON line."""
    sys.stdout.flush()
    fd = os.dup(1)
    with open(os.devnull, "w") as devnull:
        os.dup2(devnull.fileno(), 1)
        try:
            yield
        finally:
            sys.stdout.flush()
            os.dup2(fd, 1)
            os.close(fd)


def agg_stages(stages):
    out = {}
    for name, ms in stages:
        t, c = out.get(name, (0.0, 0))
        out[name] = (t + ms, c + 1)
    return out


def cpu_baseline(pts_sample, eb, p, t):
    """The reference CPU compressor (oracle/_ref, compiled from the reference's
    cpp/chain/lcps_cpp.cpp) on a bounded sample, one thread; falls back to the
    oracle's C restatement when the compiled reference is absent."""
    nbytes = pts_sample.nbytes
    ref_dir = os.path.join(ROOT, "oracle", "_ref")
    try:
        sys.path.insert(0, ref_dir)
        import lcps_cpp  # noqa: F401
        kind = "reference"
    except ImportError:
        kind = "port"
    times = []
    for _ in range(2):
        t0 = time.perf_counter()
        if kind == "reference":
            with quiet_stdout():
                lcps_cpp.lcps_compress2(pts_sample, eb, p, t, 0)
        else:
            sys.path.insert(0, os.path.join(ROOT, "oracle"))
            import oracle as O
            O.compress_lcp2(pts_sample, eb, p, t)
        times.append(time.perf_counter() - t0)
    best = min(times)
    return {"value": round(nbytes / best / 1e9, 6), "unit": "GB/s", "cores": 1, "kind": kind,
            "sample": f"{pts_sample.shape[0]} points (every {{stride}}th point of the frame, order kept), "
                      f"lcps_compress2 ec=True (Zstd-22 chooser), best of 2 runs, {best:.2f} s"}


def run_plain(args, eng):
    from chaincompressor_amd import workloads
    n = args.n
    t0 = time.perf_counter()
    pts = workloads.gaussian_clusters(n, seed=workloads.BASE_SEED + 2)
    log(f"generated {n} points in {time.perf_counter() - t0:.1f} s")
    nbytes = pts.nbytes
    dptr = eng.device_malloc(nbytes)
    eng.to_device(dptr, pts)
    eb, p, thr = args.eb, args.p, args.threshold

    def compress():
        return eng.compress_batch_device([dptr], [n], np.float32, 3, eb, p, thr)

    # one compression to size the decode input; keep a device copy of the container
    d_out, offs, sizes, st = compress()
    blob = eng.to_host(d_out + offs[0], sizes[0])
    dblob = eng.device_malloc(len(blob))
    eng.to_device(dblob, np.frombuffer(blob, dtype=np.uint8))

    def decompress():
        return eng.decompress_batch_device([blob], dblob)

    for _ in range(args.warmup):
        compress()
        decompress()
    tc = td = 0.0
    c_stages, d_stages = [], []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        compress()
        t1 = time.perf_counter()
        c_stages.append(eng.stage_times())
        decompress()
        t2 = time.perf_counter()
        d_stages.append(eng.stage_times())
        tc += t1 - t0
        td += t2 - t1
    # correctness of the timed configuration (bit-exact round trip is covered by the tests)
    dec_ptr, doffs, dns, ddims = decompress()
    dec = np.frombuffer(eng.to_host(dec_ptr, n * 3 * 8), dtype=np.float64).reshape(n, 3)
    maxerr = float(np.max(np.abs(dec - pts)))
    assert maxerr <= eb, maxerr
    eng.device_free(dptr)
    eng.device_free(dblob)
    return pts, nbytes, tc, td, c_stages, d_stages, st[0], sizes[0], maxerr


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="plain", choices=["plain"])
    ap.add_argument("--n", type=int, default=50_000_000)
    ap.add_argument("--eb", type=float, default=1e-3)
    ap.add_argument("--p", type=int, default=16)
    ap.add_argument("--threshold", type=int, default=2)
    ap.add_argument("--cpu-sample", type=int, default=2_000_000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU; --gpus above 1 is not supported"}))
        sys.exit(2)
    from chaincompressor_amd import engine
    eng = engine.Engine(0)
    pts, nbytes, tc, td, c_stages, d_stages, st, csize, maxerr = run_plain(args, eng)
    K = args.steps
    comp_gbs = nbytes * K / tc / 1e9
    dec_gbs = nbytes * K / td / 1e9
    cs = agg_stages([x for s in c_stages for x in s])
    ds = agg_stages([x for s in d_stages for x in s])
    comp_dev_ms = sum(v[0] for v in cs.values()) / K
    dec_dev_ms = sum(v[0] for k, v in ds.items() if k != "h2d") / K
    # dominant kernel stage of compression and its algorithmic bytes per launch
    n = args.n
    dom, (dom_ms, dom_cnt) = max(cs.items(), key=lambda kv: kv[1][0])
    per_launch_ms = dom_ms / dom_cnt
    key_bytes = 4  # block keys of this workload fit 32 bits (see "key_bits" below)
    alg = {
        "sort_scatter": n * 2 * (key_bytes + 4),          # read key+index, write key+index
        "sort_hist": n * key_bytes,
        "quantize": n * 3 * 4 + n * key_bytes + n * 3 * 1,  # points in, keys + u8 offsets out
        "link": n * (2 * key_bytes + 2 * 4) + n * 4,
        "explicit": n * (4 + key_bytes),
    }
    achieved = alg.get(dom, 0) / (per_launch_ms / 1e3) / 1e9 if dom in alg else None
    line = {
        "metric": "compress GB/s (one 50M-particle float32 frame, eb=1e-3; decompress GB/s alongside)",
        "value": round(comp_gbs, 3),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": round((tc + td) / K * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Gaussian-cluster frame)",
        "config": {"workload": f"plain single frame: {n} float32 xyz Gaussian clusters + 10% background",
                   "eb": args.eb, "p": args.p, "threshold": args.threshold, "container": "LCH1"},
        "compress_gbps": round(comp_gbs, 3),
        "decompress_gbps": round(dec_gbs, 3),
        "compress_device_ms": round(comp_dev_ms, 3),
        "decompress_device_ms": round(dec_dev_ms, 3),
        "compressed_bytes": int(csize),
        "bits_per_particle": round(csize * 8 / n, 3),
        "max_abs_error": maxerr,
        "stages_compress_ms": {k: round(v[0] / K, 3) for k, v in cs.items()},
        "stages_decompress_ms": {k: round(v[0] / K, 3) for k, v in ds.items()},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1) if achieved else None,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None, "traffic": None},
    }
    if not args.no_cpu_baseline:
        stride = max(1, n // args.cpu_sample)
        sample = np.ascontiguousarray(pts[::stride][:args.cpu_sample])
        cb = cpu_baseline(sample, args.eb, args.p, args.threshold)
        cb["sample"] = cb["sample"].replace("{stride}", str(stride))
        line["cpu_baseline"] = cb
    print(json.dumps(line))


if __name__ == "__main__":
    main()
