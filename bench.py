#!/usr/bin/env python3
"""Chain compressor benchmark on one MI355X.

Workloads (synthetic, seeded; chaincompressor_amd/workloads.py):

* plain    -- one frame of 50,000,000 float32 xyz particles (64 Gaussian
  clusters, sigma log-uniform in [0.005, 0.05], plus 10 % uniform background,
  cluster-major order), eb = 1e-3, p = 16, threshold = 2. A step is one
  compression of the frame (points resident in HBM -> LCH1 container in HBM)
  followed by one decompression of that container (-> float64 coordinates in
  HBM). ``value`` is the compression throughput: uncompressed input bytes
  N*3*4 per wall-clock second of the timed compressions. The same frame is
  also run at eb = 1e-2 and 1e-4. Host<->device copies are timed on their own.
* batched  -- 256 float32 frames with log-uniform counts in [1e4, 4e6], each
  half jittered lattice and half Gaussian mix, eb = 1e-3: all frames through
  one call (one launch per stage) against one call per frame.
* md       -- 128 float64 MD-like frames (fcc + thermal noise, periodic box),
  counts uniform in [1e5, 2e6], eb = 1e-4: batch decompression against one
  call per frame; with the CPU baseline, a sample of frames is decoded by the
  reference CPU decoder and compared bit for bit with the GPU output.
* serving  -- Poisson arrivals (200 req/s for 60 s), 70 % compress / 30 %
  decompress, sizes log-uniform in [5e3, 2e7], float32 80 % / float64 20 %, eb
  in {1e-2, 1e-3, 1e-4}, admitted under a 64 GB device budget; p50/p99
  latency. Then a saturation ladder: shorter traces at doubled arrival rates
  until the server no longer keeps up (or a request fails); the capacity is
  the highest aggregate rate reached, reported with its p99.

Roofline figures. Every timed region states its algorithmic bytes (the
engine's own count of what each kernel must read and write once); the
``roofline`` object takes the region with the most device time per step (HIP
events on the engine's stream) and divides its bytes by its time.
``pipeline_alg_frac`` does the same over all regions. ``model_bytes`` gives the
stage model of the compressor and decompressor in the same units:

  B_c = 2NDs + N(k+4D) + 2PN(k+4) + 2N(k+8) + 8N(D+1) + 2N + (24D+11)B + 3Y + Z
  B_d = Z + 2Y + N(12D+16) + B(12D+8) + N*D*s_out

with s the input element bytes, k the sort key bytes (4 up to 32 key bits,
else 8), P = ceil(key bits / 8) radix passes, B explicit block records, Y the
pre-entropy bytes (the LCP3 container of the same frame) and Z the
compressed bytes; its fraction divides it by the summed region time.

    python bench.py [--steps K] [--warmup W] [--workload default|plain|batched|md|serving|all]

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import contextlib
import ctypes
import json
import os
import re
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

# timed region -> the HIP kernels it launches
REGION_KERNELS = {
    "bbox": "chain::k_bbox + k_bbox_finish",
    "quantize": "chain::k_quantize",
    "sort_hist": "chain::k_sort_hist",
    "sort_scan": "chain::k_scan_u32",
    "sort_scatter": "chain::k_sort_scatter",
    "tile_link": "chain::k_tile_link",
    "link": "chain::k_link_items (single-word keys) or k_link / k_link_wide",
    "explicit": "chain::k_explicit + k_frame_counts",
    "l2_stats": "chain::k_l2_stats + k_l2_wrap",
    "l2_split": "chain::k_l2_split + k_frame_split_counts",
    "corrections": "chain::k_scan_u32 + k_corr_gather + k_frame_corr",
    "comp_stats": "chain::k_comp_stats + k_comp_reduce (one pair per component)",
    "rans_table": "chain::k_rans_table + k_prechoose + k_rans_plan",
    "rans_encode": "chain::k_rans_encode (one launch, every contending component's chunks)",
    "assemble": "chain::k_section_sizes + k_choose + k_layout + k_write_frames + k_write_components + k_copy_chunks",
    "pack": "chain::k_pack_fixed + k_pack_tagged + k_tag_seams (per component)",
    "rans_decode": "chain::k_dec_rans",
    "unpack_fixed": "chain::k_dec_fixed",
    "unpack_raw": "chain::k_dec_raw",
    "unpack_tagged": "chain::k_dec_tagged",
    "l2_counts": "chain::k_dec_tile_zero_counts + k_scan_u32",
    "l2_inverse": "chain::k_dec_l2",
    "chain_local": "chain::k_dec_chain_local + k_dec_check_counts",
    "chain_jump": "chain::k_dec_chain_jump (4 launches) + k_dec_chain_resolve",
    "dequantize": "chain::k_dec_dequant + k_dec_corr",
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


@contextlib.contextmanager
def captured_stdout():
    """Collect what native code prints on file descriptor 1 (the reference CPU
    modules printf their stage times); stdout stays free for the JSON line."""
    sys.stdout.flush()
    libc = ctypes.CDLL(None)
    fd = os.dup(1)
    box = {"text": ""}
    with tempfile.TemporaryFile(mode="w+b") as tmp:
        os.dup2(tmp.fileno(), 1)
        try:
            yield box
        finally:
            libc.fflush(None)
            os.dup2(fd, 1)
            os.close(fd)
            tmp.seek(0)
            box["text"] = tmp.read().decode(errors="replace")


def agg_regions(steps):
    """steps: list of [(name, ms, bytes)] -> {name: [ms, launches, bytes]} summed."""
    out = {}
    for s in steps:
        for name, ms, b in s:
            a = out.setdefault(name, [0.0, 0, 0])
            a[0] += ms
            a[1] += 1
            a[2] += b
    return out


def roofline(step_lists, K, profiled=False):
    """The region with the most device time per step over all step lists.
    `traffic` comes from the committed counter profile, which covers the
    plain frame only (profiled=True); other workloads report it as null."""
    regions = {}
    for steps in step_lists:
        for k, v in agg_regions(steps).items():
            a = regions.setdefault(k, [0.0, 0, 0])
            a[0] += v[0]
            a[1] += v[1]
            a[2] += v[2]
    regions.pop("h2d", None)
    if not regions:
        return None
    name, (ms, cnt, b) = max(regions.items(), key=lambda kv: kv[1][0])
    achieved = b / (ms / 1e3) / 1e9 if b and ms > 0 else None
    return {
        "bound": "hbm",
        "kernel": name,
        "hip_kernels": REGION_KERNELS.get(name, name),
        "launches_per_step": cnt / K,
        "ms_per_step": round(ms / K, 4),
        "alg_bytes_per_step": int(b / K),
        "achieved": round(achieved, 2) if achieved else None,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 5) if achieved else None,
        "traffic": traffic_from_profiles(name) if profiled else None,
    }


def region_table(steps, K):
    return {k: {"ms": round(v[0] / K, 4), "GBps": round(v[2] / (v[0] / 1e3) / 1e9, 1) if v[2] and v[0] else None}
            for k, v in agg_regions(steps).items()}


def region_sums(steps, K):
    """(summed device ms per step, stated bytes per step), host copies left out."""
    a = agg_regions(steps)
    a.pop("h2d", None)
    return sum(v[0] for v in a.values()) / K, sum(v[2] for v in a.values()) / K


def pipeline_fraction(steps, K):
    """Stated algorithmic bytes of all regions over their summed device time."""
    ms, b = region_sums(steps, K)
    return round(b / (ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4) if ms else None


def traffic_from_profiles(region):
    """HBM bytes per step of a region measured with rocprofv3 --pmc
    FETCH_SIZE / WRITE_SIZE (scaled by a copy kernel), if recorded."""
    path = os.path.join(ROOT, "profiles", "traffic_50M_f32_eb1e-3.json")
    try:
        with open(path) as fh:
            t = json.load(fh)
        return t.get("regions", {}).get(region)
    except (OSError, ValueError):
        return None


def key_bits(frame, eb, p):
    """Key bits of a frame from its bounding box: sum over dimensions of the
    bit length of the largest block id floor(floor((max - min) / 2eb) / p)."""
    f = np.asarray(frame)
    mn = f.min(axis=0).astype(np.float64)
    mx = f.max(axis=0).astype(np.float64)
    q = np.floor((mx - mn) / (2.0 * eb)).astype(np.int64)
    return int(sum(int(v // p).bit_length() for v in q))


def model_bytes(meta, s_in, s_out=8, D=3):
    """meta: [(N, key_bits, B, Y, Z)] per frame -> (B_c, B_d) summed over frames."""
    bc = bd = 0
    for N, kb, B, Y, Z in meta:
        k = 4 if kb <= 32 else 8
        P = max(1, -(-kb // 8))
        bc += (2 * N * D * s_in + N * (k + 4 * D) + 2 * P * N * (k + 4) + 2 * N * (k + 8) + 8 * N * (D + 1)
               + 2 * N + (24 * D + 11) * B + 3 * Y + Z)
        bd += Z + 2 * Y + N * (12 * D + 16) + B * (12 * D + 8) + N * D * s_out
    return int(bc), int(bd)


def model_block(meta, s_in, c_steps, d_steps, K):
    """B_c / B_d beside the engine's stated bytes and their fractions of peak
    over the summed region times."""
    bc, bd = model_bytes(meta, s_in)
    out = {"B_c": bc, "B_d": bd}
    for tag, steps, model in (("compress", c_steps, bc), ("decompress", d_steps, bd)):
        if not steps:
            continue
        ms, stated = region_sums(steps, K)
        out[f"{tag}_stated_bytes"] = int(stated)
        out[f"{tag}_device_ms"] = round(ms, 4)
        out[f"{tag}_model_frac"] = round(model / (ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4) if ms else None
        out[f"{tag}_stated_frac"] = round(stated / (ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4) if ms else None
    return out


def cpu_info():
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return {"cpu_model": model, "nproc": os.cpu_count(), "usable_cpus": usable, "threads": 1}


_STAGES = (("quantization_block", r"Quantization\+block:\s*([0-9.eE+-]+)"), ("l1", r"^l1:\s*([0-9.eE+-]+)"),
           ("l2", r"^l2:\s*([0-9.eE+-]+)"), ("encoding", r"^Encoding:\s*([0-9.eE+-]+)"))


def _stage_times(text):
    out = {}
    for name, pat in _STAGES:
        vals = [float(m.group(1)) for m in re.finditer(pat, text, re.M)]
        if vals:
            out[name] = vals[-1]
    return out


def _ref_modules():
    """The reference's own CPU modules compiled from its sources into
    oracle/_ref (checker and baseline only), or None."""
    ref_dir = os.path.join(ROOT, "oracle", "_ref")
    if ref_dir not in sys.path:
        sys.path.insert(0, ref_dir)
    try:
        import lcps_cpp
        import lcps_cpu_decompress
        return lcps_cpp, lcps_cpu_decompress
    except ImportError:
        return None


def cpu_baseline(pts_sample, eb, p, t, runs=5, warmup=1):
    """The reference CPU compressor and decoder (oracle/_ref, compiled from
    the reference's cpp/chain/lcps_cpp.cpp and
    cpp/decompress/lcps_cpu_decompress_serial.cpp) on a bounded sample, one
    thread, median of `runs` after `warmup` runs: lcps_compress2 with ec=True
    (Zstd-22 chooser, LCP2), with ec=False (pre-entropy streams only, LCP3)
    and lcps_decompress_serial. The pre-entropy / entropy split is the
    reference's own stage clock (its printed Quantization+block, l1, l2 and
    Encoding times). Falls back to the oracle's C restatement (no split) when
    the compiled reference is absent."""
    nbytes = pts_sample.nbytes
    mods = _ref_modules()
    info = cpu_info()
    if mods is None:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        blob = O.compress_lcp2(pts_sample, eb, p, t)[0]
        tc, td = [], []
        for k in range(runs + warmup):
            t0 = time.perf_counter()
            O.compress_lcp2(pts_sample, eb, p, t)
            t1 = time.perf_counter()
            O.decode(blob)
            t2 = time.perf_counter()
            if k >= warmup:
                tc.append(t1 - t0)
                td.append(t2 - t1)
        mc, md = float(np.median(tc)), float(np.median(td))
        return dict(info, value=round(nbytes / mc / 1e9, 6), unit="GB/s", cores=1, kind="port",
                    decompress_value=round(nbytes / md / 1e9, 6),
                    sample=f"{pts_sample.shape[0]} points; oracle C restatement, median of {runs}")
    lcps_cpp, lcps_dec = mods

    def timed(fn):
        walls, stages = [], []
        for k in range(runs + warmup):
            with captured_stdout() as box:
                t0 = time.perf_counter()
                r = fn()
                w = time.perf_counter() - t0
            if k >= warmup:
                walls.append(w)
                stages.append(_stage_times(box["text"]))
        med = {name: float(np.median([s[name] for s in stages if name in s]))
               for name, _ in _STAGES if any(name in s for s in stages)}
        return r, float(np.median(walls)), med

    blob, w_ec, st_ec = timed(lambda: lcps_cpp.lcps_compress2(pts_sample, eb, p, t, 0)[0])
    _, w_raw, st_raw = timed(lambda: lcps_cpp.lcps_compress2(pts_sample, eb, p, t, 0, ec=False)[0])
    _, w_dec, _ = timed(lambda: lcps_dec.lcps_decompress_serial(blob))
    pre = sum(st_ec.get(k, 0.0) for k in ("quantization_block", "l1", "l2"))
    ent = st_ec.get("encoding")

    def gbps(sec):
        return round(nbytes / sec / 1e9, 6) if sec else None

    return dict(
        info, value=gbps(w_ec), unit="GB/s", cores=1, kind="reference",
        decompress_value=gbps(w_dec),
        ec_false_value=gbps(w_raw),
        pre_entropy_value=gbps(pre), entropy_value=gbps(ent),
        stage_seconds={"ec_true": st_ec, "ec_false": st_raw},
        wall_seconds={"compress_ec_true": round(w_ec, 4), "compress_ec_false": round(w_raw, 4),
                      "decompress": round(w_dec, 4)},
        sample=f"{pts_sample.shape[0]} points (every {{stride}}th point of the frame, order kept), "
               f"lcps_compress2 ec=True (Zstd-22 chooser, LCP2) and ec=False (LCP3), lcps_decompress_serial; "
               f"median of {runs} after {warmup} warm-up; wall = pybind call incl. array conversion, "
               f"stages = the reference's own stage clock")


def cpu_md_check(eng, frames, streams, eb, p, t, count=6):
    """md frames: the reference CPU decoder on the LCP2 container of a sample
    of frames (smallest to largest; the GPU's LCP3 through the host LCP2
    export) against the GPU decode of the batched LCH1 containers, bit for
    bit; the decoder's time is the md CPU baseline."""
    mods = _ref_modules()
    if mods is None:
        return {"checked_frames": 0, "note": "compiled reference CPU decoder absent"}
    _, lcps_dec = mods
    from chaincompressor_amd import lcp2
    order = np.argsort([f.shape[0] for f in frames])
    pick = [int(order[int(round(i * (len(order) - 1) / max(1, count - 1)))]) for i in range(count)]
    exact, npts, sec = True, 0, 0.0
    for k in pick:
        gpu = eng.decompress(streams[k])
        lcp3, _ = eng.compress(frames[k], eb, p, t, fmt="lcp3")
        blob, _ = lcp2.to_lcp2(lcp3)
        t0 = time.perf_counter()
        with captured_stdout():
            ref = lcps_dec.lcps_decompress_serial(blob)
        sec += time.perf_counter() - t0
        ref = np.ascontiguousarray(ref, dtype=np.float64)
        exact = exact and ref.shape == gpu.shape and np.array_equal(ref.view(np.uint64), gpu.view(np.uint64))
        npts += frames[k].shape[0]
    return {"checked_frames": len(pick), "points": int(npts), "bit_exact_vs_reference_cpu_decoder": bool(exact),
            "cpu_decompress_gbps": round(npts * 3 * 8 / sec / 1e9, 6), "cores": 1,
            "sample": "reference lcps_decompress_serial on the GPU-written LCP3 of frames spread from the "
                      "smallest to the largest; compared with the GPU decode of the batch's LCH1 containers"}


def upload(eng, arr):
    ptr = eng.device_malloc(max(1, arr.nbytes))
    eng.to_device(ptr, arr)
    return ptr


def lcp3_sizes(eng, ptrs, ns, dtype, eb, p, thr):
    outs, _ = eng.compress_batch(None, eb, p, thr, fmt="lcp3", device_ptrs=(ptrs, ns, dtype, 3))
    return [len(o) for o in outs]


# ----------------------------------------------------------------------------
# plain: one 50M-particle frame

def run_plain(args, eng):
    from chaincompressor_amd import workloads
    n = args.n
    t0 = time.perf_counter()
    pts = workloads.gaussian_clusters(n, seed=workloads.BASE_SEED + 2)
    log(f"plain: generated {n} points in {time.perf_counter() - t0:.1f} s")
    nbytes = pts.nbytes
    dptr = upload(eng, pts)
    # host <-> device copies, pageable host memory, timed on their own
    t0 = time.perf_counter()
    eng.to_device(dptr, pts)
    h2d_s = time.perf_counter() - t0
    p, thr = args.p, args.threshold

    def one_eb(eb, steps, warmup, timed_regions):
        # the engine's own tuner picks the entropy chunking for this frame;
        # repeated compressions of the same shape replay a captured HIP graph
        if args.tuning:
            lanes, chunk, l1 = (int(x) for x in args.tuning.split(","))
            eng.set_autotune(False)
            eng.set_tuning(lanes, chunk, l1)
            tuning = eng.tuning()
        elif args.rans_chunk:
            eng.set_autotune(False)
            eng.set_tuning(64, args.rans_chunk, 1)
            tuning = eng.tuning()
        else:
            tuning, _ = eng.autotune([dptr], [n], np.float32, 3, eb, p, thr)

        def compress(graph=True):
            eng.set_graphs(graph)
            try:
                return eng.compress_batch_device([dptr], [n], np.float32, 3, eb, p, thr)
            finally:
                eng.set_graphs(False)

        d_out, offs, sizes, st = compress()
        blob = eng.to_host(d_out + offs[0], sizes[0])
        dblob = upload(eng, np.frombuffer(blob, dtype=np.uint8))

        def decompress():
            return eng.decompress_batch_device([blob], dblob)

        for _ in range(warmup):
            compress()
            decompress()
        tc = td = 0.0
        c_steps, d_steps = [], []
        for _ in range(steps):
            t0 = time.perf_counter()
            compress()
            t1 = time.perf_counter()
            decompress()
            t2 = time.perf_counter()
            if timed_regions:
                d_steps.append(eng.stage_times(with_bytes=True))
            tc += t1 - t0
            td += t2 - t1
        if timed_regions:
            # per-region device times of the same compression, launched
            # without the graph (a replayed graph is one timed region)
            for _ in range(steps):
                compress(graph=False)
                c_steps.append(eng.stage_times(with_bytes=True))
        dec_ptr, _, _, _ = decompress()
        dec = np.empty((n, 3), dtype=np.float64)
        t0 = time.perf_counter()
        eng.to_host_into(dec_ptr, dec)
        d2h_s = time.perf_counter() - t0
        maxerr = float(np.max(np.abs(dec - pts)))
        assert maxerr <= eb, maxerr
        extra = {}
        if timed_regions:
            y = lcp3_sizes(eng, [dptr], [n], np.float32, eb, p, thr)[0]
            meta = [(n, key_bits(pts, eb, p), n - int(st[0].num_chained), y, int(sizes[0]))]
            extra = {"model": model_block(meta, 4, c_steps, d_steps, steps), "key_bits": meta[0][1],
                     "explicit_records": meta[0][2], "pre_entropy_bytes": y,
                     "d2h_decoded_gbps": round(dec.nbytes / d2h_s / 1e9, 2)}
        eng.device_free(dblob)
        eng.set_tuning()
        return dict({"compress_gbps": nbytes * steps / tc / 1e9, "decompress_gbps": nbytes * steps / td / 1e9,
                     "ms_per_step": (tc + td) / steps * 1e3, "bytes": int(sizes[0]), "st": st[0],
                     "maxerr": maxerr, "c_steps": c_steps, "d_steps": d_steps, "rans": tuning}, **extra)

    if args.calibrate_copy:
        # a plain device-to-device copy of known size, for scaling HBM
        # counters (FETCH_SIZE / WRITE_SIZE) read by a profiler
        cb = 1 << 30
        a, b = eng.device_malloc(cb), eng.device_malloc(cb)
        eng.copy_device(b, a, cb)
        eng.device_free(a)
        eng.device_free(b)
    main = one_eb(args.eb, args.steps, args.warmup, True)
    main["h2d_input_gbps"] = round(nbytes / h2d_s / 1e9, 2)
    # parameter search on the device: p x threshold pairs of the same frame
    search = None
    if not args.no_search:
        t0 = time.perf_counter()
        rows, best = eng.search(None, args.eb, [4, 8, 16, 32, 64], [0, 2, 4], device=(dptr, n, 3, np.float32))
        search = {"pairs": len(rows), "seconds": round(time.perf_counter() - t0, 3),
                  "best": {"p": best[0], "threshold": best[1], "bits_per_particle": round(best[2] * 8 / n, 3)},
                  "p_values": [4, 8, 16, 32, 64], "thresholds": [0, 2, 4]}
    sweep = {}
    for eb in (() if args.no_sweep else (1e-2, 1e-4)):
        r = one_eb(eb, 2, 1, False)
        sweep[f"{eb:g}"] = {"compress_gbps": round(r["compress_gbps"], 2),
                            "decompress_gbps": round(r["decompress_gbps"], 2),
                            "bits_per_particle": round(r["bytes"] * 8 / n, 3), "max_abs_error": r["maxerr"],
                            "tuning": r["rans"]}
    eng.device_free(dptr)
    main["search"] = search
    return pts, nbytes, main, sweep


# ----------------------------------------------------------------------------
# batched: uneven frames, one call vs one call per frame

def run_batched(args, eng, frames=None, eb=1e-3, dtype=np.float32, label="batched"):
    from chaincompressor_amd import workloads
    t0 = time.perf_counter()
    if frames is None:
        frames = workloads.uneven_batch(frames=args.batch_frames)
    ns = [int(f.shape[0]) for f in frames]
    total = sum(ns)
    es = np.dtype(dtype).itemsize
    nbytes = total * 3 * es
    big = np.concatenate(frames).astype(dtype, copy=False)
    log(f"{label}: {len(frames)} frames, {total} points ({nbytes / 1e9:.2f} GB) ready in "
        f"{time.perf_counter() - t0:.1f} s")
    dptr = upload(eng, big)
    offs = np.concatenate([[0], np.cumsum(ns)[:-1]]).astype(np.int64)
    ptrs = [dptr + int(o) * 3 * es for o in offs]
    p, thr = args.p, args.threshold

    # the batched path: entropy coding and L1 strategy tuned on this batch,
    # launches replayed from a captured HIP graph
    if args.tuning:
        lanes, chunk, l1 = (int(x) for x in args.tuning.split(","))
        eng.set_autotune(False)
        eng.set_tuning(lanes, chunk, l1)
        tuned = dict(eng.tuning(), trials="fixed by --tuning")
    else:
        best, trials = eng.autotune(ptrs, ns, dtype, 3, eb, p, thr)
        tuned = dict(best, trials=[dict(t, ms=round(ms, 3), bytes=b) for t, ms, b in trials])

    def comp_batched(graph=True):
        eng.set_graphs(graph)
        try:
            return eng.compress_batch_device(ptrs, ns, dtype, 3, eb, p, thr)
        finally:
            eng.set_graphs(False)

    # containers of the batched call, kept in a device buffer for decoding
    d_out, coffs, csizes, cst = comp_batched()
    streams = [eng.to_host(d_out + o, s) for o, s in zip(coffs, csizes)]
    blob = b"".join(streams)
    dblob = upload(eng, np.frombuffer(blob, dtype=np.uint8))
    soffs = np.concatenate([[0], np.cumsum(csizes)[:-1]]).astype(np.int64).tolist()

    # the per-frame path (same tuned setting) writes the same containers
    identical = None if args.no_per_frame else True
    for k in range(0 if args.no_per_frame else len(ns)):
        o, oo, ss, _ = eng.compress_batch_device([ptrs[k]], [ns[k]], dtype, 3, eb, p, thr)
        if ss[0] != csizes[k] or eng.to_host(o + oo[0], ss[0]) != streams[k]:
            identical = False
            break

    def comp_frames():
        # the plain path: one call per frame with the batch's tuned setting
        for k in range(len(ns)):
            eng.compress_batch_device([ptrs[k]], [ns[k]], dtype, 3, eb, p, thr)

    def dec_batched():
        return eng.decompress_batch_device(streams, dblob)

    def dec_frames():
        for k in range(len(ns)):
            eng.decompress_batch_device([streams[k]], dblob + soffs[k])

    def timed(fn, steps, warmup):
        for _ in range(warmup):
            fn()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        return (time.perf_counter() - t0) / steps

    steps = args.batch_steps
    t_cb = timed(comp_batched, steps, 1)
    t_cf = None if args.no_per_frame else timed(comp_frames, steps, 1)
    t_db = timed(dec_batched, steps, 1)
    t_df = None if args.no_per_frame else timed(dec_frames, steps, 1)
    # per-region device times of the batched calls (compression without the
    # graph, so that every region has its own events)
    c_steps, d_steps = [], []
    for _ in range(steps):
        comp_batched(graph=False)
        c_steps.append(eng.stage_times(with_bytes=True))
        dec_batched()
        d_steps.append(eng.stage_times(with_bytes=True))
    # correctness of the batched decode: every coordinate within eb
    dptr_out, doffs, dns, _ = dec_batched()
    dec = np.empty((total, 3), dtype=np.float64)
    eng.to_host_into(dptr_out, dec)
    maxerr = float(np.max(np.abs(dec - big.astype(np.float64))))
    assert maxerr <= eb, maxerr
    del dec
    ys = lcp3_sizes(eng, ptrs, ns, dtype, eb, p, thr)
    meta = [(ns[k], key_bits(frames[k], eb, p), ns[k] - int(cst[k].num_chained), ys[k], int(csizes[k]))
            for k in range(len(ns))]
    eng.device_free(dptr)
    eng.device_free(dblob)
    eng.set_tuning()
    r = {
        "frames": len(ns), "points": total, "input_bytes": nbytes, "eb": eb, "dtype": np.dtype(dtype).name,
        "compress_gbps_batched": round(nbytes / t_cb / 1e9, 2),
        "compress_gbps_per_frame": round(nbytes / t_cf / 1e9, 2) if t_cf else None,
        "compress_speedup": round(t_cf / t_cb, 2) if t_cf else None,
        "decompress_gbps_batched": round(nbytes / t_db / 1e9, 2),
        "decompress_gbps_per_frame": round(nbytes / t_df / 1e9, 2) if t_df else None,
        "decompress_speedup": round(t_df / t_db, 2) if t_df else None,
        "bits_per_particle": round(sum(csizes) * 8 / total, 3),
        "containers_identical_to_per_frame": identical,
        "max_abs_error": maxerr, "steps": steps,
        "autotune": tuned, "graph_captures_replays": list(eng.graph_stats()),
        "regions_compress": region_table(c_steps, steps),
        "regions_decompress": region_table(d_steps, steps),
        "pipeline_alg_frac": {"compress": pipeline_fraction(c_steps, steps),
                              "decompress": pipeline_fraction(d_steps, steps)},
        "roofline": roofline([c_steps, d_steps], steps),
        "model": model_block(meta, es, c_steps, d_steps, steps),
        "per_frame_path": "one call per frame with the batch's tuned setting, no graph; "
                          "per-frame decompression reads the batched path's containers",
    }
    return r, {"frames": frames, "streams": streams}


# ----------------------------------------------------------------------------
# serving: Poisson arrivals through the request server

class ServingBench:
    def __init__(self, args, eng):
        from chaincompressor_amd import workloads
        self.args = args
        t0 = time.perf_counter()
        self.pool_max = 20_000_000
        self.pools = {np.float32: workloads.gaussian_clusters(self.pool_max, seed=workloads.BASE_SEED + 5),
                      np.float64: workloads.gaussian_clusters(self.pool_max, seed=workloads.BASE_SEED + 6,
                                                              dtype=np.float64)}
        # decompress requests replay streams compressed beforehand: for every
        # dtype and eb, 12 log-spaced sizes; a request takes the nearest size
        self.lib_sizes = np.unique(np.round(np.exp(np.linspace(np.log(5e3), np.log(2e7), 12))).astype(np.int64))
        self.library = {}
        for dt, pool in self.pools.items():
            for eb in (1e-2, 1e-3, 1e-4):
                for n in self.lib_sizes:
                    blob, _ = eng.compress(pool[:int(n)], eb, args.p, args.threshold)
                    self.library[(dt, eb, int(n))] = blob
        log(f"serving: pools and {len(self.library)} streams ready in {time.perf_counter() - t0:.1f} s")

    def run(self, rate, seconds, seed_offset=0):
        import queue
        import threading
        from chaincompressor_amd import engine, workloads
        args = self.args
        trace = workloads.serving_trace(rate=rate, seconds=seconds, seed=workloads.BASE_SEED + 5 + seed_offset)
        rng = np.random.default_rng(workloads.BASE_SEED + 7 + seed_offset)
        srv = engine.Server(0, hbm_budget_bytes=args.serve_budget_gb << 30, workers=args.serve_workers,
                            max_batch=args.serve_max_batch)
        pending = queue.Queue()
        starts = rng.integers(0, self.pool_max, size=len(trace))
        offered = 0

        def submitter():
            nonlocal offered
            base = time.perf_counter()
            for k, (t_arr, kind, n, dt, eb) in enumerate(trace):
                delay = base + t_arr - time.perf_counter()
                if delay > 0:
                    time.sleep(delay)
                if kind == "compress":
                    a = int(min(starts[k], self.pool_max - n))
                    tk = srv.submit_compress(self.pools[dt][a:a + n], eb, args.p, args.threshold)
                    b = n * 3 * np.dtype(dt).itemsize
                else:
                    m = int(self.lib_sizes[np.argmin(np.abs(np.log(self.lib_sizes) - np.log(n)))])
                    tk = srv.submit_decompress(self.library[(dt, eb, m)], dtype=dt)
                    b = m * 3 * np.dtype(dt).itemsize
                offered += b
                pending.put((tk, kind, b))
            pending.put(None)

        th = threading.Thread(target=submitter)
        th.start()
        lat = {"compress": [], "decompress": []}
        nbytes = 0
        first, last, failed, batches = None, 0.0, 0, []
        while True:
            item = pending.get()
            if item is None:
                break
            tk, kind, b = item
            r = srv.poll(tk)
            if r.status != 0:
                failed += 1
                continue
            lat[kind].append((r.t_end - r.t_submit) * 1e3)
            nbytes += b
            first = r.t_submit if first is None else min(first, r.t_submit)
            last = max(last, r.t_end)
            batches.append(r.batch)
        th.join()
        arena = srv.arena_allocations()
        srv.close()
        allv = np.array(lat["compress"] + lat["decompress"])

        def pct(v, q):
            return round(float(np.percentile(v, q)), 3) if len(v) else None

        wall = last - first if first is not None else 0.0
        return {
            "requests": len(trace), "failed": failed, "seconds": seconds, "rate_per_s": rate,
            "p50_ms": pct(allv, 50), "p99_ms": pct(allv, 99),
            "compress_p50_ms": pct(lat["compress"], 50), "compress_p99_ms": pct(lat["compress"], 99),
            "decompress_p50_ms": pct(lat["decompress"], 50), "decompress_p99_ms": pct(lat["decompress"], 99),
            "aggregate_gbps": round(nbytes / wall / 1e9, 3) if wall else None,
            "offered_gbps": round(offered / seconds / 1e9, 3),
            "drain_seconds": round(wall, 2),
            "mean_batch": round(float(np.mean(batches)), 2) if batches else None,
            "max_batch": int(max(batches)) if batches else None,
            "arena_allocations": arena,
        }


def run_serving(args, eng, seconds):
    sb = ServingBench(args, eng)
    main = sb.run(args.serve_rate, seconds)
    main.update({"workers": args.serve_workers, "hbm_budget_gb": args.serve_budget_gb,
                 "latency": "submit to result on the host (host input -> host output, PCIe included)"})
    log(f"serving: {main['requests']} requests, p50 {main['p50_ms']} ms, p99 {main['p99_ms']} ms, "
        f"{main['aggregate_gbps']} GB/s")
    # saturation ladder: doubled arrival rates until the server no longer
    # keeps up with the offered load (or a request fails)
    ladder = []
    rate = args.serve_rate * 2
    for step in range(args.saturation_steps):
        r = sb.run(rate, args.saturation_seconds, seed_offset=10 + step)
        ladder.append({k: r[k] for k in ("rate_per_s", "requests", "failed", "offered_gbps", "aggregate_gbps",
                                          "drain_seconds", "p50_ms", "p99_ms", "mean_batch", "max_batch")})
        log(f"serving saturation: {rate:g} req/s offered {r['offered_gbps']} GB/s -> {r['aggregate_gbps']} GB/s, "
            f"p99 {r['p99_ms']} ms, mean batch {r['mean_batch']}")
        if r["failed"] or (r["aggregate_gbps"] or 0) < 0.9 * r["offered_gbps"]:
            break
        rate *= 2
    if ladder:
        cap = max(ladder + [main], key=lambda x: x["aggregate_gbps"] or 0)
        main["saturation"] = ladder
        main["capacity_gbps"] = cap["aggregate_gbps"]
        main["p99_at_capacity_ms"] = cap["p99_ms"]
        main["capacity_rate_per_s"] = cap["rate_per_s"]
        main["saturation_note"] = (f"traces of {args.saturation_seconds:g} s at doubled arrival rates; stops at the "
                                   "first rate whose aggregate falls below 90% of the offered load or that fails "
                                   "a request; capacity = the highest aggregate reached")
    return main


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="default", choices=["default", "plain", "batched", "md", "serving", "all"])
    ap.add_argument("--n", type=int, default=50_000_000)
    ap.add_argument("--eb", type=float, default=1e-3)
    ap.add_argument("--p", type=int, default=16)
    ap.add_argument("--threshold", type=int, default=2)
    ap.add_argument("--batch-frames", type=int, default=256)
    ap.add_argument("--batch-steps", type=int, default=3)
    ap.add_argument("--serve-seconds", type=float, default=60.0, help="length of the arrival trace")
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--serve-workers", type=int, default=4)
    ap.add_argument("--serve-max-batch", type=int, default=32)
    ap.add_argument("--serve-budget-gb", type=int, default=64)
    ap.add_argument("--saturation-steps", type=int, default=4, help="rungs of the serving saturation ladder")
    ap.add_argument("--saturation-seconds", type=float, default=6.0)
    ap.add_argument("--cpu-sample", type=int, default=2_000_000)
    ap.add_argument("--cpu-full", action="store_true",
                    help="CPU baseline on the whole plain frame, one timed run each (several minutes)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-sweep", action="store_true", help="skip the eb = 1e-2 / 1e-4 runs of the plain frame")
    ap.add_argument("--no-search", action="store_true", help="skip the (p, threshold) search of the plain frame")
    ap.add_argument("--rans-chunk", type=int, default=0,
                    help="symbols per rANS chunk for the plain frame instead of the tuner's choice (profiling runs)")
    ap.add_argument("--no-per-frame", action="store_true",
                    help="batched/md: skip the one-call-per-frame comparator (profiling runs)")
    ap.add_argument("--tuning", default="",
                    help="batched/md: fixed lanes,chunk_symbols,l1_tiles instead of the tuner (profiling runs)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="one 1 GiB device-to-device copy before the plain frame (profiler counter scale)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU; --gpus above 1 is not supported "
                                   "(frames and requests shard across GPUs without exchange; see the design notes)"}))
        sys.exit(2)
    from chaincompressor_amd import engine, workloads
    eng = engine.Engine(0)
    wl = args.workload
    K = args.steps
    line = {
        "metric": "compress/decompress GB/s, batched vs plain at fixed eb; serving p50/p99 ms",
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded generators in chaincompressor_amd/workloads.py)",
    }
    t_start = time.perf_counter()
    pts = None
    md_ctx = None
    if wl in ("default", "plain", "all"):
        pts, nbytes, m, sweep = run_plain(args, eng)
        n = args.n
        line.update({
            "value": round(m["compress_gbps"], 3),
            "ms_per_step": round(m["ms_per_step"], 3),
            "config": {"workload": f"plain single frame: {n} float32 xyz Gaussian clusters + 10% background",
                       "eb": args.eb, "p": args.p, "threshold": args.threshold, "container": "LCH1",
                       "step": "compress (HBM -> HBM) then decompress (HBM -> HBM, float64)",
                       "engine": "autotuned entropy coding and L1 strategy, compression replayed from a HIP graph"},
            "compress_gbps": round(m["compress_gbps"], 3),
            "decompress_gbps": round(m["decompress_gbps"], 3),
            "compressed_bytes": m["bytes"],
            "rans_setting": m["rans"],
            "bits_per_particle": round(m["bytes"] * 8 / n, 3),
            "max_abs_error": m["maxerr"],
            "host_copies": {"h2d_input_gbps": m["h2d_input_gbps"], "d2h_decoded_gbps": m["d2h_decoded_gbps"],
                            "memory": "pageable host arrays, one hipMemcpy each; not part of value"},
            "eb_sweep": sweep,
            "parameter_search": m["search"],
            "regions_compress": region_table(m["c_steps"], K),
            "regions_decompress": region_table(m["d_steps"], K),
            "pipeline_alg_frac": {"compress": pipeline_fraction(m["c_steps"], K),
                                  "decompress": pipeline_fraction(m["d_steps"], K)},
            "model": dict(m["model"], key_bits=m["key_bits"], explicit_records=m["explicit_records"],
                          pre_entropy_bytes_Y=m["pre_entropy_bytes"]),
            "roofline": roofline([m["c_steps"], m["d_steps"]], K, profiled=True),
        })
        log(f"plain done at {time.perf_counter() - t_start:.0f} s")
    if wl in ("default", "batched", "all"):
        line["batched"], _ = run_batched(args, eng)
        log(f"batched done at {time.perf_counter() - t_start:.0f} s")
        if "value" not in line:
            line["value"] = line["batched"]["compress_gbps_batched"]
            line["config"] = {"workload": "batched: 256 uneven float32 frames", "eb": 1e-3}
    if wl in ("default", "md", "all"):
        frames = workloads.md_fcc_batch()
        line["md"], md_ctx = run_batched(args, eng, frames=frames, eb=1e-4, dtype=np.float64, label="md")
        log(f"md done at {time.perf_counter() - t_start:.0f} s")
        if "value" not in line:
            line["value"] = line["md"]["decompress_gbps_batched"]
            line["config"] = {"workload": "md: 128 float64 fcc frames", "eb": 1e-4}
    if wl in ("default", "serving", "all"):
        line["serving"] = run_serving(args, eng, args.serve_seconds)
        log(f"serving done at {time.perf_counter() - t_start:.0f} s")
        if "value" not in line:
            line["value"] = line["serving"]["aggregate_gbps"]
            line["config"] = {"workload": "serving: Poisson 200 req/s, mixed requests", "budget_gb": 64}
    if not args.no_cpu_baseline:
        if pts is not None:
            if args.cpu_full:
                cb = cpu_baseline(pts, args.eb, args.p, args.threshold, runs=1, warmup=0)
                cb["sample"] = cb["sample"].replace("every {stride}th point of the frame, order kept",
                                                    "the whole frame")
            else:
                stride = max(1, args.n // args.cpu_sample)
                sample = np.ascontiguousarray(pts[::stride][:args.cpu_sample])
                cb = cpu_baseline(sample, args.eb, args.p, args.threshold)
                cb["sample"] = cb["sample"].replace("{stride}", str(stride))
            line["cpu_baseline"] = cb
        if md_ctx is not None:
            line["md"]["cpu_check"] = cpu_md_check(eng, md_ctx["frames"], md_ctx["streams"], 1e-4, args.p,
                                                   args.threshold)
        log(f"cpu baseline done at {time.perf_counter() - t_start:.0f} s")
    print(json.dumps(line))


if __name__ == "__main__":
    main()
