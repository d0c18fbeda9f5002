"""Write the golden vectors under tests/golden/ from the compiled reference.

Runs only in a container that holds the reference sources: it loads the two
reference CPU modules that ``make -C oracle ref`` compiles into oracle/_ref/
(lcps_cpp from cpp/chain/lcps_cpp.cpp and lcps_cpu_decompress from
cpp/decompress/lcps_cpu_decompress_serial.cpp) and records, per case, the
inputs and what the reference returns:

* ``lcp3``    lcps_compress2(..., ec=False) bytes (raw integer streams),
* ``lcp2``    lcps_compress2(..., ec=True) bytes (Zstd-22 / RVN1 chooser),
* ``stats``   the rest of lcps_compress2's return list,
* ``decoded`` lcps_cpu_decompress.lcps_decompress_serial(lcp2) (float64),
* ``decoded3`` lcps_cpp.lcps_decompress2(lcp3).

Large frames keep the LCP2 bytes and SHA-256 digests of the rest.

    make -C oracle ref && python tests/golden/make_golden.py
"""
from __future__ import annotations

import contextlib
import hashlib
import os
import sys
import xml.etree.ElementTree as ET

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle", "_ref"))
sys.path.insert(0, ROOT)

import lcps_cpp  # noqa: E402  (compiled reference, oracle/_ref)
import lcps_cpu_decompress  # noqa: E402

from chaincompressor_amd import workloads  # noqa: E402

REFERENCE = "/root/reference"


@contextlib.contextmanager
def quiet_stdout():
    """The reference printf()s stage times and codec names; keep them out."""
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


def u8(b: bytes) -> np.ndarray:
    return np.frombuffer(b, dtype=np.uint8).copy()


def sha(b) -> str:
    return hashlib.sha256(bytes(b)).hexdigest()


def reference_run(pts, eb, p, t, bl=True, l1=True, l2=True):
    with quiet_stdout():
        r2 = lcps_cpp.lcps_compress2(pts, eb, p, t, 0, bl, l1, l2, True)
        r3 = lcps_cpp.lcps_compress2(pts, eb, p, t, 0, bl, l1, l2, False)
        dec = lcps_cpu_decompress.lcps_decompress_serial(r2[0])
        dec3 = lcps_cpp.lcps_decompress2(r3[0])
    assert np.array_equal(dec.view(np.uint64), dec3.view(np.uint64))
    return r2, r3, dec


def small_cases():
    rng = np.random.default_rng(7)
    f32, f64 = np.float32, np.float64
    cases = []

    def add(name, pts, eb, p, t, **kw):
        cases.append((name, np.ascontiguousarray(pts), eb, p, t, kw))

    add("uniform_f32_d3", rng.random((1000, 3)).astype(f32), 1e-3, 16, 2)
    add("uniform_f64_d3_scaled", rng.random((1000, 3)) * 100.0, 1e-2, 4, -3)
    add("d1", rng.random((3000, 1)).astype(f32), 1e-3, 16, 2)
    add("d2", rng.random((5000, 2)).astype(f32), 1e-3, 8, 1)
    add("d4", rng.random((2000, 4)).astype(f32), 1e-3, 16, 2)
    add("n1", rng.random((1, 3)).astype(f32), 1e-3, 16, 2)
    add("n2_same", np.array([[0.25, 0.5, 0.75], [0.25, 0.5, 0.75]], dtype=f64), 1e-3, 16, 0)
    add("one_block", (rng.random((500, 3)) * 1e-2).astype(f32), 1e-3, 16, 2)
    add("p1", (rng.random((4000, 3)) * 0.2).astype(f32), 1e-3, 1, 2)
    add("huge_p", rng.random((1000, 3)).astype(f32), 1e-3, 1 << 40, 2)
    add("neg_threshold", (rng.random((4000, 3)) * 0.3).astype(f32), 1e-3, 4, -10)
    add("big_threshold", (rng.random((4000, 3)) * 0.3).astype(f32), 1e-3, 4, 40)
    add("corrections_f64", 1e6 + rng.random((2000, 3)) * 1e-6, 1e-10, 16, 2)
    add("corrections_f32_tiny_eb", (rng.random((1500, 3)) + 100.0).astype(f32), 1e-9, 1, 0)
    add("duplicates", np.repeat(rng.random((50, 3)), 40, axis=0).astype(f32), 1e-3, 16, 2)
    add("negative_coords", rng.normal(size=(3000, 3)).astype(f32), 1e-2, 16, 2)
    add("ablate_l2", (rng.random((6000, 3)) * 0.3).astype(f32), 1e-3, 16, 2, l2=False)
    add("ablate_bl", (rng.random((6000, 3)) * 0.3).astype(f32), 1e-3, 16, 2, bl=False)
    add("ablate_l1", (rng.random((6000, 3)) * 0.3).astype(f32), 1e-3, 16, 2, l1=False)
    cl = workloads.gaussian_clusters(20000, seed=11, clusters=8)
    add("clusters_chain_heavy", cl, 1e-3, 16, 2)
    add("clusters_p256", cl, 1e-3, 256, 4)
    sk = np.floor(rng.exponential(3.0, size=(8000, 3))).astype(f32) * 0.002
    add("skewed_small_alphabet", sk, 1e-3, 2, 1)
    add("zeros", np.zeros((100, 3), dtype=f64), 1e-3, 3, 0)
    add("signed_zero_min", np.array([[0.0, 1.0], [-0.0, 2.0], [0.5, -0.0]], dtype=f64), 1e-3, 2, 0)
    add("subnormal_eb", rng.random((64, 2)) * 1e-300, 1e-305, 4, 0)
    return cases


def write_small():
    out = {}
    names = []
    for name, pts, eb, p, t, kw in small_cases():
        r2, r3, dec = reference_run(pts, eb, p, t, **kw)
        names.append(name)
        out[f"{name}__points"] = pts
        out[f"{name}__params"] = np.array([eb, p, t, kw.get("bl", True), kw.get("l1", True), kw.get("l2", True)],
                                          dtype=object)
        out[f"{name}__lcp2"] = u8(r2[0])
        if len(r3[0]) <= 65536:
            out[f"{name}__lcp3"] = u8(r3[0])
        out[f"{name}__lcp3_sha"] = np.array(sha(r3[0]))
        out[f"{name}__stats"] = np.array(list(r2[1:]), dtype=np.int64)
        out[f"{name}__stats3"] = np.array(list(r3[1:]), dtype=np.int64)
        if dec.size <= 6000:
            out[f"{name}__decoded"] = dec
        out[f"{name}__decoded_sha"] = np.array(sha(dec.tobytes()))
    out["names"] = np.array(names)
    params = {k: v for k, v in out.items() if k.endswith("__params")}
    for k in params:
        out[k] = np.array([float(x) if i == 0 else int(x) for i, x in enumerate(out[k])], dtype=np.float64)
    np.savez_compressed(os.path.join(HERE, "small_cases.npz"), **out)
    print("small cases:", len(names))


def read_dambreak():
    folder = os.path.join(REFERENCE, "data", "DamBreak")
    files = sorted(f for f in os.listdir(folder) if f.startswith("0654") and f.endswith(".vtp"))
    root = ET.parse(os.path.join(folder, files[0])).getroot()
    for da in root.iter("DataArray"):
        if da.attrib.get("Name") == "Points":
            arr = np.array(da.text.split(), dtype=np.float32)
            return arr.reshape(-1, 3)
    raise RuntimeError("no Points array")


def large_entry(out, name, pts, eb, p, t, store_points):
    r2, r3, dec = reference_run(pts, eb, p, t)
    if store_points:
        out[f"{name}__points"] = pts
    out[f"{name}__points_sha"] = np.array(sha(np.ascontiguousarray(pts).tobytes()))
    out[f"{name}__params"] = np.array([eb, p, t], dtype=np.float64)
    out[f"{name}__lcp2"] = u8(r2[0])
    out[f"{name}__stats"] = np.array(list(r2[1:]), dtype=np.int64)
    out[f"{name}__lcp3_sha"] = np.array(sha(r3[0]))
    out[f"{name}__lcp3_len"] = np.array(len(r3[0]))
    out[f"{name}__decoded_sha"] = np.array(sha(dec.tobytes()))


def write_large():
    out = {}
    names = []
    db = read_dambreak()
    assert db.shape == (80703, 3)
    rng_max = (db.max(axis=0) - db.min(axis=0)).astype(np.float64)
    # reader.py:269-272 equivalent_eb(maximum, qb-1) with the GPU (p, t) table (reader.py:343-344)
    ps, ts = [2, 4, 16, 16], [2, 2, 16, 2]
    for k, qb in enumerate([11, 12, 13, 14]):
        eb = float(np.max(rng_max) / (2 ** ((qb - 1) + 1) - 2))
        name = f"dambreak_qb{qb}"
        large_entry(out, name, db, eb, ps[k], ts[k], store_points=(k == 0))
        names.append(name)
    frame = workloads.dambreak_like()
    large_entry(out, "cfg1_dambreak_like", frame, 1e-3, 16, 2, store_points=False)
    names.append("cfg1_dambreak_like")
    out["names"] = np.array(names)
    np.savez_compressed(os.path.join(HERE, "large_cases.npz"), **out)
    print("large cases:", names)


if __name__ == "__main__":
    write_small()
    write_large()
