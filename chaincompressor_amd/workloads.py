"""Seeded synthetic particle frames used by the benchmark and the tests.

Each generator is deterministic for a given seed (numpy PCG64 via
``default_rng``). None of them is needed to use the codec.

* ``dambreak_like``   -- 80,703 float32 points: a jittered cubic lattice
  (spacing 0.01, jitter U(+-0.002)) filling a 0.6 x 0.3 fluid column bottom-up
  in lattice order (x fastest, then y, then z: 44 full layers of 1,800 sites
  plus 1,503 sites), the top layer lifted by 0.05*sin(2*pi*x/0.6), then 5 % of
  the indices exchanged in disjoint random pairs.
* ``gaussian_clusters`` -- a union of 64 Gaussian clusters (sigma log-uniform
  in [0.005, 0.05], centres uniform in the unit cube, equal weights) plus 10 %
  uniform background, clipped to [0, 1); cluster-major order, each cluster's
  points independent draws (so in random order), background last.
* ``uneven_batch``    -- frames with log-uniform particle counts, each half a
  jittered lattice filling the unit cube and half a 16-cluster Gaussian mix.
* ``md_fcc_batch``    -- fcc lattice sites (4-atom basis, lattice constant 1)
  in an m^3 periodic box, m = ceil((N/4)^(1/3)), the first N sites in ID order,
  plus N(0, 0.05^2) displacements wrapped into the box; float64.
* ``serving_trace``   -- Poisson arrivals with mixed request kinds, sizes,
  dtypes and error bounds.
"""
from __future__ import annotations

import math

import numpy as np

BASE_SEED = 20261203


def dambreak_like(seed: int = BASE_SEED + 1, n: int = 80_703) -> np.ndarray:
    rng = np.random.default_rng(seed)
    spacing = 0.01
    nx, ny = 60, 30
    k = np.arange(n)
    ix = k % nx
    iy = (k // nx) % ny
    iz = k // (nx * ny)
    pts = np.stack([(ix + 0.5) * spacing, (iy + 0.5) * spacing, (iz + 0.5) * spacing], axis=1)
    pts += rng.uniform(-0.002, 0.002, size=pts.shape)
    top = iz == iz.max()
    pts[top, 2] += 0.05 * np.sin(2.0 * np.pi * pts[top, 0] / 0.6)
    nswap = int(round(0.05 * n)) // 2
    chosen = rng.choice(n, size=2 * nswap, replace=False)
    a, b = chosen[:nswap], chosen[nswap:]
    tmp = pts[a].copy()
    pts[a] = pts[b]
    pts[b] = tmp
    return pts.astype(np.float32)


def gaussian_clusters(n: int, seed: int = BASE_SEED + 2, clusters: int = 64,
                      background: float = 0.10, dtype=np.float32) -> np.ndarray:
    """Cluster-major: every cluster's points are consecutive and in random
    order (independent draws), the uniform background comes last."""
    rng = np.random.default_rng(seed)
    dtype = np.dtype(dtype).type
    nb = int(round(background * n))
    nc = n - nb
    counts = rng.multinomial(nc, np.full(clusters, 1.0 / clusters))
    centres = rng.random((clusters, 3))
    sigmas = np.exp(rng.uniform(math.log(0.005), math.log(0.05), size=clusters))
    out = np.empty((n, 3), dtype=dtype)
    hi = np.nextafter(dtype(1.0), dtype(0.0))
    pos = 0
    for c in range(clusters):
        m = int(counts[c])
        blk = out[pos:pos + m]
        rng.standard_normal((m, 3), dtype=dtype, out=blk)
        blk *= dtype(sigmas[c])
        blk += centres[c].astype(dtype)
        np.clip(blk, 0.0, hi, out=blk)
        pos += m
    bg = out[pos:]
    rng.random((nb, 3), dtype=dtype, out=bg)
    np.clip(bg, 0.0, hi, out=bg)
    return out


def _lattice_unit_cube(n: int, rng, dtype=np.float64) -> np.ndarray:
    side = max(1, int(math.ceil(n ** (1.0 / 3.0))))
    spacing = 1.0 / side
    k = np.arange(n, dtype=np.int64)
    pts = np.empty((n, 3), dtype=dtype)
    pts[:, 0] = k % side
    pts[:, 1] = (k // side) % side
    pts[:, 2] = k // (side * side)
    pts += 0.5
    pts *= spacing
    pts += rng.uniform(-0.2 * spacing, 0.2 * spacing, size=pts.shape).astype(dtype)
    return pts


def _map(fn, items, workers):
    if workers <= 1 or len(items) <= 1:
        return [fn(x) for x in items]
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=workers) as ex:
        return list(ex.map(fn, items))


def uneven_batch(frames: int = 256, lo: float = 1e4, hi: float = 4e6, seed: int = BASE_SEED + 3,
                 dtype=np.float32, workers: int = 8) -> list:
    """Frame f uses its own generator seeded by (seed, f), so the batch does
    not depend on the number of worker threads."""
    sizes = np.round(np.exp(np.random.default_rng(seed).uniform(math.log(lo), math.log(hi), size=frames)))
    top = np.nextafter(np.dtype(dtype).type(1.0), np.dtype(dtype).type(0.0))

    def one(f):
        n = int(sizes[f])
        rng = np.random.default_rng([seed, f])
        half = n // 2
        pts = np.empty((n, 3), dtype=dtype)
        pts[:half] = _lattice_unit_cube(half, rng, dtype)
        pts[half:] = gaussian_clusters(n - half, seed=seed * 1000 + f, clusters=16, dtype=dtype)
        np.clip(pts, 0.0, top, out=pts)
        return pts

    return _map(one, list(range(frames)), workers)


def md_fcc_frame(n: int, rng, sigma: float = 0.05) -> np.ndarray:
    m = int(math.ceil((n / 4.0) ** (1.0 / 3.0)))
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    k = np.arange(n, dtype=np.int64)
    cell = k // 4
    pts = np.empty((n, 3), dtype=np.float64)
    pts[:, 0] = cell % m
    pts[:, 1] = (cell // m) % m
    pts[:, 2] = cell // (m * m)
    pts += basis[k % 4]
    pts += rng.normal(0.0, sigma, size=pts.shape)
    return np.mod(pts, float(m))


def md_fcc_batch(frames: int = 128, lo: int = 100_000, hi: int = 2_000_000,
                 seed: int = BASE_SEED + 4, workers: int = 8) -> list:
    """Frame f uses its own generator seeded by (seed, f)."""
    sizes = np.random.default_rng(seed).integers(lo, hi + 1, size=frames)
    return _map(lambda f: md_fcc_frame(int(sizes[f]), np.random.default_rng([seed, f])),
                list(range(frames)), workers)


def serving_trace(rate: float = 200.0, seconds: float = 60.0, seed: int = BASE_SEED + 5,
                  lo: float = 5e3, hi: float = 2e7):
    """List of (arrival_s, kind, n, dtype, eb); kind is 'compress' or 'decompress'."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= seconds:
            break
        kind = "compress" if rng.random() < 0.7 else "decompress"
        n = int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))
        dt = np.float32 if rng.random() < 0.8 else np.float64
        eb = float(rng.choice([1e-2, 1e-3, 1e-4]))
        out.append((t, kind, n, dt, eb))
    return out
