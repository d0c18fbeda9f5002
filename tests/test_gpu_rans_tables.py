"""rANS table bytes of the LCH1 writer against the oracle, on frames built so
that the in-block offsets go rANS with the table shapes the compact table
format treats specially: zero runs longer than 64 symbols (one byte per 64
zeros) and probabilities of 2^14 or more (two extra bytes).

Points are 1-D with eb = 0.5 (so q = floor(x - min)) and p = 4096, so the
offset of point i is exactly the value placed in it."""
import numpy as np
import pytest

import oracle as O
from chaincompressor_amd import container

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    from chaincompressor_amd import engine
    return engine.Engine(0)


def _frame(rel, seed):
    rng = np.random.default_rng(seed)
    n = rel.size
    blocks = rng.integers(0, 64, size=n)
    x = blocks.astype(np.float64) * 4096.0 + rel.astype(np.float64) + 0.5
    x[0] = 0.5   # minimum 0: q = floor(x)
    rel[0] = 0
    return x.reshape(n, 1)


def _sparse_runs(n, seed):
    # 31 distinct offsets 0, 100, ..., 3000: zero runs of 99 symbols
    rng = np.random.default_rng(seed)
    vals = np.arange(31) * 100
    w = 1.0 / (1.0 + np.arange(31))
    return vals[rng.choice(31, size=n, p=w / w.sum())]


def _heavy_symbol(n, seed):
    # one offset holds 60 % of the points, 150 others share the rest: 151
    # distinct values give 15-bit precision, the heavy one about 0.6 * 2^15
    rng = np.random.default_rng(seed)
    rel = rng.integers(1, 151, size=n)
    rel[rng.random(n) < 0.6] = 0
    return rel


@pytest.mark.parametrize("maker,seed", [(_sparse_runs, 1), (_heavy_symbol, 2), (_sparse_runs, 3)])
def test_rans_table_shapes(eng, maker, seed):
    n = 200_000
    pts = _frame(maker(n, seed), seed)
    print("case", maker.__name__, seed, flush=True)
    b, _ = eng.compress(pts, 0.5, 4096, 2, fmt="lch1")
    ref, kinds = O.compress_streams(pts, 0.5, 4096, 2).lch1()
    assert container.describe(ref)["components"]["rel"]["kind"] == "RANS"
    assert b == ref, kinds
    got = eng.decompress(b)
    assert np.array_equal(got.view(np.uint64), O.decode(b).view(np.uint64))
    assert np.all(np.abs(got - pts) <= 0.5)
