"""hpaFusedCompressMap (the FA compress scan of one condition, plain columns and FE maps, the
condition evaluated by the count pass) against hpaCompressMap with the condition materialised as
a byte mask by the compare primitives: equal outputs, bit for bit, for conditions selecting some,
none and all rows, over sizes with a partial last tile, one row and no rows; and hpaFusedCompress
(plain columns) against hpaCompressMany the same way."""
import numpy as np
import pytest

from horsepower_amd.vnode import H_E, H_I, H_L, expr

pytestmark = pytest.mark.gpu

REV = expr([("I", 1), ("COL", 0), "SUB", ("COL", 1), "MUL"])           # (1 - disc) * price


@pytest.mark.parametrize("n", [1_000_003, 4096 * 7, 17, 1, 0])
@pytest.mark.parametrize("cut", [-1, 500, 2000])
def test_fused_compress_map_vs_mask(gpu, n, cut):
    r = np.random.default_rng(7 * n + cut + 10)
    disc = np.round(r.uniform(0, 0.1, n), 2)
    price = np.round(r.uniform(900, 105000, n), 2)
    day = r.integers(0, 2000, n).astype(np.int32)
    key = r.integers(-2**31, 2**31 - 1, n).astype(np.int32)
    big = r.integers(-2**62, 2**62, n).astype(np.int64)
    m = gpu.mark()
    D, P, S, K, G = gpu.vec(disc), gpu.vec(price), gpu.vec(day), gpu.vec(key), gpu.vec(big)
    cond = expr([("COL", 2), ("I", cut), "GT"])
    (zk, zg), (zr,) = gpu.fused_compress_map(cond, [K, G], [(H_E, REV)], [D, P, S])
    mask = gpu.pfnGt(S, gpu.scalar(cut, H_I))
    (rk, rg), (rr,) = gpu.compress_map(mask, [K, G], [(H_E, REV)], [D, P])
    got = gpu.read_many([zk, zg, zr])
    ref = gpu.read_many([rk, rg, rr])
    sel = day > cut
    assert got[0].size == int(sel.sum())
    for a, b in zip(got, ref):
        assert a.dtype == b.dtype and np.array_equal(a, b)
    assert np.array_equal(got[0], key[sel]) and np.array_equal(got[1], big[sel])
    # plain columns through hpaFusedCompress with a two-column condition
    cond2 = expr([("COL", 0), ("I", cut), "GT", ("COL", 1), ("I", 0), "LT", "AND"])
    zs = gpu.fused_compress([K, G], [S, K], cond2)
    sel2 = (day > cut) & (key < 0)
    out = gpu.read_many(zs)
    assert np.array_equal(out[0], key[sel2]) and np.array_equal(out[1], big[sel2])
    gpu.release(m)
