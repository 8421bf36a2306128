"""Grouped reductions over a run-compressed i32 group (keys in runs of consecutive rows, over
2^24 rows: the group's flat list is read run by run).  Every reduction kind is checked against
the group's own flat list read back: integer sums, mins and maxes exactly, FP64 sums and
averages within relative 1e-12 of math.fsum per group.  The same values through the
interpreter's each(each_right(index, ...)) sequence must agree too."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
RTOL = 1e-12


def per_group(vals, counts, flat):
    starts = np.concatenate([[0], np.cumsum(counts)[:-1]])
    rows = vals[flat]
    return starts, rows


@pytest.mark.parametrize("recur", [False, True])
def test_group_reduce_over_runs(gpu, recur):
    r = np.random.default_rng(31 + recur)
    nk = 5_000_000
    keys = r.permutation(nk).astype(np.int64) * 3 - nk
    if recur:                                   # keys that come back in later runs: several runs a group
        keys = np.concatenate([keys, keys[r.integers(0, nk, nk // 4)]])
    x = np.repeat(keys, r.integers(1, 8, keys.size)).astype(np.int32)
    n = x.size
    assert n > (1 << 24)
    fv = np.round(r.uniform(0, 1e5, n), 2)
    iv = r.integers(-10**6, 10**6, n).astype(np.int32)
    m = gpu.mark()
    g = gpu.pfnGroup(gpu.vec(x))
    vals = gpu.pfnValues(g)
    _, counts, flat = gpu.read_groups(g)
    starts = np.concatenate([[0], np.cumsum(counts)[:-1]])
    fx, ix = gpu.vec(fv), gpu.vec(iv)
    fs, fa, fmn, fmx = (gpu.read(gpu.group_reduce(k, fx, vals)) for k in (0, 1, 2, 3))
    isum, imn, imx = (gpu.read(gpu.group_reduce(k, ix, vals)) for k in (0, 2, 3))
    frows, irows = fv[flat], iv[flat].astype(np.int64)
    assert np.array_equal(isum, np.add.reduceat(irows, starts))
    assert np.array_equal(imn, np.minimum.reduceat(irows, starts))
    assert np.array_equal(imx, np.maximum.reduceat(irows, starts))
    assert np.array_equal(fmn, np.minimum.reduceat(frows, starts))
    assert np.array_equal(fmx, np.maximum.reduceat(frows, starts))
    # exact sums of a sample of groups, plus every group against numpy's pairwise sum
    pick = r.choice(counts.size, 20000, replace=False)
    exact = np.array([math.fsum(frows[starts[k]:starts[k] + counts[k]]) for k in pick])
    assert np.allclose(fs[pick], exact, rtol=RTOL, atol=0)
    assert np.allclose(fa[pick], exact / counts[pick], rtol=RTOL, atol=0)
    assert np.allclose(fs, np.add.reduceat(frows, starts), rtol=1e-10, atol=0)
    # the interpreter's sequence over the same group lists
    seq = gpu.read(gpu.pfnRaze(gpu.each(gpu.each_right(fx, vals, "pfnIndex"), "pfnSum")))
    assert np.allclose(seq, fs, rtol=RTOL, atol=0)
    gpu.release(m)
