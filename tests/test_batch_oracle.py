"""The oracle running the batched workload's per-query plan, pinned against a brute-force
group-by (numpy unique/bincount and exactly rounded sums)."""
import math

import numpy as np
import pytest

from horsepower_amd import datagen as D
from horsepower_amd import queries as Q


def brute(t, thr):
    m = t["flt"] < thr
    k, v = t["key"][m], t["val"][m]
    keys = np.unique(k.astype(np.uint32)).astype(np.int32)       # pfnGroup: unsigned key order
    counts = np.array([(k == x).sum() for x in keys], dtype=np.int64)
    sums = np.array([math.fsum(v[k == x]) for x in keys])
    return keys, counts, sums


@pytest.mark.parametrize("n,card,thr", [(2000, 7, 50), (5000, 300, 20), (1, 1, 99)])
def test_batch_plain_oracle(orc, n, card, thr):
    t = D.gen_batch_query(3, n, card)
    o = Q.batch_plain(orc, Q.load_table(orc, t), thr)
    keys, counts, sums = brute(t, thr)
    assert np.array_equal(o["key"], keys)
    assert np.array_equal(o["count"], counts)
    assert np.allclose(o["sum"], sums, rtol=1e-12, atol=0)


def test_batch_plain_oracle_negative_keys(orc):
    """Keys on both sides of zero: pfnGroup orders i32 keys as unsigned values, so the
    negative keys come after the non-negative ones."""
    t = D.gen_batch_query(5, 4000, 40)
    t["key"] = t["key"] - 20
    o = Q.batch_plain(orc, Q.load_table(orc, t), 60)
    keys, counts, sums = brute(t, 60)
    assert keys[0] == 0 and keys[-1] == -1
    assert np.array_equal(o["key"], keys) and np.array_equal(o["count"], counts)
    assert np.allclose(o["sum"], sums, rtol=1e-12, atol=0)


def test_batch_plain_empty_selection_raises(orc):
    """No row passes: raze of the empty group list is an error in the reference
    (h_primitive.cc:937-938), so the per-query plan raises; the batched plan returns no groups."""
    from horsepower_amd.vnode import HorseError
    t = D.gen_batch_query(3, 300, 5)
    with pytest.raises(HorseError) as e:
        Q.batch_plain(orc, Q.load_table(orc, t), 0)
    assert e.value.code == 11
