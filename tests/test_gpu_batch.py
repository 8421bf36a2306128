"""Parity of the batched plan (hpaBatch*: one persistent launch over many filter -> group ->
sum queries) with the oracle running each query's per-query plan, across the tunable
configurations and the autotuner; and of the per-query GPU plans with the oracle."""
import numpy as np
import pytest

from horsepower_amd import datagen as D
from horsepower_amd import queries as Q
from horsepower_amd.backend import Batch

pytestmark = pytest.mark.gpu
RTOL = 1e-12

# (rows, key range, threshold): small ranges take the LDS path, large ones the global-table
# path; one row; an empty selection; a range right at the LDS budget.
SHAPES = [(20000, 4, 50), (1, 1, 99), (70000, 3000, 30), (150000, 200000, 70),
          (5000, 17, 0), (333333, 50, 95), (2000, 2000, 60), (100000, 2730, 45)]


def make_tables():
    ts = [D.gen_batch_query(q, n, card, seed=77) for q, (n, card, _) in enumerate(SHAPES)]
    neg = D.gen_batch_query(99, 40000, 600, seed=77)
    neg["key"] = (neg["key"] - 300).astype(np.int32)          # keys on both sides of zero
    return ts + [neg], [thr for _, _, thr in SHAPES] + [55]


def expected(orc, ts, thrs):
    out = []
    for t, thr in zip(ts, thrs):
        if not (t["flt"] < thr).any():
            out.append(None)
            continue
        out.append(Q.batch_plain(orc, Q.load_table(orc, t), thr))
    return out


def check(res, exp, tag):
    for q, (r, o) in enumerate(zip(res, exp)):
        if o is None:
            assert len(r["key"]) == 0 and len(r["count"]) == 0, (tag, q)
            continue
        assert np.array_equal(r["key"], o["key"]), (tag, q)
        assert np.array_equal(r["count"], o["count"]), (tag, q)
        assert np.allclose(r["sum"], o["sum"], rtol=RTOL, atol=0), (tag, q)


def test_batch_vs_oracle(gpu, orc):
    ts, thrs = make_tables()
    exp = expected(orc, ts, thrs)
    m = gpu.mark()
    b = Batch(gpu, [Q.load_table(gpu, t) for t in ts], thrs)
    try:
        for graph in (False, True, True):
            b.run(graph)
            check(b.results(), exp, ("graph", graph))
        for cfg in [(256, 4096, 0, 0), (1024, 65536, 2730, 2048), (512, 16384, 512, 1024), (64, 256, 100, 0),
                    (256, 8192, 0, 2048), (128, 8192, 2730, 1024)]:
            b.set_config(*cfg)
            b.run(True)
            check(b.results(), exp, cfg)
        for q in (0, 3, len(ts) - 1):                 # the node form of one query's results
            zk, zc, zs = b.result_nodes(q)
            r = b.results()[q]
            assert np.array_equal(gpu.read(zk), r["key"]) and np.array_equal(gpu.read(zc), r["count"])
            assert np.array_equal(gpu.read(zs), r["sum"])
        ms = b.autotune(1)
        assert len(ms) == 24 and all(t > 0 for t in ms)
        b.run(True)
        check(b.results(), exp, ("tuned", b.config()))
    finally:
        b.close()
        gpu.release(m)


def test_batch_pipelined_fetch(gpu, orc):
    """hpaBatchFetchBegin/End: results of run k copied to the host while run k+1 computes, two
    fetches in flight; every run's results equal the oracle's; a run with more groups than the
    snapshot took fails in End with HPA_E_LENGTH."""
    from horsepower_amd.vnode import HorseError
    ts, thrs = make_tables()
    exp = expected(orc, ts, thrs)
    m = gpu.mark()
    b = Batch(gpu, [Q.load_table(gpu, t) for t in ts], thrs)
    try:
        b.run(True)
        b.fetch_begin()
        for k in range(6):
            b.run(True)
            b.fetch_begin()
            check(b.fetch_end(), exp, ("pipelined", k))
        check(b.fetch_end(), exp, "last")
        with pytest.raises(HorseError):
            b.fetch_end()                                   # nothing begun
    finally:
        b.close()
        gpu.release(m)


def test_batch_graph_replays_with_eager_work_between(gpu, orc):
    """Replays of the captured plan, with eager primitive work and synchronisations between
    them, all give the eager result."""
    ts, thrs = make_tables()
    m = gpu.mark()
    T = [Q.load_table(gpu, t) for t in ts]
    b = Batch(gpu, T, thrs)
    try:
        b.run(False)
        ref = [{k: v.copy() for k, v in d.items()} for d in b.results()]      # results() are views
        for _ in range(3):
            Q.batch_fused(gpu, T[0], thrs[0])
            b.run(True)
            r = b.results()
            for q in range(len(ts)):
                for k in Q.BATCH_NAMES:
                    assert np.array_equal(r[q][k], ref[q][k]) or \
                        np.allclose(r[q][k], ref[q][k], rtol=RTOL, atol=0), (q, k)
    finally:
        b.close()
        gpu.release(m)


@pytest.mark.parametrize("n,card,thr", [(50000, 10, 50), (200000, 30000, 40), (1, 1, 99)])
def test_batch_per_query_plans(gpu, orc, n, card, thr):
    t = D.gen_batch_query(11, n, card)
    o = Q.batch_plain(orc, Q.load_table(orc, t), thr)
    Tg = Q.load_table(gpu, t)
    for fn in (Q.batch_plain, Q.batch_fused):
        g = fn(gpu, Tg, thr)
        assert np.array_equal(g["key"], o["key"]), fn.__name__
        assert np.array_equal(g["count"], o["count"]), fn.__name__
        assert np.allclose(g["sum"], o["sum"], rtol=RTOL, atol=0), fn.__name__
