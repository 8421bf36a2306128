"""Serving: device table generation against its host twin, and each request's result the
same alone and in the concurrent mix (and the same as the per-query plans and the oracle)."""
import numpy as np
import pytest

from horsepower_amd import datagen as D
from horsepower_amd import queries as Q
from horsepower_amd import serve as S

pytestmark = pytest.mark.gpu
RTOL = 1e-12


@pytest.fixture(scope="module")
def tables(gpu):
    dim, fact = gpu.gen_serve_tables(60000, 1004)
    return {"dim": dim, "fact": fact}


def test_device_generator_matches_host(gpu):
    dim, fact = gpu.gen_serve_tables(5003, 77)
    hd, hf = D.gen_serve(5003, 77)
    for k in D.SERVE_DIM:
        assert np.array_equal(gpu.read(dim[k]), hd[k]), k
    for k in D.SERVE_FACT:
        a, b = gpu.read(fact[k]), hf[k]
        assert np.array_equal(a.view(np.uint8) if a.dtype.itemsize == 1 else a, b), k


def test_q6_shared_scan_equals_single(gpu, tables):
    reqs = [r for r in S.make_requests(60, seed=5) if r.kind == "q6"][:8]
    many = Q.q6_many(gpu, tables["fact"], [r.params for r in reqs])
    for r, m in zip(reqs, many):
        one = Q.q6_fused(gpu, tables["fact"], **r.params)["revenue"]
        plain = Q.q6_plain(gpu, tables["fact"], **r.params)["revenue"]
        assert m["revenue"] == one
        assert abs(plain - one) <= RTOL * abs(one)


def test_serve_alone_vs_mix(gpu, orc, tables):
    reqs_mix = S.make_requests(40, rate=None, seed=11)
    reqs_alone = S.make_requests(40, rate=None, seed=11)
    srv = S.Server(gpu, tables, slots=4, pool_bytes=2 << 30, merge=True)
    try:
        est = srv.calibrate()
        assert set(est) == {"q6", "q1", "q3"}
        srv.serve(reqs_mix, realtime=False)
        assert all(r.done is not None for r in reqs_mix)
        assert srv.peak_in_use <= srv.pool_bytes
    finally:
        srv.close()
    one = S.Server(gpu, tables, slots=1, pool_bytes=2 << 30, merge=False)
    try:
        one.calibrate()
        one.serve(reqs_alone, realtime=False)
    finally:
        one.close()
    assert S.same_results(reqs_mix, reqs_alone, RTOL) == []
    # and the per-query plans on the default context, and the oracle for a few requests
    serial = S.serve_serial(gpu, tables, S.make_requests(40, rate=None, seed=11), plans="plain", realtime=False)
    assert S.same_results(reqs_mix, serial, RTOL) == []
    dim, fact = D.gen_serve(60000, 1004)
    To = {"dim": Q.load_table(orc, dim), "fact": Q.load_table(orc, fact)}
    for r in reqs_mix[:12]:
        o = S.PLANS["plain"][r.kind](orc, To, r.params)
        for k, v in o.items():
            g = np.asarray(r.result[k])
            if g.dtype.kind == "f":
                assert np.allclose(g, np.asarray(v), rtol=RTOL, atol=0), (r.rid, r.kind, k)
            else:
                assert np.array_equal(g, np.asarray(v)), (r.rid, r.kind, k)


def test_serve_budget_and_requeue(gpu, tables):
    """A pool that holds about one request at a time still serves everything; a request whose
    estimate is too small is requeued with a larger arena slice."""
    reqs = S.make_requests(16, rate=None, seed=3)
    srv = S.Server(gpu, tables, slots=4, pool_bytes=256 << 20, merge=True)
    try:
        est = srv.calibrate()
        assert max(est.values()) < 200 << 20
        srv.est["q3"] = 8 << 20                       # too small: HPA_E_MEMORY, then requeued
        srv.serve(reqs, realtime=False)
        assert all(r.done is not None for r in reqs)
        assert any(r.tries > 0 for r in reqs if r.kind == "q3")
        assert srv.peak_in_use <= 256 << 20
    finally:
        srv.close()
