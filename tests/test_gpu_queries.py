"""End-to-end parity of the benchmark queries (plain interpreter sequence and the fused
compiled form) on the GPU against the oracle running the same plans."""
import math

import numpy as np
import pytest

from horsepower_amd import datagen as D
from horsepower_amd import queries as Q

pytestmark = pytest.mark.gpu
RTOL = 1e-12


def close(a, b, key):
    a, b = np.asarray(a), np.asarray(b)
    if a.dtype.kind == "f":
        assert np.allclose(a, b, rtol=RTOL, atol=0), (key, a, b)
    else:
        assert np.array_equal(a, b), (key, a, b)


@pytest.mark.parametrize("n", [1000, 1_000_003])
def test_q6(gpu, orc, n):
    t = D.gen_q6(n)
    Tg, To = Q.load_table(gpu, t), Q.load_table(orc, t)
    o = Q.q6_plain(orc, To)["revenue"]
    m = (t["ship_day"] >= D.Q6_LO) & (t["ship_day"] < D.Q6_HI) & (t["quantity"] < 24) & \
        (t["discount"] >= 0.05) & (t["discount"] <= 0.07)
    exact = math.fsum(t["price"][m] * t["discount"][m])
    for fn in (Q.q6_plain, Q.q6_fused):
        g = fn(gpu, Tg)["revenue"]
        assert abs(g - o) <= RTOL * abs(o), fn.__name__
        assert abs(g - exact) <= RTOL * abs(exact), fn.__name__


@pytest.mark.parametrize("n", [5000, 600_001])
def test_q1(gpu, orc, n):
    t = D.gen_q1(n)
    Tg, To = Q.load_table(gpu, t), Q.load_table(orc, t)
    o = Q.q1_plain(orc, To)
    for fn in (Q.q1_plain, Q.q1_fused):
        g = fn(gpu, Tg)
        for k in Q.Q1_NAMES:
            close(g[k], o[k], (fn.__name__, k))


@pytest.mark.parametrize("nd", [3000, 200_000])
def test_q3(gpu, orc, nd):
    dm, f = D.gen_q3(nd)
    o = Q.q3_plain(orc, Q.load_table(orc, dm), Q.load_table(orc, f))
    Dg, Fg = Q.load_table(gpu, dm), Q.load_table(gpu, f)
    for fn in (Q.q3_plain, Q.q3_fused):
        g = fn(gpu, Dg, Fg)
        assert g["rows"] == o["rows"] and g["rows"] > 0, fn.__name__
        for k in Q.Q3_NAMES:
            close(g[k], o[k], (fn.__name__, k))


@pytest.mark.parametrize("n", [3000, 1_000_003])
def test_q1_one_pass(gpu, orc, n):
    t = D.gen_q1(n)
    o = Q.q1_plain(orc, Q.load_table(orc, t))
    g = Q.q1_chf(gpu, Q.load_table(gpu, t))
    for k in Q.Q1_NAMES:
        close(g[k], o[k], k)


def test_q1_one_pass_large(gpu):
    """34M rows, where the one-pass kernel runs three workgroups per CU.  The oracle's plan is
    sized for a few million rows (at 34M it stops with a segmentation fault), so the check is
    the GPU's interpreter sequence (group + per-group sums, checked against the oracle at 1M
    rows in test_q1) and exact per-group row counts from numpy."""
    n = 34_000_001
    t = D.gen_q1(n)
    T = Q.load_table(gpu, t)
    o = Q.q1_plain(gpu, T)
    g = Q.q1_chf(gpu, T)
    for k in Q.Q1_NAMES:
        close(g[k], o[k], k)
    sel = t["ship_day"] <= D.Q1_SHIP_MAX
    code = t["return_flag"][sel].astype(np.int64) * 256 + t["line_status"][sel]
    _, counts = np.unique(code, return_counts=True)
    assert np.asarray(g["count_order"]).tolist() == counts.tolist()


def test_fused_group_agg_overflow_and_single_key(gpu, orc):
    from horsepower_amd.vnode import H_C, expr
    r = np.random.default_rng(5)
    n = 200_000
    k0 = r.integers(0, 200, n).astype(np.uint8)           # > 8 codes per workgroup: overflow
    v = np.round(r.uniform(0, 100, n), 2)
    res = gpu.fused_group_agg([gpu.vec(k0, H_C)], [gpu.vec(v)], None, [("sum", expr([("COL", 0)]))])
    assert res is None
    k1 = r.choice(np.frombuffer(b"xyz", np.uint8), n)
    zf, (zk,), (zs, zc) = gpu.fused_group_agg([gpu.vec(k1, H_C)], [gpu.vec(v)], None,
                                              [("sum", expr([("COL", 0)])), ("len", None)])
    keys, counts, flat = orc.read_groups(orc.pfnGroup(orc.vec(k1, H_C)))
    assert gpu.read(zk).tolist() == sorted(set(k1.tolist()))
    assert gpu.read(zf).tolist() == keys.tolist()
    assert gpu.read(zc).tolist() == counts.tolist()
    sums = [math.fsum(v[g]) for g in np.split(flat, np.cumsum(counts)[:-1])]
    assert np.allclose(gpu.read(zs), sums, rtol=RTOL, atol=0)


@pytest.mark.parametrize("keys,cap", [(b"xyz", 64), (b"xyz", 2), (b"q", 64), (b"abcdefgh", 8)])
def test_fused_group_agg_host(gpu, keys, cap):
    """The host-delivering form returns what the node form returns, read back: with every group
    staged, with more groups than `cap` (read from the nodes), with a single group."""
    from horsepower_amd.vnode import H_C, expr
    r = np.random.default_rng(len(keys) + cap)
    n = 300_000
    k0 = r.choice(np.frombuffer(keys, np.uint8), n)
    k1 = r.choice(np.frombuffer(b"A" if len(keys) == 1 else b"AB", np.uint8), n)
    v = np.round(r.uniform(0, 100, n), 2)
    w = r.integers(0, 50, n).astype(np.int32)
    spec = [("sum", expr([("COL", 0)])), ("avg", expr([("COL", 1)])), ("sum", expr([("COL", 1)])),
            ("max", expr([("COL", 0)])), ("len", None)]
    cond = expr([("COL", 1), ("I", 40), "LT"])
    ks, xs = [gpu.vec(k0, H_C), gpu.vec(k1, H_C)], [gpu.vec(v), gpu.vec(w)]
    _, zk, za = gpu.fused_group_agg(ks, xs, cond, spec)
    want = gpu.read_many(zk + za)
    hk, ha = gpu.fused_group_agg_host(ks, xs, cond, spec, cap=cap)
    for a, b in zip(hk + ha, want):
        assert a.dtype == b.dtype and np.array_equal(a, b)


@pytest.mark.parametrize("ngroups", [64, 65, 128])
def test_fused_group_agg_many_groups(gpu, ngroups):
    """Groups in contiguous row ranges: each workgroup meets a few codes, the whole column up to
    128 (the merge slices the partials up to 64 groups and takes them in one block beyond)."""
    from horsepower_amd.vnode import H_C, expr
    r = np.random.default_rng(ngroups)
    n = 1 << 22
    idx = (np.arange(n, dtype=np.int64) * ngroups) // n
    k0 = (65 + idx % 16).astype(np.uint8)
    k1 = (65 + idx // 16).astype(np.uint8)
    v = np.round(r.uniform(0, 100, n), 2)
    w = r.integers(-50, 50, n).astype(np.int32)
    spec = [("sum", expr([("COL", 0)])), ("avg", expr([("COL", 0)])), ("min", expr([("COL", 1)])),
            ("max", expr([("COL", 0)])), ("sum", expr([("COL", 1)])), ("len", None)]
    res = gpu.fused_group_agg([gpu.vec(k0, H_C), gpu.vec(k1, H_C)], [gpu.vec(v), gpu.vec(w)], None, spec)
    assert res is not None
    zf, (za, zb), (zs, zav, zmn, zmx, zis, zc) = res
    want = sorted(set(zip(k0.tolist(), k1.tolist())))
    assert len(want) == ngroups
    assert list(zip(gpu.read(za).tolist(), gpu.read(zb).tolist())) == want
    code = k0.astype(np.int64) * 256 + k1
    first, sums, mins, maxs, isums, cnts = [], [], [], [], [], []
    for a, b in want:
        rows = np.flatnonzero(code == a * 256 + b)
        first.append(rows[0]); cnts.append(rows.size); sums.append(math.fsum(v[rows]))
        mins.append(w[rows].min()); maxs.append(v[rows].max()); isums.append(int(w[rows].sum()))
    assert gpu.read(zf).tolist() == first
    assert gpu.read(zc).tolist() == cnts
    assert np.allclose(gpu.read(zs), sums, rtol=RTOL, atol=0)
    assert np.allclose(gpu.read(zav), np.array(sums) / np.array(cnts), rtol=RTOL, atol=0)
    assert gpu.read(zmn).tolist() == mins and gpu.read(zmx).tolist() == maxs
    assert gpu.read(zis).tolist() == isums


def test_fused_group_agg_key_order(gpu):
    """The one-pass group emits its groups in ascending key order with char bytes compared as
    signed values, the order pfnOrderBy gives the key columns (bytes >= 128 sort first)."""
    from horsepower_amd.vnode import H_C, expr
    r = np.random.default_rng(6)
    n = 50_000
    a = r.choice(np.array([3, 100, 200, 250], np.uint8), n)
    b = r.choice(np.array([1, 130], np.uint8), n)
    v = np.round(r.uniform(0, 100, n), 2)
    _, (za, zb), (zs,) = gpu.fused_group_agg([gpu.vec(a, H_C), gpu.vec(b, H_C)], [gpu.vec(v)], None,
                                             [("sum", expr([("COL", 0)]))])
    got = list(zip(gpu.read(za).view(np.int8).tolist(), gpu.read(zb).view(np.int8).tolist()))
    sa, sb = a.view(np.int8), b.view(np.int8)
    want = sorted(set(zip(sa.tolist(), sb.tolist())))
    assert got == want
    sums = [math.fsum(v[(sa == pa) & (sb == pb)]) for pa, pb in want]
    assert np.allclose(gpu.read(zs), sums, rtol=RTOL, atol=0)


@pytest.mark.parametrize("n,sel", [(0, 0.5), (1, 1.0), (5000, 0.5), (1_000_003, 0.54), (200_000, 1.0), (70_000, 0.0)])
def test_compress_map(gpu, orc, n, sel):
    """hpaCompressMap: plain columns compressed and FE maps of other columns evaluated at the
    selected rows, against the unfused compress + map sequence on the oracle (per-operation IEEE,
    so bit for bit)."""
    from horsepower_amd.vnode import H_B, H_E, H_I, H_L, expr
    r = np.random.default_rng(n + 3)
    m = (r.random(n) < sel).astype(np.uint8)
    k = r.integers(0, 1 << 30, n).astype(np.int32)
    d = r.integers(0, 11, n) / 100.0
    p = r.integers(90000, 10500000, n) / 100.0
    q = r.integers(1, 50, n).astype(np.int32)
    rev = expr([("I", 1), ("COL", 0), "SUB", ("COL", 1), "MUL"])
    big = expr([("COL", 2), ("I", 24), "LT"])
    (zk,), (zr, zb) = gpu.compress_map(gpu.vec(m), [gpu.vec(k)], [(H_E, rev), (H_B, big)],
                                       [gpu.vec(d), gpu.vec(p), gpu.vec(q)])
    mo = orc.vec(m)
    ok = orc.read(orc.pfnCompress(mo, orc.vec(k)))
    orv = orc.read(orc.pfnMul(orc.pfnMinus(orc.scalar(1, H_L), orc.pfnCompress(mo, orc.vec(d))), orc.pfnCompress(mo, orc.vec(p))))
    assert np.array_equal(gpu.read(zk), ok)
    assert np.array_equal(gpu.read(zr), orv)
    assert np.array_equal(gpu.read(zb), (q[m != 0] < 24).astype(np.uint8))


@pytest.mark.parametrize("cuts", [(2300, 2434), (2434, 2300, 2300), (1000, 1263, 1264, 2500), (2600, -1), (2434,)])
def test_q1_many(gpu, orc, cuts):
    """Several Q1 requests in one bucketed pass (the serving loop's shared Q1 scan) equal each
    request run alone (q1_chf) and the oracle's interpreter sequence: keys and counts exactly,
    FP64 within 1e-12.  Cutoffs repeat, straddle the status boundary (1263), lie past every row
    and below every row (no group)."""
    t = D.gen_q1(700_001, seed=1009)
    Tg, To = Q.load_table(gpu, t), Q.load_table(orc, t)
    m = gpu.mark()
    many = Q.q1_many(gpu, Tg, [dict(ship_max=c) for c in cuts])
    for c, r in zip(cuts, many):
        alone = Q.q1_chf(gpu, Tg, c)
        for k in Q.Q1_NAMES:
            close(r[k], alone[k], (c, k))
        if c < t["ship_day"].min():
            # no row selected: no group (the interpreter's raze of an empty group list is an error)
            assert all(np.asarray(r[k]).size == 0 for k in Q.Q1_NAMES)
            continue
        o = Q.q1_plain(orc, To, c)
        for k in Q.Q1_NAMES:
            close(r[k], o[k], (c, k))
    gpu.release(m)
