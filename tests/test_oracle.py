"""CPU tests pinning the oracle (oracle/horse_oracle.c) before it is trusted as the checker.

Each check is against something other than the oracle's own formula: the reference's golden
vectors (tests/golden/kat.json), textbook routines (numpy, math.fsum), brute force on small
inputs, and closed forms of the reference's ordering rules.
"""
import math

import numpy as np
import pytest

import kat
from horsepower_amd import datagen as D
from horsepower_amd import queries as Q
from horsepower_amd.vnode import H_B, H_C, H_E, H_I, H_L, HorseError

rng = np.random.default_rng(7)


@pytest.mark.parametrize("e", kat.KAT, ids=[e["name"] for e in kat.KAT])
def test_kat(orc, e):
    if e["fn"] not in kat.SUPPORTED:
        pytest.skip("primitive outside the hot path")
    kat.check(orc, e)


def test_epsilon_compare_kat(orc):
    # src/horseir/v0/backend/test_pfns.h:62-73: 0.06+0.01 == 0.07 under the 1e-13 epsilon
    a = orc.vec(np.array([0.06 + 0.01, 0.1 + 0.2], np.float64))
    b = orc.vec(np.array([0.07, 0.3], np.float64))
    assert orc.read(orc.pfnEq(a, b)).tolist() == [1, 1]
    assert orc.read(orc.pfnLt(a, b)).tolist() == [0, 0]
    assert ((0.06 + 0.01) == 0.07) is False          # exact compare would say no
    c = orc.vec(np.array([1.0, 1.0 + 2e-13], np.float64))
    assert orc.read(orc.pfnEq(c, orc.scalar(1.0, H_E))).tolist() == [1, 0]


def test_int_compare_textbook(orc):
    x = rng.integers(-50, 50, 1000).astype(np.int32)
    y = rng.integers(-50, 50, 1000).astype(np.int32)
    for name, f in (("pfnLt", np.less), ("pfnLeq", np.less_equal), ("pfnGt", np.greater),
                    ("pfnGeq", np.greater_equal), ("pfnEq", np.equal), ("pfnNeq", np.not_equal)):
        got = orc.read(orc.new(name, orc.vec(x), orc.vec(y)))
        assert np.array_equal(got, f(x, y).astype(np.uint8)), name


def test_arith_textbook(orc):
    x = rng.uniform(-1e6, 1e6, 1000)
    y = rng.uniform(1, 1e3, 1000)
    for name, f in (("pfnPlus", np.add), ("pfnMinus", np.subtract), ("pfnMul", np.multiply), ("pfnDiv", np.divide)):
        assert np.array_equal(orc.read(orc.new(name, orc.vec(x), orc.vec(y))), f(x, y)), name
    xi = np.array([2**31 - 1, -2**31, 5], np.int32)
    assert orc.read(orc.pfnPlus(orc.vec(xi), orc.scalar(1, H_I))).tolist() == [-2**31, -2**31 + 1, 6]
    with pytest.raises(HorseError) as ei:
        orc.pfnDiv(orc.vec(xi), orc.vec(np.array([1, 0, 2], np.int32)))
    assert ei.value.code == 4


def test_logic_nand_nor_always_true(orc):
    # logicNand/Nor return ~(x&y) / ~(x|y) as C++ bool: nonzero for every 0/1 input
    a = orc.vec(np.array([0, 0, 1, 1], np.uint8), H_B)
    b = orc.vec(np.array([0, 1, 0, 1], np.uint8), H_B)
    assert orc.read(orc.pfnAnd(a, b)).tolist() == [0, 0, 0, 1]
    assert orc.read(orc.pfnOr(a, b)).tolist() == [0, 1, 1, 1]
    assert orc.read(orc.pfnXor(a, b)).tolist() == [0, 1, 1, 0]
    assert orc.read(orc.pfnNand(a, b)).tolist() == [1, 1, 1, 1]
    assert orc.read(orc.pfnNor(a, b)).tolist() == [1, 1, 1, 1]


@pytest.mark.parametrize("dt", [np.int32, np.float64, np.uint8, np.int64])
def test_compress_where_textbook(orc, dt):
    n = 50000
    m = (rng.random(n) < 0.3).astype(np.uint8)
    y = (rng.integers(0, 100, n)).astype(dt)
    typ = H_C if dt == np.uint8 else None
    got = orc.read(orc.pfnCompress(orc.vec(m, H_B), orc.vec(y, typ)))
    assert np.array_equal(got, y[m.astype(bool)])
    assert np.array_equal(orc.read(orc.pfnWhere(orc.vec(m, H_B))), np.nonzero(m)[0])


def test_compress_edges(orc):
    y = orc.vec(np.arange(5, dtype=np.int32))
    assert orc.read(orc.pfnCompress(orc.vec(np.zeros(5, np.uint8), H_B), y)).tolist() == []
    assert orc.read(orc.pfnCompress(orc.vec(np.ones(5, np.uint8), H_B), y)).tolist() == [0, 1, 2, 3, 4]
    assert orc.read(orc.pfnCompress(orc.vec(np.array([0, 0, 1, 0, 0], np.uint8), H_B), y)).tolist() == [2]
    with pytest.raises(HorseError):
        orc.pfnCompress(orc.vec(np.ones(4, np.uint8), H_B), y)


def test_sum_exact_reference(orc):
    x = rng.uniform(0, 1e5, 200001)
    s = orc.read_scalar(orc.pfnSum(orc.vec(x)))
    exact = math.fsum(x)
    assert abs(s - exact) <= 1e-12 * abs(exact)
    import oracle
    assert abs(oracle.neumaier(x) - exact) <= 1e-15 * abs(exact)
    xi = rng.integers(-10**6, 10**6, 10001).astype(np.int32)
    assert orc.read_scalar(orc.pfnSum(orc.vec(xi))) == int(xi.astype(np.int64).sum())
    assert orc.read(orc.pfnAvg(orc.vec(xi)))[0] == int(xi.astype(np.int64).sum()) * 1.0 / xi.shape[0]


def test_minmax_textbook(orc):
    x = rng.normal(size=1001)
    assert orc.read_scalar(orc.pfnMin(orc.vec(x))) == x.min()
    assert orc.read_scalar(orc.pfnMax(orc.vec(x))) == x.max()
    e = orc.vec(np.zeros(0, np.int32))
    assert orc.read_scalar(orc.pfnMin(e)) == np.iinfo(np.int64).max
    assert orc.read_scalar(orc.pfnMax(orc.vec(np.zeros(0, np.float64)))) == np.finfo(np.float64).tiny


def brute_group(keys, order_key):
    """groups in ascending order_key(value); key = first row; rows ascending."""
    first, rows = {}, {}
    for i, k in enumerate(keys.tolist()):
        rows.setdefault(k, []).append(i)
        first.setdefault(k, i)
    ks = sorted(rows, key=order_key)
    return [first[k] for k in ks], [rows[k] for k in ks]


def test_group_int_unsigned_order(orc):
    x = rng.integers(-20, 20, 3000).astype(np.int32)
    keys, counts, flat = orc.read_groups(orc.pfnGroup(orc.vec(x)))
    bk, br = brute_group(x, lambda v: v & 0xffffffff)
    assert keys.tolist() == bk
    assert [g.tolist() for g in np.split(flat, np.cumsum(counts)[:-1])] == br


def test_group_long_regimes(orc):
    # no key has (x>>32) > 0: sort on the low 32 bits only; runs cut on the full value
    x = np.array([-1, 4294967295, -1, 7, 7], np.int64)
    keys, counts, flat = orc.read_groups(orc.pfnGroup(orc.vec(x)))
    assert keys.tolist() == [3, 0, 1, 2] and counts.tolist() == [2, 1, 1, 1]
    assert flat.tolist() == [3, 4, 0, 1, 2]
    y = rng.integers(-5, 5, 500).astype(np.int64) * (1 << 33)
    keys, counts, flat = orc.read_groups(orc.pfnGroup(orc.vec(y)))
    bk, br = brute_group(y, lambda v: v & 0xffffffffffffffff)
    assert keys.tolist() == bk


def test_group_trie(orc):
    a = rng.choice(np.frombuffer(b"ANR", np.uint8), 2000)
    b = rng.choice(np.frombuffer(b"FO", np.uint8), 2000)
    keys, counts, flat = orc.read_groups(orc.pfnGroup(orc.list([orc.vec(a, H_C), orc.vec(b, H_C)])))
    codes = a.astype(np.int64) * 256 + b
    bk, br = brute_group(codes, lambda v: v)
    assert keys.tolist() == bk
    assert [g.tolist() for g in np.split(flat, np.cumsum(counts)[:-1])] == br


def test_order_single_int_radix(orc):
    x = rng.integers(-30, 30, 2000).astype(np.int32)
    up = orc.read(orc.pfnOrderBy(orc.vec(x), orc.bools([1])))
    st = np.argsort(x.view(np.uint32), kind="stable")
    assert np.array_equal(up, st)
    dn = orc.read(orc.pfnOrderBy(orc.vec(x), orc.bools([0])))
    assert np.array_equal(dn, st[::-1])


def test_order_case1_packed(orc):
    a = rng.integers(0, 5, 1000).astype(np.int32)
    b = rng.integers(-3, 3, 1000).astype(np.int32)
    got = orc.read(orc.pfnOrderBy(orc.list([orc.vec(a), orc.vec(b)]), orc.bools([1, 1])))
    k = np.where(b < 0, np.int64(-1) << 32 | (b.astype(np.int64) & 0xffffffff),
                 (a.astype(np.int64) << 32) | (b.astype(np.int64) & 0xffffffff))
    hi = ((k >> 32) > 0).any()
    kk = k.view(np.uint64) if hi else (k & 0xffffffff).astype(np.uint64)
    assert np.array_equal(got, np.argsort(kk, kind="stable"))


def test_order_list_lexicographic(orc):
    n = 3000
    a = rng.integers(0, 50, n).astype(np.float64) / 7
    b = rng.permutation(n).astype(np.int32)        # distinct: no full-key ties
    got = orc.read(orc.pfnOrderBy(orc.list([orc.vec(a), orc.vec(b)]), orc.bools([0, 1])))
    ref = sorted(range(n), key=lambda i: (-a[i], b[i]))
    assert got.tolist() == ref


def test_order_quicksort_is_a_sort(orc):
    x = rng.integers(0, 20, 5000).astype(np.float64)
    p = orc.read(orc.pfnOrderBy(orc.vec(x), orc.bools([1])))
    assert sorted(p.tolist()) == list(range(5000))
    assert np.all(np.diff(x[p]) >= 0)


def brute_join(x, y):
    build, probe, swap = (y, x, True) if len(x) > len(y) else (x, y, False)
    pairs = [(j, i) for j in range(len(probe)) for i in range(len(build)) if build[i] == probe[j]]
    if swap:
        return [p[0] for p in pairs], [p[1] for p in pairs]
    return [p[1] for p in pairs], [p[0] for p in pairs]


@pytest.mark.parametrize("nx,ny", [(40, 300), (300, 40), (100, 100), (5, 50), (50, 5)])
def test_join_order_closed_form(orc, nx, ny):
    x = rng.integers(0, 30, nx).astype(np.int32)
    y = rng.integers(0, 30, ny).astype(np.int32)
    j0, j1 = orc.read_list(orc.pfnJoinIndex(orc.vec(x), orc.vec(y), orc.sym("eq")))
    if min(nx, ny) >= 10:
        b0, b1 = brute_join(x, y)
        assert j0.tolist() == b0 and j1.tolist() == b1
    else:
        # linear path: for each probe row, the first matching build row (h_libs.cc:1542-1575)
        small, big, swap = (y, x, True) if nx > ny else (x, y, False)
        exp = [(next(i for i in range(len(small)) if small[i] == big[j]), j) for j in range(len(big)) if big[j] in small]
        e0 = [p[1] if swap else p[0] for p in exp]
        e1 = [p[0] if swap else p[1] for p in exp]
        assert j0.tolist() == e0 and j1.tolist() == e1


def test_index_of_first_occurrence(orc):
    x = rng.integers(0, 40, 300).astype(np.int32)
    y = rng.integers(0, 50, 500).astype(np.int32)
    got = orc.read(orc.pfnIndexOf(orc.vec(x), orc.vec(y)))
    first = {}
    for i, v in enumerate(x.tolist()):
        first.setdefault(v, i)
    assert got.tolist() == [first.get(v, len(x)) for v in y.tolist()]
    xs = np.unique(x)                               # ordered x: binary search
    got = orc.read(orc.pfnIndexOf(orc.vec(xs), orc.vec(y)))
    pos = {v: i for i, v in enumerate(xs.tolist())}
    assert got.tolist() == [pos.get(v, len(xs)) for v in y.tolist()]


def test_member_textbook(orc):
    x = rng.integers(0, 100, 1000).astype(np.int64)
    s = rng.integers(0, 100, 40).astype(np.int64)
    assert np.array_equal(orc.read(orc.pfnMember(orc.vec(x), orc.vec(s))), np.isin(x, s).astype(np.uint8))


def test_q6_matches_numpy(orc):
    t = D.gen_q6(100000)
    r = Q.q6_plain(orc, Q.load_table(orc, t))
    m = (t["ship_day"] >= D.Q6_LO) & (t["ship_day"] < D.Q6_HI) & (t["quantity"] < 24) & \
        (t["discount"] >= 0.05) & (t["discount"] <= 0.07)
    exact = math.fsum(t["price"][m] * t["discount"][m])
    assert abs(r["revenue"] - exact) <= 1e-12 * exact


def test_q1_matches_numpy(orc):
    t = D.gen_q1(60000)
    r = Q.q1_plain(orc, Q.load_table(orc, t))
    m = t["ship_day"] <= D.Q1_SHIP_MAX
    code = t["return_flag"][m].astype(np.int64) * 256 + t["line_status"][m]
    groups = sorted(set(code.tolist()))
    assert len(groups) == 4
    for gi, cde in enumerate(groups):
        sel = code == cde
        p, d, tx = t["price"][m][sel], t["discount"][m][sel], t["tax"][m][sel]
        assert r["count_order"][gi] == sel.sum()
        assert r["sum_qty"][gi] == t["quantity"][m][sel].astype(np.int64).sum()
        assert abs(r["sum_charge"][gi] - math.fsum(p * (1 - d) * (1 + tx))) <= 1e-12 * r["sum_charge"][gi]


def test_heap_full_returns_memory_error(orc):
    """A primitive whose result does not fit the oracle's heap returns HPA_E_MEMORY (100)
    instead of writing through a null payload (the reference stops with "Heap full!!",
    h_memory.cc:110)."""
    import oracle
    dll = oracle.load()
    n = 1 << 20
    m = orc.vec(np.ones(n, np.uint8))
    m_np = np.ones(n, np.uint8)
    m_np[::3] = 0
    mask = orc.vec(m_np)
    col = orc.vec(np.arange(n, dtype=np.float64))
    ints = orc.vec(rng.integers(0, 1000, n).astype(np.int32))
    mark = orc.mark()
    old = dll.orc_set_heap_top(dll.orc_get_heap() + (1 << 16))
    try:
        for name, args in (("pfnCompress", (mask, col)), ("pfnGroup", (ints,)), ("pfnWhere", (mask,)),
                           ("pfnMul", (col, col)), ("pfnOrderBy", None)):
            with pytest.raises(HorseError) as ei:
                if args is None:
                    orc.pfnOrderBy(orc.list([ints, col]), orc.bools([True, False]))
                else:
                    orc.new(name, *args)
            assert ei.value.code == 100, name
    finally:
        dll.orc_set_heap_top(old)
        orc.release(mark)
    assert orc.read(orc.pfnSum(m)).tolist() == [n]


def _join_multi_brute(xs, ys, ops):
    """All pairs (i, j) satisfying every column's operator, probe-major: sorted by (j, i)."""
    f = {"eq": np.equal, "neq": np.not_equal, "lt": np.less, "leq": np.less_equal,
         "gt": np.greater, "geq": np.greater_equal}
    ok = np.ones((len(xs[0]), len(ys[0])), bool)
    for a, b, op in zip(xs, ys, ops):
        ok &= f[op](a.astype(np.result_type(a, b))[:, None], b.astype(np.result_type(a, b))[None, :])
    i, j = np.nonzero(ok)
    o = np.lexsort((i, j))
    return i[o], j[o]


@pytest.mark.parametrize("ops", [("eq", "lt"), ("lt", "eq", "neq"), ("eq", "eq", "geq"), ("gt", "leq", "eq")])
def test_join_multi_brute_force(orc, ops):
    """pfnJoinIndexMultiple (h_primitive.cc:2095-2110): the first eq column hash-joined (build on
    x, probe y, no swap), the rest filtered pair by pair; the pairs are exactly the nested-loop
    matches in probe-major order."""
    r = np.random.default_rng(len(ops) * 7 + len(ops[0]))
    for nx, ny in ((40, 70), (300, 25), (5, 9)):
        xs, ys = [], []
        for q, op in enumerate(ops):
            if op == "eq" and q == ops.index("eq"):
                xs.append(r.integers(0, 12, nx).astype(np.int32)); ys.append(r.integers(0, 12, ny).astype(np.int32))
            elif q % 2:
                xs.append(np.round(r.uniform(0, 4, nx), 1)); ys.append(r.integers(0, 4, ny).astype(np.int64))
            else:
                xs.append(r.integers(0, 4, nx).astype(np.int32)); ys.append(r.integers(0, 4, ny).astype(np.int16))
        z = orc.pfnJoinIndex(orc.list([orc.vec(a) for a in xs]), orc.list([orc.vec(b) for b in ys]),
                             orc.vec(np.array([orc.symbol(o) for o in ops], np.int64), 9))
        got = orc.read_list(z)
        i, j = _join_multi_brute(xs, ys, ops)
        assert got[0].tolist() == i.tolist() and got[1].tolist() == j.tolist(), (ops, nx, ny)


def test_join_single_pair_after_swap(orc):
    """A one-pair join result is two inline scalars; swapping the sides (swap2,
    h_system.cc:47-56, when x is the longer input) keeps each value with its side."""
    x = np.arange(100, 140, dtype=np.int32)
    y = np.array([117], np.int32)
    z = orc.read_list(orc.pfnJoinIndex(orc.vec(x), orc.vec(y), orc.sym("eq")))
    assert z[0].tolist() == [17] and z[1].tolist() == [0]


def test_enum_closed_form(orc):
    """pfnEnum (h_primitive.cc:2198-2222): the index of every value among the keys (index_of,
    first occurrence), E_ENUM_INDEX (10) for a value that is not a key; keys/values/fetch give
    back the keys, the index vector and the values."""
    r = np.random.default_rng(21)
    keys = r.permutation(50_000).astype(np.int32) * 3 + 1
    vals = keys[r.integers(0, keys.size, 200_000)]
    e = orc.pfnEnum(orc.vec(keys), orc.vec(vals))
    pos = {int(k): i for i, k in enumerate(keys)}
    assert orc.read(orc.pfnValues(e)).tolist() == [pos[int(v)] for v in vals]
    assert np.array_equal(orc.read(orc.pfnKeys(e)), keys)
    assert np.array_equal(orc.read(orc.pfnFetch(e)), vals)
    bad = vals.copy()
    bad[777] = 2                                     # not of the form 3k + 1
    with pytest.raises(HorseError) as ei:
        orc.pfnEnum(orc.vec(keys), orc.vec(bad))
    assert ei.value.code == 10
