"""GPU parity of every hot-path primitive against the oracle on the same seeded inputs.

Integer, byte and index results are compared bit-exactly; FP64 aggregates within relative
1e-12 of the oracle and of an exactly rounded sum (math.fsum).  Element-wise FP64 arithmetic
is per-operation IEEE on both sides and compared exactly.
"""
import math

import numpy as np
import pytest

import kat
from horsepower_amd.vnode import H_B, H_C, H_D, H_E, H_I, H_L, HorseError

pytestmark = pytest.mark.gpu
RTOL = 1e-12

SIZES = [0, 1, 2, 7, 1000, 4097, 100003]


def rng(seed):
    return np.random.default_rng(seed)


def both(gpu, orc, fn):
    return fn(gpu), fn(orc)


@pytest.mark.parametrize("e", kat.KAT, ids=[e["name"] for e in kat.KAT])
def test_kat_gpu(gpu, e):
    if e["fn"] not in kat.SUPPORTED:
        pytest.skip("primitive outside the hot path")
    print("kat", e["name"], flush=True)
    kat.check(gpu, e)


@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("dt", [np.int32, np.float64, np.int64, np.int16])
def test_compare_vector_scalar(gpu, orc, n, dt):
    r = rng(n + 1)
    x = (r.integers(-100, 100, n) / (7 if dt == np.float64 else 1)).astype(dt)
    y = (r.integers(-100, 100, n) / (7 if dt == np.float64 else 1)).astype(dt)
    for op in ("pfnLt", "pfnLeq", "pfnGt", "pfnGeq", "pfnEq", "pfnNeq"):
        for B in ():
            pass
        g = gpu.read(gpu.new(op, gpu.vec(x), gpu.vec(y)))
        o = orc.read(orc.new(op, orc.vec(x), orc.vec(y)))
        assert np.array_equal(g, o), op
        s = dt(3) if dt != np.float64 else 0.5
        g = gpu.read(gpu.new(op, gpu.vec(x), gpu.vec(np.array([s], dt))))
        o = orc.read(orc.new(op, orc.vec(x), orc.vec(np.array([s], dt))))
        assert np.array_equal(g, o), op


def test_compare_epsilon_and_promotion(gpu, orc):
    a = np.array([0.06 + 0.01, 0.1 + 0.2, 1.0, 1.0 + 2e-13, -3.0], np.float64)
    b = np.array([0.07, 0.3, 1.0, 1.0, -3.0 + 5e-14], np.float64)
    for op in ("pfnLt", "pfnLeq", "pfnEq", "pfnNeq", "pfnGt", "pfnGeq"):
        assert np.array_equal(gpu.read(gpu.new(op, gpu.vec(a), gpu.vec(b))),
                              orc.read(orc.new(op, orc.vec(a), orc.vec(b)))), op
    xi = np.arange(-5, 5, dtype=np.int32)
    for B in (gpu, orc):
        B._r = B.read(B.pfnLt(B.vec(xi), B.scalar(0.5, H_E)))          # i32 promoted to f64
    assert np.array_equal(gpu._r, orc._r)


@pytest.mark.parametrize("n", SIZES)
def test_arith(gpu, orc, n):
    r = rng(n + 2)
    x = r.uniform(-1e3, 1e3, n)
    y = r.uniform(1, 1e3, n)
    xi = r.integers(-2**31, 2**31 - 1, n).astype(np.int32)
    yi = r.integers(1, 100, n).astype(np.int32)
    for op in ("pfnPlus", "pfnMinus", "pfnMul", "pfnDiv"):
        assert np.array_equal(gpu.read(gpu.new(op, gpu.vec(x), gpu.vec(y))),
                              orc.read(orc.new(op, orc.vec(x), orc.vec(y)))), op
        assert np.array_equal(gpu.read(gpu.new(op, gpu.vec(xi), gpu.vec(yi))),
                              orc.read(orc.new(op, orc.vec(xi), orc.vec(yi)))), op
        g = gpu.read(gpu.new(op, gpu.scalar(1, H_I), gpu.vec(y)))          # q1: 1 - discount
        o = orc.read(orc.new(op, orc.scalar(1, H_I), orc.vec(y)))
        assert np.array_equal(g, o), op


def test_arith_div_zero(gpu):
    with pytest.raises(HorseError) as ei:
        gpu.pfnDiv(gpu.vec(np.arange(10, dtype=np.int32)), gpu.vec(np.arange(10, dtype=np.int32)))
    assert ei.value.code == 4


@pytest.mark.parametrize("n", SIZES)
def test_logic(gpu, orc, n):
    r = rng(n + 3)
    a = r.integers(0, 2, n).astype(np.uint8)
    b = r.integers(0, 2, n).astype(np.uint8)
    for op in ("pfnAnd", "pfnOr", "pfnXor", "pfnNand", "pfnNor"):
        assert np.array_equal(gpu.read(gpu.new(op, gpu.vec(a, H_B), gpu.vec(b, H_B))),
                              orc.read(orc.new(op, orc.vec(a, H_B), orc.vec(b, H_B)))), op


@pytest.mark.parametrize("n", SIZES + [1 << 20])
@pytest.mark.parametrize("p", [0.0, 0.02, 0.5, 1.0])
def test_compress_where(gpu, orc, n, p):
    r = rng(n + 4)
    m = (r.random(n) < p).astype(np.uint8)
    cols = [r.integers(-1000, 1000, n).astype(np.int32), r.uniform(0, 1, n),
            r.integers(65, 90, n).astype(np.uint8), r.integers(0, 9, n).astype(np.int64)]
    for y in cols:
        typ = H_C if y.dtype == np.uint8 else None
        g = gpu.read(gpu.pfnCompress(gpu.vec(m, H_B), gpu.vec(y, typ)))
        o = orc.read(orc.pfnCompress(orc.vec(m, H_B), orc.vec(y, typ)))
        assert np.array_equal(g, o)
    assert np.array_equal(gpu.read(gpu.pfnWhere(gpu.vec(m, H_B))), np.nonzero(m)[0])
    zs = gpu.compress_many(gpu.vec(m, H_B), [gpu.vec(cols[0]), gpu.vec(cols[1])])
    assert np.array_equal(gpu.read(zs[0]), cols[0][m.astype(bool)])
    assert np.array_equal(gpu.read(zs[1]), cols[1][m.astype(bool)])


@pytest.mark.parametrize("n", SIZES)
def test_index_gather(gpu, orc, n):
    r = rng(n + 5)
    x = r.uniform(0, 1, max(n, 1))
    idx = r.integers(0, max(n, 1), n).astype(np.int64)
    assert np.array_equal(gpu.read(gpu.pfnIndex(gpu.vec(x), gpu.vec(idx))), x[idx])
    bad = np.array([0, len(x)], np.int64)
    with pytest.raises(HorseError) as ei:
        gpu.pfnIndex(gpu.vec(x), gpu.vec(bad))
    assert ei.value.code == 3


@pytest.mark.parametrize("n", [0, 1, 3, 5000])
def test_index_many(gpu, n):
    """The multi-column gather gives what pfnIndex gives column by column."""
    r = rng(n + 17)
    m = max(n, 1) + 3
    cols = [r.uniform(0, 1, m), r.integers(-9, 9, m).astype(np.int32),
            r.integers(0, 255, m).astype(np.uint8), r.integers(-2**40, 2**40, m).astype(np.int64)]
    typs = [H_E, H_I, H_C, H_L]
    xs = [gpu.vec(c, t) for c, t in zip(cols, typs)]
    idx = gpu.vec(r.integers(0, m, n).astype(np.int64))
    many = [gpu.read(z) for z in gpu.index_many(xs, idx)]
    for a, x in zip(many, xs):
        assert np.array_equal(a, gpu.read(gpu.pfnIndex(x, idx)))
    with pytest.raises(HorseError) as ei:
        gpu.index_many(xs, gpu.vec(np.array([0, m], np.int64)))
    assert ei.value.code == 3


@pytest.mark.parametrize("n", [0, 1, 5, 4096, 100003, 3_000_001])
def test_sum_avg_minmax(gpu, orc, n):
    r = rng(n + 6)
    x = np.round(r.uniform(0, 1e5, n), 2)
    g = gpu.read_scalar(gpu.pfnSum(gpu.vec(x)))
    exact = math.fsum(x)
    assert abs(g - exact) <= RTOL * abs(exact) + 1e-300
    assert abs(g - orc.read_scalar(orc.pfnSum(orc.vec(x)))) <= RTOL * abs(exact) + 1e-300
    xi = r.integers(-10**6, 10**6, n).astype(np.int32)
    assert gpu.read_scalar(gpu.pfnSum(gpu.vec(xi))) == int(xi.astype(np.int64).sum())
    if n:
        assert gpu.read_scalar(gpu.pfnMin(gpu.vec(x))) == x.min()
        assert gpu.read_scalar(gpu.pfnMax(gpu.vec(xi))) == xi.max()
        a = gpu.read_scalar(gpu.pfnAvg(gpu.vec(x)))
        assert abs(a - exact / n) <= RTOL * abs(exact / n)
    else:
        assert gpu.read_scalar(gpu.pfnMin(gpu.vec(xi))) == np.iinfo(np.int64).max


def groups_of(B, z):
    k, c, f = B.read_groups(z)
    return k.tolist(), c.tolist(), f.tolist()


@pytest.mark.parametrize("n,card", [(0, 1), (1, 1), (1000, 7), (100000, 50), (200000, 150000), (1 << 20, 1 << 16)])
def test_group_int(gpu, orc, n, card):
    r = rng(n + card)
    x = r.integers(-card // 2, card - card // 2, n).astype(np.int32)
    assert groups_of(gpu, gpu.pfnGroup(gpu.vec(x))) == groups_of(orc, orc.pfnGroup(orc.vec(x)))


@pytest.mark.parametrize("nkeys,maxrun,recur", [(1_000_000, 30, True), (1_300_000, 30, False), (150_000, 200, True),
                                                 (2000, 20000, True)])
def test_group_int_clustered(gpu, orc, nkeys, maxrun, recur):
    """Keys in runs of consecutive rows, over 2^24 rows (the run-compressed group): runs of
    1..maxrun rows, cut at the 64-row cap, runs far longer than the cap, keys that come back in
    later runs, negative keys (the group order is the unsigned order of the bit pattern).  Runs
    of a few rows at 80M rows: test_group_int_clustered_large."""
    r = rng(nkeys + maxrun)
    keys = r.permutation(nkeys).astype(np.int64) * 7 - nkeys * 3
    if recur:
        keys = np.concatenate([keys, keys[r.integers(0, nkeys, nkeys // 3)]])
    x = np.repeat(keys, r.integers(1, maxrun + 1, keys.size)).astype(np.int32)
    assert x.size > (1 << 24)
    gk, gc, gf = gpu.read_groups(gpu.pfnGroup(gpu.vec(x)))
    ok, oc, of = orc.read_groups(orc.pfnGroup(orc.vec(x)))
    assert np.array_equal(gk, ok) and np.array_equal(gc, oc) and np.array_equal(gf, of)


def test_group_int_clustered_large(gpu):
    """80M rows clustered in runs (past 2^26 rows the run cap shrinks to 32 rows): checked by the
    properties that define the stable-sort group -- the flat list is a permutation of the rows,
    keys are equal inside a group and strictly ascending (unsigned) from group to group, rows
    ascend inside a group, and each group's key is its first row."""
    r = rng(80)
    keys = (r.permutation(20_000_000).astype(np.int64) * 5 - 50_000_000).astype(np.int32)
    x = np.repeat(keys, r.integers(1, 8, keys.size))
    x[r.integers(0, x.size, 1000)] = 12345                      # a key that recurs far apart
    n = x.size
    assert n > (1 << 26)
    k, c, f = gpu.read_groups(gpu.pfnGroup(gpu.vec(x)))
    assert f.size == n and c.sum() == n and k.size == c.size
    seen = np.zeros(n, bool)
    seen[f] = True
    assert seen.all()
    starts = np.cumsum(c) - c
    head = np.zeros(n, bool)
    head[starts] = True
    xs = x[f].view(np.uint32)
    assert np.array_equal(xs[1:] != xs[:-1], head[1:])        # key changes exactly at group heads
    assert (np.diff(xs[starts].astype(np.int64)) > 0).all()    # groups in ascending unsigned order
    assert (np.diff(f)[~head[1:]] > 0).all()                   # rows ascend inside a group
    assert np.array_equal(k, f[starts])


@pytest.mark.parametrize("n", [1, 700, 5000, 70000])
def test_group_along_order(gpu, orc, n):
    """Lists and non-integer vectors group along pfnOrderBy's order with every flag true: two
    i32 columns take its case 1 (packed (x0 << 32) | x1, x1 sign-extended, so negative x1 values
    of different x0 share a key and equal rows can form separate runs), one i32 column its
    unsigned radix order, chars their byte order."""
    r = rng(n + 16)
    a = r.integers(-3, 3, n).astype(np.int32)
    b = r.integers(-2, 2, n).astype(np.int32)
    ch = r.integers(0, 256, n).astype(np.uint8)
    for cols in ([a, b], [a], [b, a], [a, b, a]):
        zg = gpu.pfnGroup(gpu.list([gpu.vec(v) for v in cols]))
        zo = orc.pfnGroup(orc.list([orc.vec(v) for v in cols]))
        assert groups_of(gpu, zg) == groups_of(orc, zo), len(cols)
    assert groups_of(gpu, gpu.pfnGroup(gpu.vec(ch, H_C))) == groups_of(orc, orc.pfnGroup(orc.vec(ch, H_C)))


@pytest.mark.parametrize("n,clustered", [(300_000, False), (20_000_000, True)])
def test_index_by_group_keys(gpu, n, clustered):
    """index(x, keys(group(x))) on an i32 column made in the arena returns the groups' key values
    (kept by the group, row path and run-compressed path); the same keys over another column of
    the same length are gathered as usual."""
    r = rng(n)
    if clustered:
        base = r.permutation(n // 4).astype(np.int32) - 5
        x = np.repeat(base, r.integers(1, 8, base.size))[:n]
    else:
        x = r.integers(-1000, 100000, n).astype(np.int32)
    other = r.integers(0, 7, x.size).astype(np.int32)
    xa = gpu.pfnPlus(gpu.vec(x), gpu.scalar(0, H_I))            # arena columns
    oa = gpu.pfnPlus(gpu.vec(other), gpu.scalar(0, H_I))
    keys = gpu.pfnKeys(gpu.pfnGroup(xa))
    k = gpu.read(keys)
    assert np.array_equal(gpu.read(gpu.pfnIndex(xa, keys)), x[k])
    assert np.array_equal(gpu.read(gpu.pfnIndex(oa, keys)), other[k])


@pytest.mark.parametrize("n", [5000, 70001])
def test_sorted_inputs(gpu, orc, n):
    """Keys already in the sort's order skip the radix scatter passes.  Inputs: sorted in the
    unsigned bit-pattern order the group uses; sorted as signed values (negatives then sit below
    the positives, a descent in that order); sorted but for one descent at a lane, wave or
    4096-row tile boundary; i64 keys; single-column orders of sorted and reverse-sorted columns."""
    r = rng(n + 77)
    v = r.integers(-10**6, 10**6, n).astype(np.int32)
    uns = np.sort(v.view(np.uint32)).view(np.int32)
    cases = [uns, np.sort(v)]
    uq = np.unique(v.view(np.uint32)).view(np.int32)              # strictly ascending (unsigned)
    for p in (64, 1024, 4096, 3 * 4096 + 1024 + 64):
        if p < uq.size:
            w = uq.copy()
            w[p - 1], w[p] = uq[p], uq[p - 1]                     # the one descent: row p
            cases.append(w)
    for x in cases:
        assert groups_of(gpu, gpu.pfnGroup(gpu.vec(x))) == groups_of(orc, orc.pfnGroup(orc.vec(x)))
    l64 = np.sort(r.integers(0, 2**40, n).astype(np.int64))
    assert groups_of(gpu, gpu.pfnGroup(gpu.vec(l64))) == groups_of(orc, orc.pfnGroup(orc.vec(l64)))
    for x in (np.sort(v), np.sort(v)[::-1].copy(), l64, np.sort(r.random(n))):
        for up in (1, 0):
            g = gpu.read(gpu.pfnOrderBy(gpu.vec(x), gpu.bools([up])))
            o = orc.read(orc.pfnOrderBy(orc.vec(x), orc.bools([up])))
            assert np.array_equal(g, o), (x.dtype, up)


@pytest.mark.parametrize("span", [9, 10, 11, 16, 19, 20, 21, 26, 30, 32])
@pytest.mark.parametrize("low", [0, 3])
def test_radix_digit_plans(gpu, orc, span, low):
    """i32 keys whose varying bits span `span` bits above `low` constant low bits: the radix sort
    covers them with a 10-bit first digit and then 10-bit or byte digits (2 passes up to 20 bits,
    3 up to 30), skipping constant digits.  Group and ascending/descending order against the
    oracle; 32-bit spans include negative keys (unsigned bit-pattern order)."""
    n = 300_000
    r = rng(span * 10 + low)
    hi = min(span + low, 32)
    v = r.integers(0, 1 << (hi - low), n, dtype=np.uint64) << np.uint64(low)
    x = (v & np.uint64(0xFFFFFFFF)).astype(np.uint32).view(np.int32)
    x[r.integers(0, n, n // 4)] = x[r.integers(0, n, n // 4)]              # repeated keys
    assert groups_of(gpu, gpu.pfnGroup(gpu.vec(x))) == groups_of(orc, orc.pfnGroup(orc.vec(x)))
    for up in (1, 0):
        g = gpu.read(gpu.pfnOrderBy(gpu.vec(x), gpu.bools([up])))
        o = orc.read(orc.pfnOrderBy(orc.vec(x), orc.bools([up])))
        assert np.array_equal(g, o), up


def test_group_long(gpu, orc):
    for x in (np.array([-1, 4294967295, -1, 7, 7], np.int64),
              rng(9).integers(-5, 5, 5000).astype(np.int64) * (1 << 33),
              rng(10).integers(0, 1000, 5000).astype(np.int64)):
        assert groups_of(gpu, gpu.pfnGroup(gpu.vec(x))) == groups_of(orc, orc.pfnGroup(orc.vec(x)))


@pytest.mark.parametrize("n", [1, 9, 1000, 300000])
def test_group_trie(gpu, orc, n):
    r = rng(n + 11)
    a = r.choice(np.frombuffer(b"ANR", np.uint8), n)
    b = r.choice(np.frombuffer(b"FO", np.uint8), n)
    zg = gpu.pfnGroup(gpu.list([gpu.vec(a, H_C), gpu.vec(b, H_C)]))
    zo = orc.pfnGroup(orc.list([orc.vec(a, H_C), orc.vec(b, H_C)]))
    assert groups_of(gpu, zg) == groups_of(orc, zo)


@pytest.mark.parametrize("n,lo,span", [(1, 7, 1), (4097, 0, 1), (5000, 1000, 256), (300001, 3, 255),
                                       (70000, 2**31 - 300, 256), (20000, -10, 200), (9000, 0, 257),
                                       (16385, 100, 7), (50000, 0, 13), (65, 5, 3), (33000, 9, 17)])
def test_group_int_small_range(gpu, orc, n, lo, span):
    """Key ranges of at most 256 values take the counting-sort group (packed per-thread counts up to
    16 bins, ballot ranks above); ranges that wrap in the unsigned order or hold 257 values take the
    radix group.  All match the oracle."""
    x = (lo + rng(n + span).integers(0, span, n)).astype(np.int32)
    assert groups_of(gpu, gpu.pfnGroup(gpu.vec(x))) == groups_of(orc, orc.pfnGroup(orc.vec(x)))


@pytest.mark.parametrize("n,na,nb", [(4096, 16, 16), (12345, 17, 16), (300001, 5, 50), (33, 256, 1), (70001, 1, 1),
                                     (16383, 3, 2), (40000, 4, 4), (100, 2, 2)])
def test_group_trie_bins(gpu, orc, n, na, nb):
    """Trie groups over na x nb byte pairs: up to 256 pairs in the counting-sort group, more in the
    radix group."""
    r = rng(n + na + nb)
    a = r.choice(r.choice(256, na, replace=False), n).astype(np.uint8)
    b = r.choice(r.choice(256, nb, replace=False), n).astype(np.uint8)
    zg = gpu.pfnGroup(gpu.list([gpu.vec(a, H_C), gpu.vec(b, H_C)]))
    zo = orc.pfnGroup(orc.list([orc.vec(a, H_C), orc.vec(b, H_C)]))
    assert groups_of(gpu, zg) == groups_of(orc, zo)


def test_group_small_sampled_bins_miss(gpu, orc):
    """From 2^20 rows on, the counting-sort group takes its bins from a strided sample of the
    input; values held only by rows outside the sample are caught by the count pass and the group
    is redone from the whole input (counting sort again, or radix when the bins then overflow)."""
    n = (1 << 20) + 5
    a = np.full(n, ord("A"), np.uint8)
    b = np.full(n, ord("F"), np.uint8)
    a[17] = ord("Z")
    b[n - 1] = ord("O")
    zg = gpu.pfnGroup(gpu.list([gpu.vec(a, H_C), gpu.vec(b, H_C)]))
    zo = orc.pfnGroup(orc.list([orc.vec(a, H_C), orc.vec(b, H_C)]))
    assert groups_of(gpu, zg) == groups_of(orc, zo)
    x = np.full(n, 5, np.int32)
    x[40] = 9
    assert groups_of(gpu, gpu.pfnGroup(gpu.vec(x))) == groups_of(orc, orc.pfnGroup(orc.vec(x)))
    x[40] = 100000
    x[n - 2] = 7
    assert groups_of(gpu, gpu.pfnGroup(gpu.vec(x))) == groups_of(orc, orc.pfnGroup(orc.vec(x)))
    r = rng(21)
    x = r.integers(0, 6, n).astype(np.int32)                 # every value in the sample: no redo
    assert groups_of(gpu, gpu.pfnGroup(gpu.vec(x))) == groups_of(orc, orc.pfnGroup(orc.vec(x)))


def test_group_list_and_float(gpu, orc):
    r = rng(12)
    a = r.integers(0, 5, 3000).astype(np.int64)
    b = r.integers(0, 4, 3000).astype(np.int32)
    zg = gpu.pfnGroup(gpu.list([gpu.vec(a), gpu.vec(b)]))
    zo = orc.pfnGroup(orc.list([orc.vec(a), orc.vec(b)]))
    assert groups_of(gpu, zg) == groups_of(orc, zo)
    d = r.permutation(2000).astype(np.float64) / 3          # distinct keys
    assert groups_of(gpu, gpu.pfnGroup(gpu.vec(d))) == groups_of(orc, orc.pfnGroup(orc.vec(d)))
    for n in (3000, 300_000):                                # keys with ties: the quicksort's order
        t = r.integers(0, 50, n) / 4.0
        assert groups_of(gpu, gpu.pfnGroup(gpu.vec(t))) == groups_of(orc, orc.pfnGroup(orc.vec(t)))
        dt = (19940101 + r.integers(0, 500, n)).astype(np.int32)
        assert groups_of(gpu, gpu.pfnGroup(gpu.vec(dt, H_D))) == groups_of(orc, orc.pfnGroup(orc.vec(dt, H_D)))
        zg = gpu.pfnGroup(gpu.list([gpu.vec(t), gpu.vec(dt, H_D)]))
        zo = orc.pfnGroup(orc.list([orc.vec(t), orc.vec(dt, H_D)]))
        assert groups_of(gpu, zg) == groups_of(orc, zo)


@pytest.mark.parametrize("n", [0, 1, 9, 10, 5000, 300001])
@pytest.mark.parametrize("up", [1, 0])
def test_order_single(gpu, orc, n, up):
    r = rng(n + 13 + up)
    for x, typ in ((r.integers(-50, 50, n).astype(np.int32), None), (r.integers(-2**40, 2**40, n).astype(np.int64), None),
                   (r.integers(0, 1000, n).astype(np.int64), None), (r.permutation(n).astype(np.float64) - n / 2, None),
                   (r.integers(65, 91, n).astype(np.uint8), H_C),
                   (r.integers(0, 40, n) / 8.0, None),                                  # f64 with ties
                   ((19940101 + r.integers(0, 300, n)).astype(np.int32), H_D)):         # dates with ties
        g = gpu.read(gpu.pfnOrderBy(gpu.vec(x, typ), gpu.bools([up])))
        o = orc.read(orc.pfnOrderBy(orc.vec(x, typ), orc.bools([up])))
        assert np.array_equal(g, o), (x.dtype, typ)


@pytest.mark.parametrize("n", [2, 1000, 1025, 30000, 131072, 131073, 200000])
def test_order_lists(gpu, orc, n):
    r = rng(n + 14)
    a = r.integers(0, 20, n).astype(np.int32)
    b = r.integers(-5, 5, n).astype(np.int32)
    for fl in ([1, 1], [0, 0]):                                  # packed case 1
        g = gpu.read(gpu.pfnOrderBy(gpu.list([gpu.vec(a), gpu.vec(b)]), gpu.bools(fl)))
        o = orc.read(orc.pfnOrderBy(orc.list([orc.vec(a), orc.vec(b)]), orc.bools(fl)))
        assert np.array_equal(g, o)
    rev = r.permutation(n).astype(np.float64) / 7              # revenue desc, day asc (q3.hir:73)
    day = r.integers(0, 2400, n).astype(np.int32)
    g = gpu.read(gpu.pfnOrderBy(gpu.list([gpu.vec(rev), gpu.vec(day)]), gpu.bools([0, 1])))
    o = orc.read(orc.pfnOrderBy(orc.list([orc.vec(rev), orc.vec(day)]), orc.bools([0, 1])))
    assert np.array_equal(g, o)
    rev_t = np.round(r.integers(0, max(n // 4, 2), n) * 0.5, 1)   # tied revenues, tied dates
    dte = (19950101 + r.integers(0, 60, n)).astype(np.int32)
    for fl in ([0, 1], [1, 0]):
        g = gpu.read(gpu.pfnOrderBy(gpu.list([gpu.vec(rev_t), gpu.vec(dte, H_D)]), gpu.bools(fl)))
        o = orc.read(orc.pfnOrderBy(orc.list([orc.vec(rev_t), orc.vec(dte, H_D)]), orc.bools(fl)))
        assert np.array_equal(g, o), fl
    c1 = r.integers(0, 30, n).astype(np.int64)
    c2 = r.permutation(n).astype(np.int64)
    g = gpu.read(gpu.pfnOrderBy(gpu.list([gpu.vec(c1), gpu.vec(c2)]), gpu.bools([1, 0])))
    o = orc.read(orc.pfnOrderBy(orc.list([orc.vec(c1), orc.vec(c2)]), orc.bools([1, 0])))
    assert np.array_equal(g, o)
    # full-row ties on integer columns (the reference merge-sorts those: rows in ascending
    # row order) and a constant column
    t1 = r.integers(0, 3, n).astype(np.int64) - 1
    t2 = np.full(n, 7, np.int32)
    t3 = r.integers(-2, 2, n).astype(np.int32)
    cols = lambda B: B.list([B.vec(t1), B.vec(t2), B.vec(t3)])
    for fl in ([1, 0, 1], [0, 1, 0]):
        g = gpu.read(gpu.pfnOrderBy(cols(gpu), gpu.bools(fl)))
        o = orc.read(orc.pfnOrderBy(cols(orc), orc.bools(fl)))
        assert np.array_equal(g, o), fl


@pytest.mark.parametrize("n", [500, 3000, 60000])
def test_order_lists_sorted_runs(gpu, orc, n):
    """The reference reverses a run in place when its column is already non-decreasing and the
    flag is descending (lib_order_by_sorted), which puts rows tied on every column in
    descending row order, and can flip them back one level down.  Integer columns (merge-sorted
    in the reference, so the tie order is defined): a sorted first column, columns sorted
    inside the runs of the previous one, constant columns, both flags."""
    r = rng(n + 15)
    rows = np.arange(n)
    cases = [
        [np.sort(r.integers(0, 10, n)).astype(np.int64), r.integers(0, 3, n).astype(np.int32), np.zeros(n, np.int32)],
        [r.integers(0, 4, n).astype(np.int32), (rows // 3).astype(np.int64), np.full(n, 5, np.int64)],
        [np.sort(r.integers(0, 6, n)).astype(np.int32), (-(rows // 4)).astype(np.int64), (rows % 2).astype(np.int32)],
        [r.integers(0, 3, n).astype(np.int64), r.integers(0, 2, n).astype(np.int32), (rows // 7).astype(np.int32)],
        # many first-column values (small buckets of the bucket-rank path), full ties on the rest
        [r.integers(0, max(n // 3, 1), n).astype(np.int64), r.integers(0, 2, n).astype(np.int32), np.zeros(n, np.int32)],
    ]
    for cols in cases:
        for fl in ([0, 0, 0], [0, 1, 0], [1, 0, 0], [0, 0, 1], [1, 1, 0]):
            g = gpu.read(gpu.pfnOrderBy(gpu.list([gpu.vec(c) for c in cols]), gpu.bools(fl)))
            o = orc.read(orc.pfnOrderBy(orc.list([orc.vec(c) for c in cols]), orc.bools(fl)))
            assert np.array_equal(g, o), (fl, [c.dtype for c in cols])


@pytest.mark.parametrize("n,sparse", [(300_000, False), (300_000, True), (2_000_000, False)])
def test_repeated_build_keys(gpu, orc, n, sparse):
    """Build sides whose keys repeat in many rows -- a constant column, a two-valued one -- for
    index_of (first occurrence), member and join, on the dense and the hashed tables."""
    r = rng(n + (7 if sparse else 0))
    scale = 100_000_007 if sparse else 1
    for x in (np.full(n, 5 * scale, np.int32), (r.integers(0, 2, n) * scale).astype(np.int32)):
        y = (r.integers(0, 4, n // 2) * scale).astype(np.int32)
        assert np.array_equal(gpu.read(gpu.pfnIndexOf(gpu.vec(x), gpu.vec(y))),
                              orc.read(orc.pfnIndexOf(orc.vec(x), orc.vec(y))))
        assert np.array_equal(gpu.read(gpu.pfnMember(gpu.vec(y), gpu.vec(x))),
                              orc.read(orc.pfnMember(orc.vec(y), orc.vec(x))))
    xb = np.concatenate([np.full(50, 9 * scale, np.int32), r.integers(0, 100, 2000).astype(np.int32) * scale])
    yb = (r.integers(0, 100, 3000) * scale).astype(np.int32)
    g = gpu.read_list(gpu.pfnJoinIndex(gpu.vec(xb), gpu.vec(yb), gpu.sym("eq")))
    o = orc.read_list(orc.pfnJoinIndex(orc.vec(xb), orc.vec(yb), orc.sym("eq")))
    assert np.array_equal(g[0], o[0]) and np.array_equal(g[1], o[1])


@pytest.mark.parametrize("nx,ny,card", [(40, 300, 30), (300, 40, 30), (100, 100, 10), (5, 50, 8),
                                        (50, 5, 8), (100000, 400000, 150000), (0, 10, 3), (20, 0, 3)])
def test_join(gpu, orc, nx, ny, card):
    r = rng(nx + ny)
    x = r.integers(0, card, nx).astype(np.int32)
    y = r.integers(0, card, ny).astype(np.int32)
    g = gpu.read_list(gpu.pfnJoinIndex(gpu.vec(x), gpu.vec(y), gpu.sym("eq")))
    o = orc.read_list(orc.pfnJoinIndex(orc.vec(x), orc.vec(y), orc.sym("eq")))
    assert np.array_equal(g[0], o[0]) and np.array_equal(g[1], o[1])


@pytest.mark.parametrize("nb,ny,swap", [(10, 40, False), (5000, 20000, False), (20000, 200000, True), (100000, 800000, False)])
def test_join_unique_build(gpu, orc, nb, ny, swap):
    """Unique build keys (the packed-table path), including -1 and keys that never match; the
    build side is x, or y after the swap."""
    r = rng(nb + ny)
    keys = (r.permutation(4 * nb) - 1).astype(np.int32)[:nb]
    probe = r.integers(-2, 4 * nb, ny).astype(np.int32)
    x, y = (probe, keys) if swap else (keys, probe)
    g = gpu.read_list(gpu.pfnJoinIndex(gpu.vec(x), gpu.vec(y), gpu.sym("eq")))
    o = orc.read_list(orc.pfnJoinIndex(orc.vec(x), orc.vec(y), orc.sym("eq")))
    assert len(o[0]) > 0
    assert np.array_equal(g[0], o[0]) and np.array_equal(g[1], o[1])


@pytest.mark.parametrize("nx,ny", [(0, 5), (1, 10), (300, 500), (100000, 300000)])
def test_index_of_member(gpu, orc, nx, ny):
    r = rng(nx + ny + 1)
    x = r.integers(0, max(nx // 2, 1), nx).astype(np.int32)
    y = r.integers(0, max(nx, 1), ny).astype(np.int32)
    for xx in (x, np.unique(x)):
        g = gpu.read(gpu.pfnIndexOf(gpu.vec(xx), gpu.vec(y)))
        o = orc.read(orc.pfnIndexOf(orc.vec(xx), orc.vec(y)))
        assert np.array_equal(g, o)
    xf = x.astype(np.float64) / 4
    yf = y.astype(np.float64) / 4
    assert np.array_equal(gpu.read(gpu.pfnIndexOf(gpu.vec(xf), gpu.vec(yf))),
                          orc.read(orc.pfnIndexOf(orc.vec(xf), orc.vec(yf))))
    for s in (x, x[:2], x[:1]):
        assert np.array_equal(gpu.read(gpu.pfnMember(gpu.vec(y), gpu.vec(s))),
                              orc.read(orc.pfnMember(orc.vec(y), orc.vec(s))))


@pytest.mark.parametrize("nb,spread", [(5000, 40), (5000, 400), (100003, 40)])
def test_guarded_tables(gpu, orc, nb, spread):
    """Build keys spanning 8..512 slots per key take the direct table behind a presence bitmap:
    index_of (unique and repeated keys) and join (unique build; a repeated build falls to the
    sort path), probed in ascending order as group keys come and in random order, with keys
    absent inside the range and keys outside it."""
    r = rng(nb + spread)
    lo = -7
    span = spread * nb
    u = np.unique(r.integers(1, span - 1, nb + nb // 4))[: nb - 2]
    keys = np.concatenate([[lo, lo + span - 1], lo + u]).astype(np.int32)
    r.shuffle(keys)
    absent = np.setdiff1d(np.arange(lo, lo + span, max(1, span // 5000)), keys)[:3000]
    probe = np.concatenate([r.choice(keys, 3 * nb), absent, [lo - 1, lo - 5, lo + span, 2**31 - 1]]).astype(np.int32)
    for y in (np.sort(probe), r.permutation(probe)):
        for x in (keys, np.concatenate([keys, keys[: nb // 3]])):
            assert np.array_equal(gpu.read(gpu.pfnIndexOf(gpu.vec(x), gpu.vec(y))),
                                  orc.read(orc.pfnIndexOf(orc.vec(x), orc.vec(y))))
            g = gpu.read_list(gpu.pfnJoinIndex(gpu.vec(x), gpu.vec(y), gpu.sym("eq")))
            o = orc.read_list(orc.pfnJoinIndex(orc.vec(x), orc.vec(y), orc.sym("eq")))
            assert len(o[0]) > 0
            assert np.array_equal(g[0], o[0]) and np.array_equal(g[1], o[1])


@pytest.mark.parametrize("nb,ny", [(5000, 20000), (60000, 100000)])
def test_hash_sparse_keys(gpu, orc, nb, ny):
    """Build keys spread over the whole i32 range take the packed hash table (dense key ranges
    take the direct table): index_of with repeated keys, member, and a unique-build join."""
    r = rng(nb + ny + 3)
    keys = r.integers(-2**31, 2**31 - 1, nb).astype(np.int32)
    rep = np.concatenate([keys, keys[: nb // 3]])                      # first occurrences matter
    probe = np.where(r.random(ny) < 0.5, r.choice(keys, ny), r.integers(-2**31, 2**31 - 1, ny)).astype(np.int32)
    assert np.array_equal(gpu.read(gpu.pfnIndexOf(gpu.vec(rep), gpu.vec(probe))),
                          orc.read(orc.pfnIndexOf(orc.vec(rep), orc.vec(probe))))
    assert np.array_equal(gpu.read(gpu.pfnMember(gpu.vec(probe), gpu.vec(rep))),
                          orc.read(orc.pfnMember(orc.vec(probe), orc.vec(rep))))
    u = np.unique(keys)
    g = gpu.read_list(gpu.pfnJoinIndex(gpu.vec(u), gpu.vec(probe), gpu.sym("eq")))
    o = orc.read_list(orc.pfnJoinIndex(orc.vec(u), orc.vec(probe), orc.sym("eq")))
    assert len(o[0]) > 0
    assert np.array_equal(g[0], o[0]) and np.array_equal(g[1], o[1])


def test_dense_keys_repeated_build_join(gpu, orc):
    """A dense build side with repeated keys: the direct table flags the repeat and the join takes
    its sort-based path."""
    r = rng(17)
    x = r.integers(100, 3000, 5000).astype(np.int32)
    y = r.integers(50, 3100, 20000).astype(np.int32)
    g = gpu.read_list(gpu.pfnJoinIndex(gpu.vec(x), gpu.vec(y), gpu.sym("eq")))
    o = orc.read_list(orc.pfnJoinIndex(orc.vec(x), orc.vec(y), orc.sym("eq")))
    assert np.array_equal(g[0], o[0]) and np.array_equal(g[1], o[1])


def test_index_of_bool_and_char(gpu, orc):
    for x, y in ((np.array([1, 1, 0], np.uint8), np.array([0, 1], np.uint8)),
                 (np.array([1, 1], np.uint8), np.array([0, 1], np.uint8))):
        assert np.array_equal(gpu.read(gpu.pfnIndexOf(gpu.vec(x, H_B), gpu.vec(y, H_B))),
                              orc.read(orc.pfnIndexOf(orc.vec(x, H_B), orc.vec(y, H_B))))
    cx = np.frombuffer(b"hello world", np.uint8)
    cy = np.frombuffer(b"lowered!", np.uint8)
    assert np.array_equal(gpu.read(gpu.pfnIndexOf(gpu.vec(cx, H_C), gpu.vec(cy, H_C))),
                          orc.read(orc.pfnIndexOf(orc.vec(cx, H_C), orc.vec(cy, H_C))))


@pytest.mark.parametrize("n,card", [(1000, 4), (200000, 6), (200000, 50000), (3_000_000, 4)])
def test_grouped_aggregates(gpu, orc, n, card):
    r = rng(n + card + 2)
    k = r.integers(0, card, n).astype(np.int32)
    v = np.round(r.uniform(0, 1e5, n), 2)
    q = r.integers(1, 50, n).astype(np.int32)
    res = {}
    for name, B in (("g", gpu), ("o", orc)):
        grp = B.pfnGroup(B.vec(k))
        vals = B.pfnValues(grp)
        out = {}
        for col, arr in (("v", v), ("q", q)):
            x = B.vec(arr)
            for fn in ("pfnSum", "pfnAvg", "pfnLen", "pfnMin", "pfnMax"):
                out[col + fn] = B.read(B.pfnRaze(B.each(B.each_right(x, vals, "pfnIndex"), fn)))
        res[name] = out
    for key in res["g"]:
        g, o = res["g"][key], res["o"][key]
        if g.dtype == np.float64:
            assert np.allclose(g, o, rtol=RTOL, atol=0), key
        else:
            assert np.array_equal(g, o), key
    # the FP templates on the device agree with the interpreter form
    grp = gpu.pfnGroup(gpu.vec(k))
    vals = gpu.pfnValues(grp)
    for kind, fn in ((0, "pfnSum"), (1, "pfnAvg"), (2, "pfnMin"), (3, "pfnMax")):
        fp = gpu.read(gpu.group_reduce(kind, gpu.vec(v), vals))
        assert np.allclose(fp, res["g"]["v" + fn], rtol=RTOL, atol=0)


def test_group_reduce_row_bounds(gpu):
    """A group's row lists are known to index the grouped column; reducing a shorter column
    through them still checks every row and fails with HPA_E_INDEX (code 3)."""
    k = (np.arange(100_000) % 7).astype(np.int32)
    vals = gpu.pfnValues(gpu.pfnGroup(gpu.vec(k)))
    v = np.round(rng(31).uniform(0, 100, k.size), 2)
    full = gpu.read(gpu.group_reduce(0, gpu.vec(v), vals))
    assert np.allclose(full, [math.fsum(v[k == g]) for g in range(7)], rtol=1e-12, atol=0)
    with pytest.raises(HorseError) as ei:
        gpu.group_reduce(0, gpu.vec(v[: k.size // 2]), vals)
    assert ei.value.code == 3


def test_group_offsets_after_rewind(gpu):
    """A group keeps its flat-list offsets for the reductions over it; after the arena is rewound
    a new group with as many groups but other sizes lands at the same addresses and must be
    reduced over its own offsets."""
    r = rng(41)
    keys = (np.arange(300) * 1000 - 7).astype(np.int32)       # a range over 256: the radix path
    m = gpu.mark()
    s1 = r.integers(1, 60, 300)
    for sizes in (s1, r.permutation(s1)):                   # same row count: same addresses
        k = r.permutation(np.repeat(keys, sizes)).astype(np.int32)
        v = np.round(r.uniform(0, 100, k.size), 2)
        vals = gpu.pfnValues(gpu.pfnGroup(gpu.vec(k)))
        got = gpu.read(gpu.group_reduce(0, gpu.vec(v), vals))
        want = [math.fsum(v[k == key]) for key in np.sort(keys.view(np.uint32)).view(np.int32)]
        assert np.allclose(got, want, rtol=1e-12, atol=0)
        gpu.release(m)


def test_arena_rewind_and_contexts(gpu):
    """Heap offset rewinds the arena; a context whose arena is too small returns HPA_E_MEMORY
    (the reference exits on a full heap) and leaves the default context usable."""
    lib = gpu.lib
    x = gpu.vec(np.arange(200_000, dtype=np.int64))   # before the mark: rewinding frees later nodes
    m = gpu.mark()
    gpu.pfnPlus(x, x)
    assert lib.dll.hpaGetHeapOffset() > m
    gpu.release(m)
    assert lib.dll.hpaGetHeapOffset() == m
    ctx = lib.dll.hpaCtxCreate(1 << 20)
    assert ctx
    lib.dll.hpaCtxSetCurrent(ctx)
    try:
        with pytest.raises(HorseError) as ei:
            gpu.pfnPlus(x, x)                      # a 1.6 MB output: more than the 1 MiB arena
        assert ei.value.code == 100
        small = gpu.pfnPlus(gpu.host.vec(np.arange(10, dtype=np.int64)), gpu.host.vec(np.arange(10, dtype=np.int64)))
        assert np.array_equal(gpu.read(small), 2 * np.arange(10))
    finally:
        lib.dll.hpaCtxSetCurrent(None)
        lib.dll.hpaCtxDestroy(ctx)
    assert np.array_equal(gpu.read(gpu.pfnPlus(x, x)), 2 * np.arange(200_000))
    gpu.release(m)


def test_where_one_match_leaves_no_index_bound(gpu):
    """pfnWhere with at most one match returns an inline scalar and rewinds the arena; the row
    bound of the freed index buffer must not pass to the next vector allocated at its address,
    so pfnIndex through that vector still checks its rows (HPA_E_INDEX, code 3)."""
    x = gpu.vec(np.arange(500, dtype=np.int64))
    src = gpu.vec(np.array([0, 1, 1000], np.int64))       # row 1000 is out of range for x
    zero = gpu.scalar(0, H_L)
    m = gpu.mark()
    mask2 = np.zeros(64, np.uint8)
    mask2[[5, 9]] = 1
    addr = gpu.pfnWhere(gpu.vec(mask2)).contents.g        # where the index buffer lands
    gpu.release(m)
    mask1 = np.zeros(64, np.uint8)
    mask1[5] = 1
    assert gpu.read(gpu.pfnWhere(gpu.vec(mask1))).tolist() == [5]
    t = None
    for _ in range(4096):                                  # 256-byte steps up to that address
        t = gpu.pfnPlus(src, zero)
        if t.contents.g >= addr:
            break
    assert t.contents.g == addr
    with pytest.raises(HorseError) as ei:
        gpu.pfnIndex(x, t)
    assert ei.value.code == 3
    gpu.release(m)


def test_tohost_copies_and_rewinds(gpu):
    """hpaToHost copies made before a heap mark survive a rewind to it, copies made after it are
    released, and a resident payload copied to the host keeps its handle for hpaFreeResident."""
    dll = gpu.lib.dll
    one = gpu.scalar(1, H_L)
    a = gpu.pfnPlus(gpu.vec(np.arange(1000, dtype=np.int64)), one)
    assert dll.hpaToHost(a) == 0
    m = gpu.mark()
    b = gpu.pfnPlus(a, one)
    assert dll.hpaToHost(b) == 0
    assert np.array_equal(gpu.read(b), np.arange(1000) + 2)
    gpu.release(m)
    assert np.array_equal(gpu.read(a), np.arange(1000) + 1)
    assert np.array_equal(gpu.read(gpu.pfnPlus(a, one)), np.arange(1000) + 2)
    gpu.release(m)
    r = gpu.vec(np.arange(10, dtype=np.int64) * 3)
    assert dll.hpaToHost(r) == 0
    assert np.array_equal(gpu.read(r), np.arange(10) * 3)
    assert dll.hpaFreeResident(r) == 0
