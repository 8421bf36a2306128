"""Orders and groups on comparison-typed keys with ties, GPU against the oracle.

The reference orders B/J/H/F/E/D (and the other non-I/L/C types) with an unstable first-pivot
Hoare quicksort (h_libs.cc:409-442); rows whose keys tie come out in the order its partitions
leave them, and groups on such keys take their keys and row lists from that order
(h_primitive.cc:1002-1057).  Permutations, group keys and row lists must match bit for bit.
The oracle's restatement of the partition is pinned by the hand-traced vectors of
tests/golden/kat_traced.json.
"""
import numpy as np
import pytest

from horsepower_amd.vnode import H_B, H_D, H_E, H_F, H_H, H_J, HorseError

pytestmark = pytest.mark.gpu

# around the LDS segment size (4096), the tiny-segment size (32) and the 64-row chunks
SIZES = [2, 3, 10, 33, 64, 65, 1000, 4096, 4097, 4098, 9000, 100_003, 300_000]


def rng(seed):
    return np.random.default_rng(seed)


def columns(n, seed):
    """Tie-bearing columns of every comparison-sorted type: (values, type)."""
    r = rng(seed)
    return [
        (r.integers(0, 11, n) / 100.0, H_E),                          # discount-like: 11 values
        (np.round(r.integers(0, max(n // 8, 2), n) / 7.0, 6), H_E),    # about 8 rows a value
        (r.integers(0, 3, n).astype(np.float32), H_F),
        (19920101 + r.integers(0, 2406, n).astype(np.int32), H_D),     # dates with repeats
        (r.integers(-4, 4, n).astype(np.int16), H_H),
        (r.integers(0, 2, n).astype(np.uint8), H_B),
        (r.integers(-100, 100, n).astype(np.int8), H_J),
        (np.where(r.random(n) < 0.5, -0.0, 0.0) + r.integers(0, 2, n), H_E),   # -0 and +0 tie
    ]


def groups_of(B, z):
    k, c, f = B.read_groups(z)
    return k.tolist(), c.tolist(), f.tolist()


@pytest.mark.parametrize("n", SIZES)
def test_order_vector_ties(gpu, orc, n):
    for i, (x, t) in enumerate(columns(n, n + 1)):
        for up in (1, 0):
            print("order", n, i, up, flush=True)
            g = gpu.read(gpu.pfnOrderBy(gpu.vec(x, t), gpu.bools([up])))
            o = orc.read(orc.pfnOrderBy(orc.vec(x, t), orc.bools([up])))
            assert np.array_equal(g, o), (n, i, up)


@pytest.mark.parametrize("n", SIZES)
def test_group_ties(gpu, orc, n):
    for i, (x, t) in enumerate(columns(n, n + 2)):
        print("group", n, i, flush=True)
        assert groups_of(gpu, gpu.pfnGroup(gpu.vec(x, t))) == groups_of(orc, orc.pfnGroup(orc.vec(x, t))), (n, i)


def test_order_sorted_and_reversed_ties(gpu, orc):
    """Inputs the first-pivot partition handles worst: sorted and reverse-sorted keys with ties
    (every pivot the smallest or the largest of its segment), and all-equal keys."""
    r = rng(5)
    n = 20_000
    base = np.sort(r.integers(0, 2000, n)) / 4.0
    for x in (base, base[::-1].copy(), np.full(n, 2.5), np.repeat([1.0, 2.0], n // 2)):
        for up in (1, 0):
            g = gpu.read(gpu.pfnOrderBy(gpu.vec(x, H_E), gpu.bools([up])))
            o = orc.read(orc.pfnOrderBy(orc.vec(x, H_E), orc.bools([up])))
            assert np.array_equal(g, o), up


@pytest.mark.parametrize("n", [5, 33, 1000, 4097, 60_000, 300_000])
def test_order_lists_float_ties(gpu, orc, n):
    """List orders with a comparison-sorted column and rows tied on every column: the reference's
    column-by-column passes (lib_quicksort_list) decide their order."""
    r = rng(n + 3)
    rev = np.round(r.integers(0, max(n // 20, 2), n) * 0.25, 2)      # revenue-like, repeated
    day = r.integers(0, 30, n).astype(np.int32)
    prio = r.integers(0, 3, n).astype(np.uint8)
    dte = (19950101 + r.integers(0, 40, n)).astype(np.int32)
    cases = [
        ([(rev, H_E), (day, None)], [[0, 1], [1, 1], [0, 0], [1, 0]]),
        ([(day, None), (rev, H_E)], [[1, 0], [0, 1]]),
        ([(dte, H_D), (prio, H_B), (day, None)], [[1, 1, 1], [0, 1, 0]]),
        ([(prio, H_B), (rev, H_E), (dte, H_D)], [[1, 0, 1], [0, 0, 0]]),
        ([(np.sort(rev), H_E), (prio, H_B)], [[1, 1], [0, 1], [0, 0]]),         # sorted first column
    ]
    for cols, flags in cases:
        for fl in flags:
            print("list", n, [c[1] for c in cols], fl, flush=True)
            g = gpu.read(gpu.pfnOrderBy(gpu.list([gpu.vec(v, t) for v, t in cols]), gpu.bools(fl)))
            o = orc.read(orc.pfnOrderBy(orc.list([orc.vec(v, t) for v, t in cols]), orc.bools(fl)))
            assert np.array_equal(g, o), fl


@pytest.mark.parametrize("n", [1000, 200_000])
def test_group_list_float_ties(gpu, orc, n):
    """pfnGroup of a list with a float column orders it first (h_primitive.cc:1017-1020)."""
    r = rng(n + 4)
    a = r.integers(0, 5, n) / 4.0
    b = r.integers(0, 4, n).astype(np.int32)
    d = (19950101 + r.integers(0, 9, n)).astype(np.int32)
    for cols in ([(a, H_E), (b, None)], [(b, None), (a, H_E)], [(d, H_D), (a, H_E)]):
        zg = gpu.pfnGroup(gpu.list([gpu.vec(v, t) for v, t in cols]))
        zo = orc.pfnGroup(orc.list([orc.vec(v, t) for v, t in cols]))
        assert groups_of(gpu, zg) == groups_of(orc, zo)


def test_group_discount_6m(gpu, orc):
    """A group on a 6M-row decimal column with 11 values (config 0's discount)."""
    x = rng(6).integers(0, 11, 6_000_000) / 100.0
    assert groups_of(gpu, gpu.pfnGroup(gpu.vec(x, H_E))) == groups_of(orc, orc.pfnGroup(orc.vec(x, H_E)))


def test_nan_keys_rejected(gpu, orc):
    """A NaN key sends lib_partition's scan outside its segment in the reference; both sides
    return E_DOMAIN (code 1) instead."""
    x = np.array([1.0, np.nan, 2.0, 1.0])
    for B in (gpu, orc):
        for call in (lambda: B.pfnOrderBy(B.vec(x, H_E), B.bools([1])),
                     lambda: B.pfnGroup(B.vec(x, H_E)),
                     lambda: B.pfnOrderBy(B.list([B.vec(x, H_E), B.vec(np.arange(4, dtype=np.int32))]), B.bools([1, 1]))):
            with pytest.raises(HorseError) as ei:
                call()
            assert ei.value.code == 1, B.name
