"""Multi-column pfnJoinIndex on the GPU against the oracle (itself pinned by a nested-loop brute
force in tests/test_oracle.py) and against the brute force directly on small inputs."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def brute(xs, ys, ops):
    f = {"eq": np.equal, "neq": np.not_equal, "lt": np.less, "leq": np.less_equal,
         "gt": np.greater, "geq": np.greater_equal}
    ok = np.ones((len(xs[0]), len(ys[0])), bool)
    for a, b, op in zip(xs, ys, ops):
        t = np.result_type(a, b)
        ok &= f[op](a.astype(t)[:, None], b.astype(t)[None, :])
    i, j = np.nonzero(ok)
    o = np.lexsort((i, j))
    return i[o], j[o]


def columns(r, ops, nx, ny, card):
    xs, ys = [], []
    first = ops.index("eq")
    for q, op in enumerate(ops):
        if q == first:
            xs.append(r.integers(0, card, nx).astype(np.int32)); ys.append(r.integers(0, card, ny).astype(np.int32))
        elif q % 2:
            xs.append(np.round(r.uniform(0, 4, nx), 1)); ys.append(r.integers(0, 4, ny).astype(np.int64))
        else:
            xs.append(r.integers(0, 4, nx).astype(np.int32)); ys.append(r.integers(0, 4, ny).astype(np.int16))
    return xs, ys


def join(B, xs, ys, ops):
    f = B.vec(np.array([B.symbol(o) for o in ops], np.int64), 9)
    return B.read_list(B.pfnJoinIndex(B.list([B.vec(a) for a in xs]), B.list([B.vec(b) for b in ys]), f))


OPS = [("eq", "lt"), ("lt", "eq", "neq"), ("eq", "eq", "geq"), ("gt", "leq", "eq")]


@pytest.mark.parametrize("ops", OPS)
def test_join_multi_small(gpu, ops):
    r = np.random.default_rng(len(ops) * 11)
    for nx, ny in ((40, 70), (300, 25), (5, 9), (9, 1)):
        xs, ys = columns(r, ops, nx, ny, 12)
        got = join(gpu, xs, ys, ops)
        i, j = brute(xs, ys, ops)
        assert got[0].tolist() == i.tolist() and got[1].tolist() == j.tolist(), (ops, nx, ny)


@pytest.mark.parametrize("ops", OPS)
@pytest.mark.parametrize("nx,ny,card", [(100_000, 300_000, 60_000), (300_000, 100_000, 1 << 30), (50_000, 50_000, 10)])
def test_join_multi_vs_oracle(gpu, orc, ops, nx, ny, card):
    r = np.random.default_rng(nx + ny + len(ops))
    xs, ys = columns(r, ops, nx, ny, card)
    if card == 10:                       # a few hundred thousand pairs per key: keep it small
        xs = [a[:2000] for a in xs]
        ys = [b[:3000] for b in ys]
    g, o = join(gpu, xs, ys, ops), join(orc, xs, ys, ops)
    assert g[0].tolist() == o[0].tolist() and g[1].tolist() == o[1].tolist(), (ops, nx, ny)


def test_join_multi_errors(gpu, orc):
    a = np.arange(10, dtype=np.int32)
    for B in (gpu, orc):
        f2 = B.vec(np.array([B.symbol("eq"), B.symbol("lt")], np.int64), 9)
        f3 = B.vec(np.array([B.symbol("eq")] * 3, np.int64), 9)
        with pytest.raises(Exception) as ei:          # list lengths differ from the operator count
            B.pfnJoinIndex(B.list([B.vec(a), B.vec(a)]), B.list([B.vec(a), B.vec(a)]), f3)
        assert ei.value.code == 5
        fl = B.vec(np.array([B.symbol("eq"), B.symbol("lt")], np.int64), 9)
        with pytest.raises(Exception) as ei:          # the first eq column is not i32
            B.pfnJoinIndex(B.list([B.vec(a.astype(np.float64)), B.vec(a)]), B.list([B.vec(a.astype(np.float64)), B.vec(a)]), fl)
        assert ei.value.code == 99
        del f2


def test_join_single_pair_after_swap(gpu, orc):
    x = np.arange(100, 140, dtype=np.int32)
    y = np.array([117], np.int32)
    for B in (gpu, orc):
        z = B.read_list(B.pfnJoinIndex(B.vec(x), B.vec(y), B.sym("eq")))
        assert z[0].tolist() == [17] and z[1].tolist() == [0]


def test_join_wide_sparse_range(gpu, orc):
    """Unique build keys spread over a range of about 5e8 values (over 2^28, at most 512 slots a
    key: the direct table behind a presence bitmap), probes half hitting: pairs equal to the
    oracle's, and index_of and member equal too."""
    r = np.random.default_rng(77)
    nb = 1_200_000
    build = np.sort(r.choice(500_000_000, nb, replace=False)).astype(np.int32)
    r.shuffle(build)
    probe = np.where(r.random(4_000_000) < 0.5, build[r.integers(0, nb, 4_000_000)],
                     r.integers(0, 500_000_000, 4_000_000)).astype(np.int32)
    eq = gpu.sym("eq")
    g = gpu.read_list(gpu.pfnJoinIndex(gpu.vec(build), gpu.vec(probe), eq))
    o = orc.read_list(orc.pfnJoinIndex(orc.vec(build), orc.vec(probe), orc.sym("eq")))
    assert np.array_equal(g[0], o[0]) and np.array_equal(g[1], o[1])
    gi = gpu.read(gpu.pfnIndexOf(gpu.vec(build), gpu.vec(probe)))
    assert np.array_equal(gi, orc.read(orc.pfnIndexOf(orc.vec(build), orc.vec(probe))))
    gm = gpu.read(gpu.pfnMember(gpu.vec(probe), gpu.vec(build)))
    assert np.array_equal(gm, np.isin(probe, build).astype(np.uint8))
