"""HorseIR programs fused into the library's descriptors (horsepower_amd/hir.py) against the same
programs run one primitive per statement, on the GPU and on the oracle.  Integer and boolean
results bit for bit; element-wise float results bit for bit (both sides round each operation);
sums within 1e-12 of an exactly rounded sum.  Inputs keep float compares away from the
interpreter's epsilon, where the emitted (fused) code and the interpreter differ by design."""
import math

import numpy as np
import pytest

from horsepower_amd import hir as H
from horsepower_amd.vnode import H_C, H_E, H_I, H_L

pytestmark = pytest.mark.gpu
RTOL = 1e-12

Q6 = """
    lo = @geq(ship, 730:i32);  hi = @lt(ship, 1095:i32);  yr = @and(lo, hi);
    d0 = @geq(disc, 0.05:f64); d1 = @leq(disc, 0.07:f64); band = @and(d0, d1);
    few = @lt(qty, 24:i32);    m0 = @and(yr, band);        keep = @and(m0, few);
    v = @mul(price, disc);     c = @compress(keep, v);     total = @sum(c);
    return total;"""

Q1 = """
    m = @leq(ship, 2434:i32);
    f = @compress(m, flag);  st = @compress(m, status);  qt = @compress(m, qty);
    pr = @compress(m, price); ds = @compress(m, disc);   tx = @compress(m, tax);
    one = @minus(1:f64, ds);  dp = @mul(pr, one);         up = @plus(1:f64, tx);   ch = @mul(dp, up);
    k = @list(f, st);  g = @group(k);  gv = @values(g);  gk = @keys(g);
    x1 = @each_right(@index, qt, gv);  s1 = @each(@sum, x1);  sq = @raze(s1);
    x2 = @each_right(@index, dp, gv);  s2 = @each(@sum, x2);  sdp = @raze(s2);
    x3 = @each_right(@index, ch, gv);  s3 = @each(@sum, x3);  sch = @raze(s3);
    x4 = @each_right(@index, qt, gv);  s4 = @each(@len, x4);  cnt = @raze(s4);
    x5 = @each_right(@index, ds, gv);  s5 = @each(@avg, x5);  adisc = @raze(s5);
    return gk, sq, sdp, sch, cnt, adisc;"""


def cols(n, seed):
    r = np.random.default_rng(seed)
    return dict(ship=r.integers(0, 2500, n).astype(np.int32), disc=r.integers(0, 11, n) / 100.0,
                qty=r.integers(1, 51, n).astype(np.int32), price=r.integers(90000, 10500000, n) / 100.0,
                tax=r.integers(0, 9, n) / 100.0, flag=r.choice(np.frombuffer(b"ANR", np.uint8), n),
                status=r.choice(np.frombuffer(b"FO", np.uint8), n))


def bind(B, c):
    env = {}
    for k, v in c.items():
        env[k] = B.vec(v, H_C) if v.dtype == np.uint8 else B.vec(v)
    return env


@pytest.mark.parametrize("n", [5000, 6_000_000])
def test_q6_program(gpu, orc, n):
    c = cols(n, n)
    (tf,), pf = H.run(Q6, gpu, bind(gpu, c), fuse=True)
    assert pf.kinds() == ["reduce"], pf.kinds()              # one FA reduction, nothing else
    (tp,), _ = H.run(Q6, gpu, bind(gpu, c), fuse=False)
    (to,), _ = H.run(Q6, orc, bind(orc, c), fuse=False)
    sel = (c["ship"] >= 730) & (c["ship"] < 1095) & (c["disc"] >= 0.05) & (c["disc"] <= 0.07) & (c["qty"] < 24)
    exact = math.fsum(c["price"][sel] * c["disc"][sel])
    for v in (gpu.read_scalar(tf), gpu.read_scalar(tp)):
        assert abs(v - exact) <= RTOL * abs(exact)
    assert abs(orc.read_scalar(to) - exact) <= 1e-9 * abs(exact)


@pytest.mark.parametrize("n", [3000, 1_000_000])
def test_q1_program(gpu, orc, n):
    c = cols(n, n + 7)
    outs = {}
    for name, B, fuse in (("fused", gpu, True), ("plain", gpu, False), ("oracle", orc, False)):
        res, p = H.run(Q1, B, bind(B, c), fuse=fuse)
        if name == "fused":
            assert p.kinds().count("group_reduce") == 5 and "compress" in p.kinds() and "map" in p.kinds(), p.kinds()
        outs[name] = [B.read(z) for z in res]
    for a, b, o in zip(outs["fused"], outs["plain"], outs["oracle"]):
        if a.dtype.kind == "f":
            assert np.allclose(a, b, rtol=RTOL, atol=0) and np.allclose(a, o, rtol=1e-9, atol=0)
        else:
            assert np.array_equal(a, b) and np.array_equal(a, o)


OPS_NUM = ["plus", "minus", "mul"]
OPS_CMP = ["lt", "leq", "gt", "geq", "eq", "neq"]


def random_program(r, k, bound):
    """k random element-wise statements over i32 / i64 / f64 columns, each value kept below 2^30
    in magnitude (no intermediate overflows its type), ending in a returned boolean and a
    returned number.  bound: magnitude bound of each column."""
    lines, nums, bools = [], ["a", "b", "e"], []
    mag = dict(bound)
    lits = {"3:i32": 3, "2.5:f64": 2.5, "7:i64": 7}
    for j in range(k):
        t = f"t{j}"
        roll = r.random()
        if roll < 0.55:
            x, y = r.choice(nums), r.choice(nums + list(lits))
            op = r.choice(OPS_NUM)
            bx, by = mag[x], mag.get(y, lits.get(y))
            m = bx * by if op == "mul" else bx + by
            if m >= 2 ** 30:
                continue
            lines.append(f"{t} = @{op}({x}, {y});")
            nums.append(t)
            mag[t] = m
        elif roll < 0.85 or len(bools) < 2:
            x, y = r.choice(nums), r.choice(nums + ["10:i32", "0.5:f64"])
            lines.append(f"{t} = @{r.choice(OPS_CMP)}({x}, {y});")
            bools.append(t)
        else:
            x, y = r.choice(bools), r.choice(bools)
            lines.append(f"{t} = @{r.choice(['and', 'or'])}({x}, {y});")
            bools.append(t)
    if not bools:
        lines.append(f"tb = @lt({nums[-1]}, 5:i32);")
        bools.append("tb")
    lines.append(f"return {bools[-1]}, {nums[-1]};")
    return "\n".join(lines)


def test_random_elementwise_programs(gpu):
    r = np.random.default_rng(31)
    n = 100_003
    c = dict(a=r.integers(-20, 20, n).astype(np.int32), b=r.integers(-5, 5, n).astype(np.int64),
             e=r.integers(-400, 400, n) / 8.0)                    # eighths: exact, no epsilon ties
    fused_any = 0
    for k in range(40):
        text = random_program(r, int(r.integers(2, 9)), {"a": 20, "b": 5, "e": 50})
        (zb, zn), p = H.run(text, gpu, bind(gpu, c), fuse=True)
        (pb, pn), _ = H.run(text, gpu, bind(gpu, c), fuse=False)
        fused_any += p.kinds().count("map")
        for x, y in ((zb, pb), (zn, pn)):
            gx, gy = gpu.read(x), gpu.read(y)
            assert gx.dtype == gy.dtype and np.array_equal(gx, gy), text
    assert fused_any > 10


def test_division_stays_a_statement(gpu):
    """Division keeps pfnDiv's zero check: a tree over a division runs the division as a
    statement (fusing what is around it) and a zero divisor fails as the primitive does."""
    r = np.random.default_rng(2)
    n = 10_000
    c = dict(a=r.integers(1, 9, n).astype(np.int32), e=r.integers(1, 9, n) / 4.0)
    text = "t0 = @div(a, 2:i32); t1 = @plus(t0, a); t2 = @div(e, 2:f64); t3 = @mul(t2, e); return t1, t3;"
    (z1, z3), p = H.run(text, gpu, bind(gpu, c), fuse=True)
    assert p.kinds() == ["stmt"] * 4
    assert np.array_equal(gpu.read(z1), c["a"] // 2 + c["a"])
    assert np.array_equal(gpu.read(z3), (c["e"] / 2.0) * c["e"])
    c0 = dict(a=np.zeros(n, np.int32), e=c["e"])
    with pytest.raises(Exception) as ei:
        H.run("t0 = @div(e, a); t1 = @plus(t0, 1:i32); return t1;", gpu, bind(gpu, c0), fuse=True)
    assert getattr(ei.value, "code", None) == 4


def test_registered_table_program(gpu):
    """load_table / column_value statements feed the fused sections."""
    import ctypes as C
    from test_gpu_enum import table, register
    c = cols(200_000, 5)
    keep = []
    t = table(gpu, ["l_shipdate", "l_discount", "l_quantity", "l_extendedprice"],
              [c["ship"], c["disc"], c["qty"], c["price"]], keep)
    register(gpu, "hir_lineitem", t)
    text = """t0 = @load_table(`hir_lineitem:sym);
      ship = @column_value(t0, `l_shipdate:sym);  disc = @column_value(t0, `l_discount:sym);
      qty = @column_value(t0, `l_quantity:sym);   price = @column_value(t0, `l_extendedprice:sym);""" + Q6
    (tf,), p = H.run(text, gpu, {}, fuse=True)
    (tp,), _ = H.run(text, gpu, {}, fuse=False)
    assert "reduce" in p.kinds()
    assert abs(gpu.read_scalar(tf) - gpu.read_scalar(tp)) <= RTOL * abs(gpu.read_scalar(tp))
