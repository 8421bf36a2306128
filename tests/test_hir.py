"""The HorseIR front of the fusion (horsepower_amd/hir.py) on the CPU: parsing the emitted
statement form, the fusion plan it derives, and the interpreter's one-primitive-per-statement run
on the oracle against numpy.  The fused runs are GPU tests (tests/test_gpu_hir.py)."""
import numpy as np
import pytest

from horsepower_amd import hir as H

Q6_HIR = """
module demo {
  import Builtin.*;
  def main() : table {
    ship:i32 = check_cast(@column_value(li, `l_shipdate:sym), i32);
    disc:f64 = @column_value(li, `l_discount:sym);
    qty:i32  = @column_value(li, `l_quantity:sym);
    price:f64 = @column_value(li, `l_extendedprice:sym);
    lo:bool = @geq(ship, 730:i32);
    hi:bool = @lt(ship, 1095:i32);
    inyear:bool = @and(lo, hi);
    d0:bool = @geq(disc, 0.05:f64);
    d1:bool = @leq(disc, 0.07:f64);
    band:bool = @and(d0, d1);
    few:bool = @lt(qty, 24:i32);
    m0:bool = @and(inyear, band);
    keep:bool = @and(m0, few);      // used twice: materialised once
    p:f64 = @compress(keep, price);
    d:f64 = @compress(keep, disc);
    rev:f64 = @mul(p, d);
    total:f64 = @sum(rev);
    return total;
  }
}
"""


def test_parse_forms():
    st, ret = H.parse(Q6_HIR)
    assert ret == ["total"] and len(st) == 17
    assert st[0].fn == "column_value" and st[0].args[1].vals == ["l_shipdate"]
    assert st[4].fn == "geq" and st[4].args[1].vals == ["730"] and st[4].args[1].typ == "i32"
    st2, _ = H.parse('a = (1,2):i32; b = @sub_string(s, a); c = @like(b, "1_":str); d = @each(@sum, g); return c;')
    assert st2[0].fn == "literal" and st2[0].args[0].vals == ["1", "2"] and st2[0].args[0].vector
    assert st2[2].args[1].vals == [b"1_"] and isinstance(st2[3].args[0], H.Fn)
    with pytest.raises(SyntaxError):
        H.parse("a = @lt(x, ;")


def test_plan_q6_shape():
    p = H.plan(*H.parse(Q6_HIR))
    kinds = p.kinds()
    # four column reads, one two-column compress with the mask's tree as its condition (no
    # mask built: the mask has no other use), one fused sum
    assert kinds == ["stmt"] * 4 + ["compress", "reduce"], kinds
    c = p.ops[4]
    assert c[1] == "keep" and c[4][0] == "op" and c[4][1] == "and"
    assert c[2] == [("p", c[2][0][1]), ("d", c[2][1][1])]
    assert H.plan(*H.parse(Q6_HIR), fuse=False).kinds() == ["stmt"] * 17


def test_plan_fp_and_fa_reduce():
    text = """
      v = @mul(price, disc);
      c = @compress(m, v);
      s = @sum(c);
      x = @each_right(@index, price, groups);
      y = @each(@len, x);
      z = @raze(y);
      return s, z;"""
    p = H.plan(*H.parse(text))
    assert p.kinds() == ["reduce", "group_reduce"]
    assert p.ops[0][3] == ("var", "m") and p.ops[0][4][1] == "mul"
    assert p.ops[1][2] == 4


def test_plain_run_on_oracle(orc):
    r = np.random.default_rng(6)
    n = 200_000
    cols = dict(ship=r.integers(0, 2500, n).astype(np.int32), disc=r.integers(0, 11, n) / 100.0,
                qty=r.integers(1, 51, n).astype(np.int32), price=r.integers(90000, 10500000, n) / 100.0)
    text = Q6_HIR.replace("check_cast(@column_value(li, `l_shipdate:sym), i32)", "ship0").replace(
        "@column_value(li, `l_discount:sym)", "disc0").replace("@column_value(li, `l_quantity:sym)", "qty0").replace(
        "@column_value(li, `l_extendedprice:sym)", "price0")
    env = {k + "0": orc.vec(v) for k, v in cols.items()}
    (tot,), p = H.run(text, orc, env, fuse=False)
    sel = (cols["ship"] >= 730) & (cols["ship"] < 1095) & (cols["disc"] >= 0.05) & (cols["disc"] <= 0.07) & (cols["qty"] < 24)
    want = np.sum(cols["price"][sel] * cols["disc"][sel])
    assert abs(orc.read_scalar(tot) - want) <= 1e-9 * abs(want)
