"""The C caller (tests/c_driver/driver.c, built by horsepower_amd/build.py) against the oracle:
q6.hir's statement sequence and its fused form, a two-column join and a group on repeated f64
keys, all through the C-ABI from C."""
import math
import os
import subprocess

import numpy as np
import pytest

from horsepower_amd import datagen as D
from horsepower_amd import queries as Q
from horsepower_amd.vnode import H_E

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "c_driver", "hpa_c_driver")


def test_c_driver_vs_oracle(orc, tmp_path):
    assert os.path.exists(DRIVER), "the C driver is built by horsepower_amd/build.py"
    r = np.random.default_rng(3)
    t = D.gen_q6(500_000)
    cols = {"ship_day.i32": t["ship_day"], "discount.f64": t["discount"], "quantity.i32": t["quantity"],
            "price.f64": t["price"],
            "jx_key.i32": r.integers(0, 5000, 20_000).astype(np.int32), "jx_val.f64": np.round(r.uniform(0, 9, 20_000), 1),
            "jy_key.i32": r.integers(0, 5000, 60_000).astype(np.int32), "jy_val.i64": r.integers(0, 9, 60_000).astype(np.int64),
            "gkey.f64": r.integers(0, 300, 100_000) / 8.0}
    for name, a in cols.items():
        np.ascontiguousarray(a).tofile(tmp_path / name)
    print("running", DRIVER, flush=True)
    p = subprocess.run([DRIVER, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    rd = lambda name, dt: np.fromfile(tmp_path / name, dtype=dt)
    o6 = Q.q6_plain(orc, Q.load_table(orc, t))["revenue"]
    sel = (t["ship_day"] >= D.Q6_LO) & (t["ship_day"] < D.Q6_HI) & (t["quantity"] < 24) & \
        (t["discount"] >= 0.05) & (t["discount"] <= 0.07)
    exact = math.fsum(t["price"][sel] * t["discount"][sel])
    for name in ("q6.f64", "q6_fused.f64"):
        g = float(rd(name, np.float64)[0])
        assert abs(g - o6) <= 1e-12 * abs(o6) and abs(g - exact) <= 1e-12 * abs(exact), name
    f = orc.vec(np.array([orc.symbol("eq"), orc.symbol("lt")], np.int64), 9)
    oj = orc.read_list(orc.pfnJoinIndex(orc.list([orc.vec(cols["jx_key.i32"]), orc.vec(cols["jx_val.f64"])]),
                                        orc.list([orc.vec(cols["jy_key.i32"]), orc.vec(cols["jy_val.i64"])]), f))
    assert rd("join0.i64", np.int64).tolist() == oj[0].tolist()
    assert rd("join1.i64", np.int64).tolist() == oj[1].tolist()
    k, c, fl = orc.read_groups(orc.pfnGroup(orc.vec(cols["gkey.f64"], H_E)))
    assert rd("gkeys.i64", np.int64).tolist() == k.tolist()
    assert rd("gcounts.i64", np.int64).tolist() == c.tolist()
    assert rd("grows.i64", np.int64).tolist() == fl.tolist()
