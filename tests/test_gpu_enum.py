"""Enumerations and foreign keys on the GPU against the oracle: pfnEnum at the join workload's
size (1.5M order keys, about 6M lineitem foreign keys), E_ENUM_INDEX, keys/values/fetch of an
enum, and the table flow pfnAddFKey -> pfnColumnValue -> values -> index that the compiled join
query runs (tests/hyper/online-hir/q3.hir:12-23)."""
import ctypes as C

import numpy as np
import pytest

from horsepower_amd import datagen as D
from horsepower_amd.vnode import H_A, H_G, H_I, H_L, H_Q, V, V0, HorseError

pytestmark = pytest.mark.gpu


class A0(C.Structure):                      # hpa_A0: rows, columns, foreign-key list
    _fields_ = [("row", C.c_int64), ("col", C.c_int64), ("fkey", C.c_void_p)]


def table(B, names, cols, keep):
    """A table node in the reference layout (h_system.h:17-31): V0[2] = (name symbols, list of
    columns), the union pointer -> A0."""
    syms = np.array([B.symbol(n) for n in names], np.int64)
    cells = (V0 * 2)()
    kv = B.host.vec(syms, H_Q)
    lst = B.list([B.vec(c) for c in cols])
    cells[0] = kv.contents
    cells[1] = lst.contents
    a = A0(len(cols[0]), len(cols), None)
    t = C.pointer(V0())
    t.contents.typ, t.contents.len = H_A, 2
    t.contents.g = C.addressof(cells)
    t.contents.p = C.addressof(a)
    keep += [cells, kv, lst, a, t, syms]
    return t


def register(B, name, t):
    dll = B.lib.dll
    if B.name == "gpu":
        dll.hpaRegisterTable.argtypes = [C.c_int64, V]
        assert dll.hpaRegisterTable(B.symbol(name), t) == 0
    else:
        dll.orc_register_table.argtypes = [C.c_int64, V]
        assert dll.orc_register_table(B.symbol(name), t) == 0


def test_enum_q3_size(gpu, orc):
    dm, f = D.gen_q3(1_500_000)
    eg = gpu.pfnEnum(gpu.vec(dm["key"]), gpu.vec(f["fk"]))
    eo = orc.pfnEnum(orc.vec(dm["key"]), orc.vec(f["fk"]))
    vg, vo = gpu.read(gpu.pfnValues(eg)), orc.read(orc.pfnValues(eo))
    assert vg.dtype == np.int64 and np.array_equal(vg, vo)
    assert np.array_equal(gpu.read(gpu.pfnKeys(eg)), dm["key"])
    assert np.array_equal(gpu.read(gpu.pfnFetch(eg)), f["fk"])
    # the foreign-key gather: a dimension column through the enum's values
    g = gpu.read(gpu.pfnIndex(gpu.vec(dm["order_day"]), gpu.pfnValues(eg)))
    assert np.array_equal(g, dm["order_day"][vo])


@pytest.mark.parametrize("n", [0, 1, 2, 1000])
def test_enum_small_and_missing(gpu, orc, n):
    r = np.random.default_rng(n + 5)
    keys = (r.permutation(max(n, 1) * 4) * 2).astype(np.int32)
    vals = keys[r.integers(0, keys.size, n)] if n else np.zeros(0, np.int32)
    for B in (gpu, orc):
        e = B.pfnEnum(B.vec(keys), B.vec(vals))
        assert e.contents.len == n
    g = gpu.read(gpu.pfnValues(gpu.pfnEnum(gpu.vec(keys), gpu.vec(vals))))
    o = orc.read(orc.pfnValues(orc.pfnEnum(orc.vec(keys), orc.vec(vals))))
    assert g.tolist() == o.tolist()
    if n:
        bad = vals.copy()
        bad[-1] = 1                                  # odd: not a key
        for B in (gpu, orc):
            with pytest.raises(HorseError) as ei:
                B.pfnEnum(B.vec(keys), B.vec(bad))
            assert ei.value.code == 10, B.name


def test_add_fkey_and_column_value(gpu, orc):
    """pfnAddFKey on registered tables, then the compiled join query's first statements:
    column_value of the foreign-key column is the enum, values() its index, index() through it."""
    dm, f = D.gen_q3(200_000)
    out = {}
    for B in (gpu, orc):
        keep = []
        orders = table(B, ["o_orderkey", "o_orderdate", "o_segment"], [dm["key"], dm["order_day"], dm["segment"]], keep)
        lineitem = table(B, ["l_orderkey", "l_shipdate", "l_price"], [f["fk"], f["ship_day"], f["price"]], keep)
        register(B, "orders", orders)
        register(B, "lineitem", lineitem)
        sym = lambda s: B.host.vec(np.array([B.symbol(s)], np.int64), H_Q)
        xk, yk = sym("o_orderkey"), sym("l_orderkey")
        keep += [xk, yk]
        B.lib.call("pfnAddFKey", sym("orders"), xk, sym("lineitem"), yk)
        t = B.pfnLoadTable(sym("lineitem"))
        c20 = B.pfnColumnValue(t, sym("l_orderkey"))          # the enum (q3.hir:16)
        assert c20.contents.typ == 19, B.name
        c21 = B.pfnColumnValue(t, sym("l_price"))
        seg = B.pfnColumnValue(B.pfnLoadTable(sym("orders")), sym("o_segment"))
        mask = B.pfnEq(seg, B.scalar(1, H_I))
        fk_mask = B.pfnIndex(mask, B.pfnValues(c20))          # segment of each lineitem's order
        rev = B.pfnCompress(fk_mask, c21)
        out[B.name] = (B.read(B.pfnValues(c20)), B.read(fk_mask), B.read(rev), keep)
        with pytest.raises(HorseError) as ei:
            B.pfnColumnValue(t, sym("no_such_column"))
        assert ei.value.code == 51
    for a, b in zip(out["gpu"][:3], out["oracle"][:3]):
        assert np.array_equal(a, b)
