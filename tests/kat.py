"""Applies a known-answer vector from tests/golden/kat.json to a primitive backend."""
import json
import os

import numpy as np

from horsepower_amd.vnode import H_B, H_E, H_F, H_H, H_I, H_J, H_L

_GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
# kat.json: examples taken from the reference's own files (tests/golden/make_golden.py);
# kat_traced.json: small inputs traced by hand through the cited reference code, trace in "note"
KAT = (json.load(open(os.path.join(_GOLDEN, "kat.json")))
       + json.load(open(os.path.join(_GOLDEN, "kat_traced.json"))))
TYP = {"bool": H_B, "i8": H_J, "i16": H_H, "i32": H_I, "i64": H_L, "f32": H_F, "f64": H_E}
NPT = {"bool": np.uint8, "i8": np.int8, "i16": np.int16, "i32": np.int32, "i64": np.int64,
       "f32": np.float32, "f64": np.float64}
SUPPORTED = {"compress", "where", "index", "member", "min", "len", "eq", "lt", "order",
             "order_list", "group", "index_of", "plus_chain", "each_sum_raze"}


def node(B, a):
    return B.vec(np.array(a["values"], dtype=NPT[a["type"]]), TYP[a["type"]])


def run(B, e):
    fn, A = e["fn"], e["args"]
    x = [node(B, a) for a in A]
    if fn == "compress":
        return B.read(B.pfnCompress(x[0], x[1]))
    if fn == "where":
        return B.read(B.pfnWhere(x[0]))
    if fn == "index":
        return B.read(B.pfnIndex(x[0], x[1]))
    if fn == "member":
        return B.read(B.pfnMember(x[0], x[1]))
    if fn == "min":
        return B.read(B.pfnMin(x[0]))
    if fn == "len":
        return B.read(B.pfnLen(x[0]))
    if fn == "eq":
        return B.read(B.pfnEq(x[0], x[1]))
    if fn == "lt":
        return B.read(B.pfnLt(x[0], x[1]))
    if fn == "order":
        return B.read(B.pfnOrderBy(x[0], x[1]))
    if fn == "order_list":
        return B.read(B.pfnOrderBy(B.list([x[0], x[1]]), x[2]))
    if fn == "index_of":
        return B.read(B.pfnIndexOf(x[0], x[1]))
    if fn == "plus_chain":
        return B.read(B.pfnPlus(B.pfnPlus(x[0], x[1]), x[2]))
    if fn == "each_sum_raze":
        return B.read(B.pfnRaze(B.each(B.list([x[0], x[1]]), "pfnSum")))
    if fn == "group":
        keys, counts, flat = B.read_groups(B.pfnGroup(x[0]))
        return {"keys": keys.tolist(), "groups": [g.tolist() for g in np.split(flat, np.cumsum(counts)[:-1])]}
    raise KeyError(fn)


def check(B, e):
    got = run(B, e)
    exp = e["expected"]
    if "keys" in exp:
        assert got == exp, (e["name"], got, exp)
    else:
        assert got.tolist() == exp["values"], (e["name"], got, exp)
        assert got.dtype == np.dtype(NPT[exp["type"]]), (e["name"], got.dtype, exp["type"])
