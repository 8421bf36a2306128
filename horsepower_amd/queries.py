"""HorseIR query plans for the benchmark workloads, written against the pfn* ABI.

``*_plain`` functions are the interpreter's statement sequences of the reference programs
(tests/hyper/online-hir/q6.hir:1-27, q1.hir:1-72, q3.hir:1-85), restated for the synthetic
schemas in datagen.py (dates as day numbers, quantity i32, i32 keys).  ``*_fused`` functions
are what the optimising compiler emits for the same programs -- FE/FA fused loops and FP
grouped reductions (optimizer/autofusion.cc:770-935, optimizer/pattern.cc:180-219;
src/horseir/v1/compile-static/opt/q6.h:31-48, cur/q1.h) -- on the HIP fused entry points.
Both take a backend (backend.Backend) and a dict of column nodes, and return numpy results.
"""
import functools

import numpy as np

from . import datagen as D
from .vnode import H_B, H_C, H_E, H_I, H_L, expr


CHAR_COLUMNS = ("return_flag", "line_status")


def load_table(B, cols):
    """Columns as resident nodes; the flag/status bytes are char (H_C) columns."""
    return {k: B.vec(v, H_C if k in CHAR_COLUMNS else None) for k, v in cols.items()}


# ------------------------------------------------------------------------------------- Q6
def q6_plain(B, T, lo=D.Q6_LO, hi=D.Q6_HI, dlo=0.05, dhi=0.07, qmax=24):
    t18 = B.pfnGeq(T["ship_day"], B.scalar(lo, H_I))
    t19 = B.pfnLt(T["ship_day"], B.scalar(hi, H_I))
    t20 = B.pfnAnd(t18, t19)
    t21 = B.pfnLt(T["quantity"], B.scalar(qmax, H_I))
    t22 = B.pfnGeq(T["discount"], B.scalar(dlo, H_E))
    t23 = B.pfnLeq(T["discount"], B.scalar(dhi, H_E))
    t24 = B.pfnAnd(t22, t23)
    t25 = B.pfnAnd(t20, t21)
    t26 = B.pfnAnd(t25, t24)
    t32 = B.pfnCompress(t26, T["price"])
    t33 = B.pfnCompress(t26, T["discount"])
    t43 = B.pfnMul(t32, t33)
    t44 = B.pfnSum(t43)
    return {"revenue": B.read_scalar(t44)}


@functools.lru_cache(maxsize=256)
def q6_cond(lo=D.Q6_LO, hi=D.Q6_HI, dlo=0.05, dhi=0.07, qmax=24):
    """FA condition over columns (ship_day, quantity, discount, price)."""
    return expr([("COL", 0), ("I", lo), "GEQ", ("COL", 0), ("I", hi), "LT", "AND",
                 ("COL", 1), ("I", qmax), "LT", "AND",
                 ("COL", 2), ("F", dlo), "GEQ", ("COL", 2), ("F", dhi), "LEQ", "AND", "AND"])


def q6_fused(B, T, **kw):
    """q6_chf_01: filter -> compress -> mul -> sum in one pass (opt/q6.h:31-48)."""
    cols = [T["ship_day"], T["quantity"], T["discount"], T["price"]]
    body = expr([("COL", 3), ("COL", 2), "MUL"])
    z = B.fused_reduce(0, cols, q6_cond(**kw), body)
    return {"revenue": B.read_scalar(z)}


def q6_many(B, T, params):
    """Several Q6 requests (one dict of q6 constants each) in one shared scan: one pass over
    the four columns computes every request's revenue."""
    cols = [T["ship_day"], T["quantity"], T["discount"], T["price"]]
    body = expr([("COL", 3), ("COL", 2), "MUL"])
    zs = B.fused_reduce_many(0, cols, [q6_cond(**p) for p in params], body)
    return [{"revenue": B.read_scalar(z)} for z in zs]


# ------------------------------------------------------------------------------------- Q1
Q1_NAMES = ("return_flag", "line_status", "sum_qty", "sum_base_price", "sum_disc_price",
            "sum_charge", "avg_qty", "avg_price", "avg_disc", "count_order")


def q1_plain(B, T, ship_max=D.Q1_SHIP_MAX):
    one = B.scalar(1, H_I)
    t18 = B.pfnLeq(T["ship_day"], B.scalar(ship_max, H_I))
    t23 = B.pfnCompress(t18, T["quantity"])
    t24 = B.pfnCompress(t18, T["price"])
    t25 = B.pfnCompress(t18, T["discount"])
    t26 = B.pfnCompress(t18, T["tax"])
    t27 = B.pfnCompress(t18, T["return_flag"])
    t28 = B.pfnCompress(t18, T["line_status"])
    t35 = B.pfnMinus(one, t25)
    t36 = B.pfnMul(t24, t35)
    t37 = B.pfnPlus(one, t26)
    t38 = B.pfnMinus(one, t25)
    t39 = B.pfnMul(t24, t38)
    t40 = B.pfnMul(t39, t37)
    t41 = B.list([t27, t28])
    t42 = B.pfnGroup(t41)
    t43 = B.pfnKeys(t42)
    t44 = B.pfnValues(t42)
    t45 = B.pfnIndex(t27, t43)
    t46 = B.pfnIndex(t28, t43)

    def agg(col, fn):
        return B.pfnRaze(B.each(B.each_right(col, t44, "pfnIndex"), fn))

    t49 = agg(t23, "pfnSum")
    t52 = agg(t23, "pfnAvg")
    t55 = agg(t24, "pfnSum")
    t58 = agg(t24, "pfnAvg")
    t61 = agg(t36, "pfnSum")
    t64 = agg(t40, "pfnSum")
    t70 = agg(t25, "pfnAvg")
    t73 = agg(t27, "pfnLen")
    t74 = B.list([t45, t46])
    t75 = B.pfnOrderBy(t74, B.bools([1, 1]))
    out = [B.pfnIndex(v, t75) for v in (t45, t46, t49, t55, t61, t64, t52, t58, t70, t73)]
    return dict(zip(Q1_NAMES, B.read_many(out)))


def q1_fused(B, T, ship_max=D.Q1_SHIP_MAX):
    """FA scan (one mask, six columns), FE arithmetic, group, FP1/FP2/FP7 grouped reductions."""
    mask = B.pfnLeq(T["ship_day"], B.scalar(ship_max, H_I))
    qty, price, disc, tax, flag, status = B.compress_many(
        mask, [T["quantity"], T["price"], T["discount"], T["tax"], T["return_flag"], T["line_status"]])
    disc_price = B.fused_map(H_E, [price, disc], expr([("COL", 0), ("I", 1), ("COL", 1), "SUB", "MUL"]))
    charge = B.fused_map(H_E, [price, disc, tax], expr(
        [("COL", 0), ("I", 1), ("COL", 1), "SUB", "MUL", ("I", 1), ("COL", 2), "ADD", "MUL"]))
    g = B.pfnGroup(B.list([flag, status]))
    keys, groups = B.pfnKeys(g), B.pfnValues(g)
    kf, ks = B.pfnIndex(flag, keys), B.pfnIndex(status, keys)
    aggs = [B.group_reduce(0, qty, groups), B.group_reduce(0, price, groups),
            B.group_reduce(0, disc_price, groups), B.group_reduce(0, charge, groups),
            B.group_reduce(1, qty, groups), B.group_reduce(1, price, groups),
            B.group_reduce(1, disc, groups), B.group_reduce(4, flag, groups)]
    perm = B.pfnOrderBy(B.list([kf, ks]), B.bools([1, 1]))
    return dict(zip(Q1_NAMES, B.read_many(B.index_many([kf, ks] + aggs, perm))))


@functools.lru_cache(maxsize=64)
def q1_chf_spec(ship_max=D.Q1_SHIP_MAX):
    """Columns, condition and aggregates of the one-pass Q1 (x = ship_day, quantity, price,
    discount, tax).  disc_price = price*(1-discount), charge = (price*(1-discount))*(1+tax),
    the association of q1.hir's t36 and t40.  Built once per constant (the expression programs
    are read-only)."""
    cond = expr([("COL", 0), ("I", ship_max), "LEQ"])
    qty = expr([("COL", 1)])
    price = expr([("COL", 2)])
    disc = expr([("COL", 3)])
    dprice = expr([("COL", 2), ("I", 1), ("COL", 3), "SUB", "MUL"])
    charge = expr([("COL", 2), ("I", 1), ("COL", 3), "SUB", "MUL", ("I", 1), ("COL", 4), "ADD", "MUL"])
    spec = [("sum", qty), ("sum", price), ("sum", dprice), ("sum", charge),
            ("avg", qty), ("avg", price), ("avg", disc), ("len", None)]
    return cond, spec


def q1_chf(B, T, ship_max=D.Q1_SHIP_MAX):
    """Q1 in one pass over the table (filter, two-char group, eight aggregates).  The fused
    kernel emits its groups in ascending (flag, status) order, each char compared as the order
    by compares it, so q1.hir's order by (t75) is the identity on its output and is not run.
    The results come back with the kernel's own completion check (one synchronisation).  Falls
    back to q1_fused when the key cardinality exceeds the fused kernel's per-workgroup slots."""
    cond, spec = q1_chf_spec(ship_max)
    xs = [T["ship_day"], T["quantity"], T["price"], T["discount"], T["tax"]]
    r = B.fused_group_agg_host([T["return_flag"], T["line_status"]], xs, cond, spec)
    if r is None:
        return q1_fused(B, T, ship_max)
    keys, aggs = r
    return dict(zip(Q1_NAMES, keys + aggs))


@functools.lru_cache(maxsize=64)
def q1_many_spec(cuts):
    """Condition, bucket and sums of a shared Q1 pass over ascending distinct cutoffs `cuts`
    (x = ship_day, quantity, price, discount, tax): rows with ship_day <= the largest cutoff, in
    bucket = the number of cutoffs below their ship_day; the sums every Q1 output derives from."""
    cond = expr([("COL", 0), ("I", cuts[-1]), "LEQ"])
    prog = [("I", 0)]
    for c in cuts[:-1]:
        prog += [("COL", 0), ("I", c), "GT", "ADD"]
    bucket = expr(prog)
    qty, price, disc = expr([("COL", 1)]), expr([("COL", 2)]), expr([("COL", 3)])
    dprice = expr([("COL", 2), ("I", 1), ("COL", 3), "SUB", "MUL"])
    charge = expr([("COL", 2), ("I", 1), ("COL", 3), "SUB", "MUL", ("I", 1), ("COL", 4), "ADD", "MUL"])
    spec = [("sum", qty), ("sum", price), ("sum", dprice), ("sum", charge), ("sum", disc), ("len", None)]
    return cond, bucket, spec


Q1_MANY_MAX = 4


def q1_many(B, T, params):
    """Several Q1 requests (one dict with ship_max each; at most Q1_MANY_MAX distinct cutoffs) in
    one pass over the table: the pass groups on (bucket, flag, status), the bucket being how many
    of the cutoffs lie below a row's ship_day, and each request adds up the buckets at or below
    its own cutoff -- integer sums and counts exactly, FP64 sums with math.fsum of the bucket
    sums -- and divides for the averages as q1_chf's kernel does.  Groups come out in (flag,
    status) order, the order q1_chf returns."""
    import math
    cuts = tuple(sorted(set(int(p.get("ship_max", D.Q1_SHIP_MAX)) for p in params)))
    if len(cuts) > Q1_MANY_MAX:
        raise ValueError(f"q1_many: {len(cuts)} distinct cutoffs (at most {Q1_MANY_MAX})")
    if len(cuts) == 1:
        r = q1_chf(B, T, cuts[0])
        return [dict(r) for _ in params]
    cond, bucket, spec = q1_many_spec(cuts)
    xs = [T["ship_day"], T["quantity"], T["price"], T["discount"], T["tax"]]
    bbits = 1 if len(cuts) <= 2 else 2
    (kf, ks), bk, (sq, sp, sdp, sc, sd, cnt) = B.fused_group_agg_buckets([T["return_flag"], T["line_status"]], xs, cond,
                                                                        bucket, bbits, spec)
    out = []
    for p in params:
        lim = cuts.index(int(p.get("ship_max", D.Q1_SHIP_MAX)))
        acc = {}
        for g in range(bk.size):
            if bk[g] > lim:
                continue
            a = acc.setdefault((int(kf[g]), int(ks[g])), [0, [], [], [], [], 0])
            a[0] += int(sq[g]); a[1].append(float(sp[g])); a[2].append(float(sdp[g])); a[3].append(float(sc[g]))
            a[4].append(float(sd[g])); a[5] += int(cnt[g])
        # (flag, status) in the fused kernel's key order: each char compared as a signed byte
        keys = sorted(acc, key=lambda k: ((k[0] ^ 0x80) << 8) | (k[1] ^ 0x80))
        res = {n: [] for n in Q1_NAMES}
        for k in keys:
            a = acc[k]
            n = a[5]
            s_qty, s_p, s_dp, s_c, s_d = a[0], math.fsum(a[1]), math.fsum(a[2]), math.fsum(a[3]), math.fsum(a[4])
            for name, v in zip(Q1_NAMES, (k[0], k[1], s_qty, s_p, s_dp, s_c, float(s_qty) / n, s_p / n, s_d / n, n)):
                res[name].append(v)
        dts = {"return_flag": kf.dtype, "line_status": ks.dtype, "sum_qty": np.int64, "count_order": np.int64}
        out.append({n: np.array(v, dtype=dts.get(n, np.float64)) for n, v in res.items()})
    return out


# ------------------------------------------------------------------------------------- Q3
Q3_NAMES = ("orderkey", "revenue", "order_day")


def q3_plain(B, Dm, F, day=D.Q3_DAY, limit=10, segment=1):
    d_mask = B.pfnAnd(B.pfnEq(Dm["segment"], B.scalar(segment, H_I)), B.pfnLt(Dm["order_day"], B.scalar(day, H_I)))
    lk = B.pfnCompress(d_mask, Dm["key"])
    lday = B.pfnCompress(d_mask, Dm["order_day"])
    f_mask = B.pfnGt(F["ship_day"], B.scalar(day, H_I))
    rk = B.pfnCompress(f_mask, F["fk"])
    rp = B.pfnCompress(f_mask, F["price"])
    rd = B.pfnCompress(f_mask, F["discount"])
    grp = B.pfnGroup(rk)
    gkeys, gvals = B.pfnKeys(grp), B.pfnValues(grp)
    gk = B.pfnIndex(rk, gkeys)
    rev = B.pfnMul(B.pfnMinus(B.scalar(1, H_L), rd), rp)
    gsum = B.pfnRaze(B.each(B.each_right(rev, gvals, "pfnIndex"), "pfnSum"))
    j = B.pfnJoinIndex(lk, gk, B.sym("eq"))
    j0, j1 = B.pfnIndex(j, B.scalar(0, H_L)), B.pfnIndex(j, B.scalar(1, H_L))
    revenue = B.pfnIndex(gsum, j1)
    oday = B.pfnIndex(lday, j0)
    okey = B.pfnIndex(lk, j0)
    perm = B.pfnOrderBy(B.list([revenue, oday]), B.bools([0, 1]))
    rows = int(perm.contents.len)
    head = B.pfnIndex(perm, B.host.vec(np.arange(min(limit, rows), dtype=np.int64), H_L))   # limit
    res = dict(zip(Q3_NAMES, B.read_many([B.pfnIndex(v, head) for v in (okey, revenue, oday)])))
    res["rows"] = rows
    return res


# (1 - discount) * price over columns (discount, price): q3.hir's revenue, rounded per operation
Q3_REVENUE = expr([("I", 1), ("COL", 0), "SUB", ("COL", 1), "MUL"])


@functools.lru_cache(maxsize=64)
def q3_dim_cond(segment, day):
    """Q3's order filter over (segment, order_day): segment == s and order_day < day."""
    return expr([("COL", 0), ("I", segment), "EQ", ("COL", 1), ("I", day), "LT", "AND"])


@functools.lru_cache(maxsize=64)
def q3_fact_cond(day):
    """Q3's lineitem filter over (discount, price, ship_day): ship_day > day."""
    return expr([("COL", 2), ("I", day), "GT"])


def q3_fused(B, Dm, F, day=D.Q3_DAY, limit=10, segment=1):
    """Q3 as the optimising compiler emits it (src/horseir/v1/compile-static/cur/q3.h:1-56):
    one mask per table and the multi-column peephole compress of its filtered columns with the FE
    revenue map fused into the fact table's compress, group, the grouped sum as one FP loop over the group lists (q3_peephole_1,
    cur/q3.h:30-50: each group's revenue summed through its row indices, nothing gathered
    first), join and the multi-column gathers."""
    # each table's filter evaluated by its compress's count pass (the FA compress scan of one
    # condition): no mask is written
    lk, lday = B.fused_compress([Dm["key"], Dm["order_day"]], [Dm["segment"], Dm["order_day"]], q3_dim_cond(segment, day))
    # the fact compress with the FE revenue map fused: rk and (1 - discount) * price written in
    # the one pass, price and discount never materialised compressed
    (rk,), (rev,) = B.fused_compress_map(q3_fact_cond(day), [F["fk"]], [(H_E, Q3_REVENUE)],
                                         [F["discount"], F["price"], F["ship_day"]])
    grp = B.pfnGroup(rk)
    gkeys, gvals = B.pfnKeys(grp), B.pfnValues(grp)
    gk = B.pfnIndex(rk, gkeys)
    gsum = B.group_reduce(0, rev, gvals)
    j = B.pfnJoinIndex(lk, gk, B.sym("eq"))
    j0, j1 = B.pfnIndex(j, B.scalar(0, H_L)), B.pfnIndex(j, B.scalar(1, H_L))
    revenue = B.pfnIndex(gsum, j1)
    okey, oday = B.index_many([lk, lday], j0)
    perm = B.pfnOrderBy(B.list([revenue, oday]), B.bools([0, 1]))
    rows = int(perm.contents.len)
    head = B.pfnIndex(perm, B.host.vec(np.arange(min(limit, rows), dtype=np.int64), H_L))
    res = dict(zip(Q3_NAMES, B.read_many(B.index_many([okey, revenue, oday], head))))
    res["rows"] = rows
    return res


# ---------------------------------------------------------------------- batched group-by
BATCH_NAMES = ("key", "count", "sum")


def batch_plain(B, T, thr):
    """One query of the batched workload as the interpreter runs it: rows with flt < thr,
    grouped on key, FP1 sum of val and FP7 len per group (optimizer/pattern.cc:180-219);
    (key, count, sum) over the groups in pfnGroup order."""
    m = B.pfnLt(T["flt"], B.scalar(int(thr), H_I))
    k, v = B.pfnCompress(m, T["key"]), B.pfnCompress(m, T["val"])
    g = B.pfnGroup(k)
    keys, groups = B.pfnKeys(g), B.pfnValues(g)
    gk = B.pfnIndex(k, keys)
    s = B.pfnRaze(B.each(B.each_right(v, groups, "pfnIndex"), "pfnSum"))
    n = B.pfnRaze(B.each(B.each_right(k, groups, "pfnIndex"), "pfnLen"))
    return dict(zip(BATCH_NAMES, B.read_many([gk, n, s])))


def batch_fused(B, T, thr):
    """The same query with the compiler's fused forms: one mask, peephole compress of both
    columns, group, FP1/FP7 grouped reductions over the group lists."""
    m = B.pfnLt(T["flt"], B.scalar(int(thr), H_I))
    k, v = B.compress_many(m, [T["key"], T["val"]])
    g = B.pfnGroup(k)
    keys, groups = B.pfnKeys(g), B.pfnValues(g)
    gk = B.pfnIndex(k, keys)
    return dict(zip(BATCH_NAMES, B.read_many([gk, B.group_reduce(4, k, groups), B.group_reduce(0, v, groups)])))
