"""Benchmark of the HorseIR primitive layer on one MI355X.

Default workload: the hash group-by query (TPC-H Q1 shape) over a 60,000,000-row fact table
held in HBM -- filter ship_day <= 2434, group on (return_flag, line_status), eight FP64/int64
aggregates, order groups by key.  One step runs the whole query through the C-ABI and reads the
result back; inputs are resident before the timed region.  `value` is fact rows per second.

The JSON line carries:
  roofline      dominant kernel of the step (largest device time, HIP events on the launching
                stream): algorithmic bytes per launch / average launch duration vs 8 TB/s
  cpu_baseline  the oracle (C + OpenMP restatement of the reference primitives, the reference's
                interpreter statement sequence) on a bounded sample of the same workload
  plain         the same query through the interpreter's per-primitive sequence on the GPU

Workload "batch": 64 independent filter -> group -> sum queries (row counts log-uniform 1e4..1e7,
key ranges log-uniform 4..1e6) in one persistent launch replayed from a captured hipGraph, with
the autotuner's configuration; `value` is queries per second, compared with the same 64 queries
run one by one through the per-query primitive plans on the GPU.

Workload "prims": every hot-path primitive alone at the configs' sizes (60M-row columns of the
group-by table; 15M-row sorts and groups; 1.5M x 6M joins), device time per call from HIP events
around the whole call, algorithmic bytes per call, fraction of the 8 TB/s HBM peak.

Workload "serve": a Poisson stream of mixed Q6/Q1/Q3 requests with random constants over a
600M-row lineitem and 150M-row orders table generated in HBM; the server (continuous batching,
memory-budget admission, one stream per slot) against the same stream run one request at a time.

Workload "copy": a plain device copy of a 60M-element f64 column (the generated FE map of one
column), 480 MB read and 480 MB written per launch: the reference for scaling the HBM traffic
counters (FETCH_SIZE, WRITE_SIZE) of the other kernels.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload q1|q6|q3|batch|prims|serve|copy]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0

# HBM traffic per launch measured with the FETCH_SIZE / WRITE_SIZE counters (scaled by a copy of
# known size); the summaries hold the commands and the raw counter values.
TRAFFIC = {("q1", 60_000_000, "fused_group_agg"): "profiles/fga_q1_60M_traffic.md"}


def measured_traffic(workload, rows, kernel):
    f = TRAFFIC.get((workload, rows, kernel))
    path = os.path.join(ROOT, f) if f else None
    if not path or not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=50)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--workload", default="all", choices=["all", "q1", "q6", "q3", "batch", "prims", "serve", "copy", "ingest"])
    p.add_argument("--rows", type=int, default=0)
    p.add_argument("--plan", default="best", choices=["best", "fused", "plain"])
    p.add_argument("--cpu-rows", type=int, default=0)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-plain", action="store_true")
    p.add_argument("--profile-steps", type=int, default=3)
    p.add_argument("--requests", type=int, default=120)
    p.add_argument("--rate", type=float, default=0.0)
    p.add_argument("--slots", type=int, default=4)
    p.add_argument("--heavy-share", type=int, default=-1)
    p.add_argument("--heavy-slots", type=int, default=2)
    p.add_argument("--heavy-fifo", type=int, default=0)
    p.add_argument("--late-share", type=int, default=0)
    return p.parse_args()


def workload(name, rows):
    from horsepower_amd import datagen as D, queries as Q
    if name == "q1":
        n = rows or 60_000_000
        return dict(n=n, gen=lambda k: (D.gen_q1(k),), plans=dict(best=Q.q1_chf, fused=Q.q1_fused, plain=Q.q1_plain),
                    desc=f"q1 hash group-by, {n} rows, filter ship_day<=2434, 8 aggregates, order by key",
                    cpu_default=6_000_000)
    if name == "q6":
        n = rows or 60_000_000
        return dict(n=n, gen=lambda k: (D.gen_q6(k),), plans=dict(best=Q.q6_fused, fused=Q.q6_fused, plain=Q.q6_plain),
                    desc=f"q6 filter-compress-aggregate, {n} rows", cpu_default=6_000_000)
    nd = (rows or 60_000_000) // 4
    return dict(n=None, nd=nd, gen=lambda k: D.gen_q3(k), plans=dict(best=Q.q3_fused, fused=Q.q3_fused, plain=Q.q3_plain),
                desc=f"q3 join+group+top-10, {nd} orders", cpu_default=1_500_000)


def run_steps(B, plan, tables, steps, mark):
    # the collector stays off while steps run (as timeit does): a collection inside a short timed
    # region shows up as host time of the step it lands in
    import gc
    was = gc.isenabled()
    gc.disable()
    try:
        for _ in range(steps):
            plan(B, *tables)
            B.release(mark)
    finally:
        if was:
            gc.enable()


def free_tables(B, tables):
    """Release the resident column payloads of the given tables (dicts of column nodes)."""
    for T in tables:
        for node in T.values():
            v = node.contents
            if v.len > 1 and v.g and B.lib.dll.hpaIsDevicePtr(v.g):
                B.lib.dll.hpaFreeResident(node)


HEAD = {"metric": "rows/s + % HBM roofline per query; batched queries/s & p99 latency vs plain",
        "n_gpus": 1, "higher_is_better": True, "scaling": "weak", "vs_baseline": None}


def bench_query(a, B, name, rows=0, tables=None, desc=None, cpu=True, plain=True):
    """One query workload (q1 / q6 / q3): K timed steps over resident inputs, a kernel-timing
    pass, the interpreter sequence on the GPU and the oracle on a bounded sample."""
    import ctypes as C
    W = workload(name, rows)
    own = tables is None
    if own:
        t0 = time.time()
        cols = W["gen"](W["n"] if W["n"] else W["nd"])
        print(f"[bench] generated {name} in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
        from horsepower_amd import queries as Q
        tables = tuple(Q.load_table(B, c) for c in cols)
        nrows = W["n"] if W["n"] else sum(len(next(iter(c.values()))) for c in cols)
        del cols
    else:
        nrows = sum(next(iter(T.values())).contents.len for T in tables) if not W["n"] else W["n"]
        if name == "q3":
            nrows = sum(T[k].contents.len for T, k in zip(tables, ("key", "fk")))
    lib = B.lib
    plan = W["plans"][a.plan]
    B.sync()
    mark = B.mark()
    run_steps(B, plan, tables, a.warmup, mark)
    B.sync()
    t0 = time.perf_counter()
    run_steps(B, plan, tables, a.steps, mark)
    B.sync()
    dt = time.perf_counter() - t0
    ms = dt * 1000.0 / a.steps
    print(f"[bench] {name} {nrows} rows: {a.steps} steps, {ms:.3f} ms/step", file=sys.stderr, flush=True)
    kern = profile_kernels(lib, lambda: run_steps(B, plan, tables, 1, mark), a.profile_steps)
    # the dominant kernel scope among those with algorithmic bytes (a scope without them, such as
    # the reference-order quicksort, has no roofline)
    withb = {k: v for k, v in kern.items() if v["bytes"] > 0}
    dom = max(withb, key=lambda k: withb[k]["ms"]) if withb else (max(kern, key=lambda k: kern[k]["ms"]) if kern else None)
    roof = None
    if dom:
        roof = roofline_of(kern, dom)
        roof["traffic"] = measured_traffic(name, nrows, dom)
    per_kernel = {k: {"launches_per_step": v["count"] / a.profile_steps,
                      "us_per_step": round(v["ms"] * 1e3 / a.profile_steps, 2),
                      "GBs": round(v["bytes"] / (v["ms"] * 1e-3) / 1e9, 1) if v["ms"] > 0 and v["bytes"] > 0 else None}
                  for k, v in sorted(kern.items(), key=lambda kv: -kv[1]["ms"])}
    pl = None
    if plain and a.plan != "plain" and not a.no_plain:
        pp = W["plans"]["plain"]
        run_steps(B, pp, tables, 1, mark)
        B.sync()
        t1 = time.perf_counter()
        ps = max(2, a.steps // 4)
        run_steps(B, pp, tables, ps, mark)
        B.sync()
        pdt = (time.perf_counter() - t1) / ps
        pl = {"ms_per_step": round(pdt * 1e3, 3), "rows_per_s": nrows / pdt}
    cb = cpu_baseline(a, W) if cpu and not a.no_cpu else None
    B.release(mark)
    if own:
        free_tables(B, tables)
    return {"value": nrows * a.steps / dt, "unit": "rows/s", "steps": a.steps, "warmup": a.warmup,
            "ms_per_step": round(ms, 4), "dtype": "f64", "data": "synthetic (seeded counter-based generator)",
            "config": {"workload": desc or W["desc"], "rows": nrows, "plan": a.plan},
            "roofline": roof, "cpu_baseline": cb, "plain_gpu": pl, "kernels": per_kernel}


def bench_all(a):
    """The default run: the 60M-row group-by query as the headline line, and every other
    workload of the metric measured in the same process as sub-objects: the filter query at 6M
    rows, the join query at 15M and 60M lineitems, the 64-query batch and the serving stream
    over 600M + 150M rows, and each primitive alone."""
    os.environ.setdefault("HPA_ARENA_BYTES", str(96 << 30))
    from horsepower_amd.backend import GpuBackend
    B = GpuBackend()
    t_all = time.time()
    head = bench_query(a, B, "q1", a.rows or 60_000_000)
    subs = {}
    q6 = workload("q6", 6_000_000)
    subs["q6_6M"] = bench_query(a, B, "q6", 6_000_000)
    for nd, tag in ((3_750_000, "q3_15M"), (15_000_000, "q3_60M")):
        dim, fact = B.gen_serve_tables(nd, 1002)
        nf = fact["fk"].contents.len
        subs[tag] = bench_query(a, B, "q3", nd * 4, tables=(dim, fact), cpu=(tag == "q3_15M"),
                                desc=f"q3 join+group+top-10, {nd} orders, {nf} lineitems (generated in HBM)")
        free_tables(B, (dim, fact))
    print(f"[bench] query workloads done at {time.time() - t_all:.0f}s", file=sys.stderr, flush=True)
    subs["batch_64"] = bench_batch(a, B, cpu_budget=10.0)
    print(f"[bench] batch done at {time.time() - t_all:.0f}s", file=sys.stderr, flush=True)
    subs["serve_600M"] = bench_serve(a, B)
    print(f"[bench] serve done at {time.time() - t_all:.0f}s", file=sys.stderr, flush=True)
    subs["prims"] = bench_prims(a, B, cpu_budget=1.5)
    subs["ingest_15M"] = bench_ingest(a, B, n_dim=3_750_000)
    print(f"[bench] all workloads in {time.time() - t_all:.0f}s", file=sys.stderr, flush=True)
    for v in subs.values():
        for k in HEAD:
            v.pop(k, None)
    out = dict(HEAD)
    out.update(head)
    out["workloads"] = subs
    del q6
    return out


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {a.gpus}); this build runs on a single "
              "MI355X and has no multi-GPU path", file=sys.stderr)
        sys.exit(2)
    from horsepower_amd.backend import GpuBackend
    if a.workload == "all":
        out = bench_all(a)
    elif a.workload == "batch":
        out = bench_batch(a)
    elif a.workload == "prims":
        out = bench_prims(a)
    elif a.workload == "serve":
        out = bench_serve(a)
    elif a.workload == "copy":
        out = bench_copy(a)
    elif a.workload == "ingest":
        out = bench_ingest(a)
    else:
        if a.workload == "q3":
            os.environ.setdefault("HPA_ARENA_BYTES", str(64 << 30))
        out = dict(HEAD)
        out.update(bench_query(a, GpuBackend(), a.workload, a.rows))
    for k, v in HEAD.items():
        out.setdefault(k, v)
    print(json.dumps(out), flush=True)


def profile_kernels(lib, fn, steps):
    """Per-scope device time and algorithmic bytes from the library's HIP-event profiler."""
    lib.dll.hpaProfileReset()
    lib.dll.hpaProfileEnable(1)
    for _ in range(steps):
        fn()
    lib.sync()
    lib.dll.hpaProfileEnable(0)
    return read_profile(lib)


def read_profile(lib):
    import ctypes as C
    buf = C.create_string_buffer(1 << 16)
    lib.dll.hpaProfileList(buf, 1 << 16)
    kern = {}
    for line in buf.value.decode().splitlines():
        name, cnt, tms, byt = line.split()
        kern[name] = dict(count=int(cnt), ms=float(tms), bytes=float(byt))
    return kern


def roofline_of(kern, name):
    k = kern[name]
    avg_ms = k["ms"] / k["count"]
    per_launch = k["bytes"] / k["count"]
    achieved = per_launch / (avg_ms * 1e-3) / 1e9
    return {"bound": "hbm", "kernel": name, "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS,
            "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": None,
            "bytes_per_launch": per_launch, "avg_launch_us": round(avg_ms * 1e3, 2)}


def pct(xs, p):
    import numpy as np
    return float(np.percentile(np.asarray(xs), p))


def bench_batch(a, B=None, cpu_budget=25.0):
    import numpy as np
    from horsepower_amd import datagen as D, queries as Q
    from horsepower_amd.backend import Batch, GpuBackend
    nq = 64
    ns, gs, thr = D.batch_shapes(nq)
    t0 = time.time()
    cols = [D.gen_batch_query(q, int(ns[q]), int(gs[q])) for q in range(nq)]
    rows = int(ns.sum())
    print(f"[bench] generated {nq} queries, {rows} rows in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    if B is None:
        B = GpuBackend()
    tables = [Q.load_table(B, c) for c in cols]
    B.sync()
    b = Batch(B, tables, thr)
    tune = b.autotune(2)
    cfg = b.config()
    print(f"[bench] autotuned {len(tune)} candidates -> {cfg}", file=sys.stderr, flush=True)
    mark = B.mark()

    def step():
        b.run(True)
        return b.results()

    for _ in range(a.warmup):
        step()
    lat = []
    B.sync()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        t1 = time.perf_counter()
        res = step()
        lat.append(time.perf_counter() - t1)
    dt = time.perf_counter() - t0
    ms = dt * 1e3 / a.steps
    print(f"[bench] batched: {ms:.3f} ms per batch of {nq}", file=sys.stderr, flush=True)
    # a stream of batches: the results of batch k travel to the host while batch k+1 computes
    b.run(True)
    b.fetch_begin()
    B.sync()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        b.run(True)
        b.fetch_begin()
        pres = b.fetch_end()
    pres = b.fetch_end()
    pdt = time.perf_counter() - t0
    pms = pdt * 1e3 / a.steps
    pipe_equal = all(np.array_equal(pres[q]["key"], res[q]["key"]) and np.array_equal(pres[q]["sum"], res[q]["sum"])
                     for q in range(nq))
    print(f"[bench] pipelined: {pms:.3f} ms per batch of {nq} (equal={pipe_equal})", file=sys.stderr, flush=True)
    kern = profile_kernels(B.lib, step, a.profile_steps)
    roof = roofline_of(kern, "batch_graph") if "batch_graph" in kern else None

    # the same queries one by one through the per-query primitive plans
    def per_query(plan):
        out, t = [], []
        for q in range(nq):
            t1 = time.perf_counter()
            out.append(plan(B, tables[q], thr[q]))
            t.append(time.perf_counter() - t1)
            B.release(mark)
        return out, t
    per_query(Q.batch_fused)
    plain_res, plain_t = per_query(Q.batch_plain)
    fused_res, fused_t = per_query(Q.batch_fused)
    equal = True
    for q in range(nq):
        for other in (plain_res[q], fused_res[q]):
            equal &= bool(np.array_equal(res[q]["key"], other["key"]) and
                          np.array_equal(res[q]["count"], other["count"]) and
                          np.allclose(res[q]["sum"], other["sum"], rtol=1e-12, atol=0))
    print(f"[bench] per-query plain {sum(plain_t) * 1e3:.1f} ms, fused {sum(fused_t) * 1e3:.1f} ms, "
          f"equal={equal}", file=sys.stderr, flush=True)

    cpu = None
    if not a.no_cpu:
        import oracle
        ob = oracle.backend(os.cpu_count() or 1)
        om = ob.mark()
        k = 0
        t0 = time.perf_counter()
        while k < nq and time.perf_counter() - t0 < cpu_budget:
            Q.batch_plain(ob, Q.load_table(ob, cols[k]), thr[k])
            ob.release(om)
            k += 1
        cdt = time.perf_counter() - t0
        crow = int(ns[:k].sum())
        cpu = {"value": k / cdt, "unit": "queries/s", "cores": ob.threads(), "kind": "port",
               "sample": f"the first {k} of the {nq} queries ({crow} rows), one run each, in order",
               "rows_per_s": crow / cdt}
    out = {
        "metric": "rows/s + % HBM roofline per query; batched queries/s & p99 latency vs plain",
        "value": nq / (pdt / a.steps), "unit": "queries/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(pms, 4), "higher_is_better": True, "scaling": "weak",
        "timing": "a stream of batches: batch k's results are copied to host memory while batch k+1 "
                  "computes (hpaBatchFetchBegin/End); one_at_a_time waits for each batch's results",
        "one_at_a_time": {"queries_per_s": nq / (dt / a.steps), "ms_per_batch": round(ms, 4)},
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded counter-based generator)",
        "config": {"workload": f"batched group-by: {nq} queries in one persistent launch, {rows} rows, "
                               "row counts log-uniform 1e4..1e7, key ranges log-uniform 4..1e6",
                   "rows": rows, "queries": nq, "tuned": cfg},
        "rows_per_s": rows / (pdt / a.steps),
        "latency_ms": {"p50": round(pct(lat, 50) * 1e3, 3), "p99": round(pct(lat, 99) * 1e3, 3)},
        "plain_gpu": {"queries_per_s": nq / sum(plain_t), "ms_per_batch": round(sum(plain_t) * 1e3, 3),
                      "query_p50_ms": round(pct(plain_t, 50) * 1e3, 3), "query_p99_ms": round(pct(plain_t, 99) * 1e3, 3)},
        "fused_gpu": {"queries_per_s": nq / sum(fused_t), "ms_per_batch": round(sum(fused_t) * 1e3, 3)},
        "equal_results": equal and pipe_equal, "roofline": roof, "cpu_baseline": cpu,
        "kernels": {k: {"launches_per_step": v["count"] / a.profile_steps,
                        "us_per_step": round(v["ms"] * 1e3 / a.profile_steps, 2)} for k, v in kern.items()},
        "autotune_ms": [round(x, 4) for x in tune],
    }
    b.close()
    free_tables(B, tables)
    return out


def prims_cases(B, n, gpu):
    """Inputs and calls of the per-primitive measurements on backend B: n-row Q1 columns, n/4-row
    sort/group columns (a 1e6-key i32 column), a join/index_of/member pair of n/40 unique sparse
    build keys and n/10 probe keys.  Each case: (name, algorithmic bytes, call).  The fused
    reduction has no oracle counterpart and is listed on the GPU only."""
    import numpy as np
    from horsepower_amd import datagen as D, queries as Q
    from horsepower_amd.vnode import H_E, H_I, H_L, expr
    ns = max(n // 4, 1)                       # sorts, groups and hash builds: config 2's 15M rows
    t = D.gen_q1(n)
    T = Q.load_table(B, t)
    key = B.vec(D.uniform(7, 1, ns, 0, 999_999))                   # 1e6-key i32 group column
    dim_keys = D.gen_q3(ns // 10)[0]["key"]
    build = B.vec(dim_keys)                                        # unique sparse keys
    probe = B.vec(dim_keys[D.uniform(7, 2, 4 * len(dim_keys), 0, len(dim_keys) - 1, np.int64)])
    idx = B.vec(D.uniform(7, 3, n // 2, 0, n - 1, np.int64))
    m1 = B.pfnLeq(T["ship_day"], B.scalar(2434, H_I))
    m6 = B.pfnAnd(B.pfnGeq(T["ship_day"], B.scalar(730, H_I)), B.pfnLt(T["ship_day"], B.scalar(1095, H_I)))
    k1 = int(B.read_scalar(B.pfnSum(m1)))
    k6 = int(B.read_scalar(B.pfnSum(m6)))
    B.sync()
    nb, npb = len(dim_keys), 4 * len(dim_keys)
    cases = [
        ("compare_i32", 5 * n, lambda: B.pfnLeq(T["ship_day"], B.scalar(2434, H_I))),
        ("compare_f64", 9 * n, lambda: B.pfnGeq(T["discount"], B.scalar(0.05, H_E))),
        ("and", 3 * n, lambda: B.pfnAnd(m1, m6)),
        ("arith_EE", 24 * n, lambda: B.pfnMul(T["price"], T["discount"])),
        ("arith_sE", 16 * n, lambda: B.pfnMinus(B.scalar(1, H_L), T["discount"])),
        ("compress_E_96pct", n * 9 + 8 * k1, lambda: B.pfnCompress(m1, T["price"])),
        ("compress_E_14pct", n * 9 + 8 * k6, lambda: B.pfnCompress(m6, T["price"])),
        ("where_14pct", n + 8 * k6, lambda: B.pfnWhere(m6)),
        ("index_E", (n // 2) * (8 + 16), lambda: B.pfnIndex(T["price"], idx)),
        ("sum_E", 8 * n, lambda: B.pfnSum(T["price"])),
        ("group_trie", 2 * n + 8 * n + 8 * 4, lambda: B.pfnGroup(B.list([T["return_flag"], T["line_status"]]))),
        ("group_i32_1e6keys", 12 * ns + 8 * 1_000_000, lambda: B.pfnGroup(key)),
        ("order_i32", 12 * ns, lambda: B.pfnOrderBy(B.list([key]), B.bools([1]))),
        ("join_i32", 4 * (nb + npb) + 16 * npb, lambda: B.pfnJoinIndex(build, probe, B.sym("eq"))),
        ("index_of_i32", 4 * (nb + npb) + 8 * npb, lambda: B.pfnIndexOf(build, probe)),
        ("member_i32", 4 * (nb + npb) + npb, lambda: B.pfnMember(probe, build)),
    ]
    if gpu:
        cond6 = Q.q6_cond()
        cases.append(("fused_q6_reduce", 24 * n, lambda: B.fused_reduce(0, [T["ship_day"], T["quantity"], T["discount"], T["price"]],
                                                                        cond6, expr([("COL", 3), ("COL", 2), "MUL"]))))
    return cases, {"ns": ns, "nb": nb, "npb": npb,
                   "tables": [T, {"key": key, "build": build, "probe": probe, "idx": idx}]}


def bench_prims(a, B=None, cpu_budget=5.0):
    import math
    from horsepower_amd.backend import GpuBackend
    n = 60_000_000 if (B is not None or not a.rows) else a.rows
    t0 = time.time()
    if B is None:
        B = GpuBackend()
    cases, sz = prims_cases(B, n, True)
    print(f"[bench] prims inputs ready in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    lib = B.lib
    mark = B.mark()
    reps = max(a.steps // 4, 3)
    res = {}
    def timed(fn, name, nbytes, mode):
        B.sync()
        lib.dll.hpaProfileReset()
        lib.dll.hpaProfileEnable(mode)
        for _ in range(reps):
            lib.dll.hpaProfileBegin(name.encode(), float(nbytes))
            fn()
            lib.dll.hpaProfileEnd()
            B.release(mark)
        B.sync()
        lib.dll.hpaProfileEnable(0)
        return read_profile(B.lib)

    for name, nbytes, fn in cases:
        fn()
        B.release(mark)
        # the call's time from a pass without the library's own launch events; the breakdown into
        # launch scopes from a second pass with them
        outer = timed(fn, name, nbytes, 2)
        kern = timed(fn, name, nbytes, 1)
        ms = outer[name]["ms"] / max(outer[name]["count"], 1)
        gbs = nbytes / (ms * 1e-3) / 1e9
        inner = {k: round(v["ms"] * 1e3 / v["count"], 2) for k, v in kern.items() if k != name}
        res[name] = {"bytes": nbytes, "us": round(ms * 1e3, 2), "GBs": round(gbs, 1),
                     "frac": round(gbs / PEAK_HBM_GBS, 4), "launch_scopes_us": inner}
        print(f"[bench] {name:20s} {ms * 1e3:9.1f} us {gbs:8.1f} GB/s {gbs / PEAK_HBM_GBS:6.1%}  {inner}", file=sys.stderr, flush=True)
    cpu = None
    if not a.no_cpu:
        cpu = prims_cpu(a, res, cpu_budget)
    gm = math.exp(sum(math.log(v["GBs"]) for v in res.values()) / len(res))
    out = {"metric": "rows/s + % HBM roofline per query; batched queries/s & p99 latency vs plain",
           "value": round(gm, 1), "unit": "GB/s (geometric mean over primitives)", "n_gpus": 1,
           "steps": reps, "warmup": 1, "ms_per_step": None, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "mixed (i32/f64/u8 as each primitive)",
           "data": "synthetic (seeded counter-based generator)",
           "config": {"workload": f"each primitive alone: {n}-row columns, {sz['ns']}-row sorts/groups, "
                                  f"join/index_of/member {sz['nb']} x {sz['npb']}", "rows": n},
           "cpu_baseline": cpu, "primitives": res}
    B.release(mark)
    free_tables(B, sz["tables"])
    return out


def prims_cpu(a, res, budget=5.0):
    """The same primitives on the oracle (C + OpenMP restatement of the reference's, all host
    cores; the reference's serial phases stay serial) over a smaller sample: one discarded call,
    then the mean of up to three.  Adds cpu_us / cpu_GBs / gpu_over_cpu to each primitive."""
    import math
    import oracle
    threads = os.cpu_count() or 1
    ob = oracle.backend(threads)
    n = a.cpu_rows or 6_000_000
    cases, sz = prims_cases(ob, n, False)
    mark = ob.mark()
    rates = []
    for name, nbytes, fn in cases:
        fn()
        ob.release(mark)
        times = []
        t_start = time.perf_counter()
        while len(times) < 3 and time.perf_counter() - t_start < budget:
            t1 = time.perf_counter()
            fn()
            times.append(time.perf_counter() - t1)
            ob.release(mark)
        s = sum(times) / len(times)
        gbs = nbytes / s / 1e9
        rates.append(gbs)
        if name in res:
            res[name].update({"cpu_us": round(s * 1e6, 1), "cpu_GBs": round(gbs, 2),
                              "gpu_over_cpu": round(res[name]["GBs"] / gbs, 1)})
        print(f"[bench] cpu {name:20s} {s * 1e6:11.1f} us {gbs:8.2f} GB/s", file=sys.stderr, flush=True)
    gm = math.exp(sum(math.log(r) for r in rates) / len(rates))
    return {"value": round(gm, 2), "unit": "GB/s (geometric mean over primitives)", "cores": ob.threads(),
            "kind": "port", "sample": f"{n}-row columns, {sz['ns']}-row sorts/groups, join/index_of/member "
                                      f"{sz['nb']} x {sz['npb']}; one discarded call, mean of up to three"}


def bench_serve(a, B=None):
    import numpy as np
    from horsepower_amd import serve as S
    from horsepower_amd.backend import GpuBackend
    n_dim = a.rows or 150_000_000
    # the one-at-a-time paths run on the default context: its arena must hold the interpreter's
    # Q1 over 600M rows (about 50 GiB); the serving pool takes most of the rest of the 288 GB next
    # to the 24.6 GB of tables, room for three Q3s at once (about 39 GiB each)
    os.environ.setdefault("HPA_ARENA_BYTES", str(96 << 30))
    if B is None:
        B = GpuBackend()
    t0 = time.time()
    dim, fact = B.gen_serve_tables(n_dim, 1004)
    T = {"dim": dim, "fact": fact}
    nf = fact["fk"].contents.len
    print(f"[bench] serve tables: {n_dim} orders, {nf} lineitems in HBM ({time.time() - t0:.1f}s)", file=sys.stderr, flush=True)
    srv = S.Server(B, T, slots=a.slots, pool_bytes=128 << 30, merge=True,
                   heavy_share=None if a.heavy_share < 0 else a.heavy_share, heavy_slots=a.heavy_slots,
                   heavy_fifo=bool(a.heavy_fifo), late_share=a.late_share)
    est = srv.calibrate()
    print(f"[bench] calibrated arena estimates (MiB): { {k: v >> 20 for k, v in est.items()} }", file=sys.stderr, flush=True)
    # service time of the per-query fused path sets the offered load
    probe = S.make_requests(12, rate=None, seed=99)
    S.serve_serial(B, T, probe, plans="fused", realtime=False)
    mean_s = float(np.mean([r.done - r.admit for r in probe]))
    rate = a.rate or 0.9 / mean_s
    print(f"[bench] serial fused mean service {mean_s * 1e3:.2f} ms -> rate {rate:.1f}/s", file=sys.stderr, flush=True)
    N = a.requests
    runs = {}
    runs["server"] = S.summarize(srv.serve(S.make_requests(N, rate, seed=2024)))
    print(f"[bench] server: {runs['server']}", file=sys.stderr, flush=True)
    # the tail-first policy on the same stream: no short request that arrived after a running
    # heavy one runs beside it
    fifo, late = srv.heavy_fifo, srv.late_share
    srv.heavy_fifo, srv.late_share = True, 0
    runs["server_tail_first"] = S.summarize(srv.serve(S.make_requests(N, rate, seed=2024)))
    srv.heavy_fifo, srv.late_share = fifo, late
    print(f"[bench] server, tail first: {runs['server_tail_first']}", file=sys.stderr, flush=True)
    ser = S.serve_serial(B, T, S.make_requests(N, rate, seed=2024), plans="fused")
    runs["serial_fused"] = S.summarize(ser)
    print(f"[bench] serial fused: {runs['serial_fused']}", file=sys.stderr, flush=True)
    closed_srv = srv.serve(S.make_requests(N, None, seed=2025), realtime=False)
    runs["server_closed"] = S.summarize(closed_srv)
    closed_ser = S.serve_serial(B, T, S.make_requests(N, None, seed=2025), plans="fused", realtime=False)
    runs["serial_fused_closed"] = S.summarize(closed_ser)
    plain = S.serve_serial(B, T, S.make_requests(N, None, seed=2025), plans="plain", realtime=False)
    runs["serial_plain_closed"] = S.summarize(plain)
    print(f"[bench] closed loop: server {runs['server_closed']['queries_per_s']:.1f} q/s, serial fused "
          f"{runs['serial_fused_closed']['queries_per_s']:.1f} q/s, serial plain {runs['serial_plain_closed']['queries_per_s']:.1f} q/s",
          file=sys.stderr, flush=True)
    equal = S.same_results(closed_srv, closed_ser) == [] and S.same_results(closed_srv, plain) == []
    srv.close()
    out = {"metric": "rows/s + % HBM roofline per query; batched queries/s & p99 latency vs plain",
           "value": runs["server"]["queries_per_s"], "unit": "queries/s", "n_gpus": 1, "steps": N, "warmup": 12,
           "ms_per_step": runs["server"]["p50_ms"], "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded counter-based generator, in HBM)",
           "config": {"workload": f"serving: Poisson mix of q6/q1/q3 requests (45/35/20%) with random constants, "
                                  f"{nf} lineitem + {n_dim} order rows resident", "rate_per_s": rate,
                      "slots": a.slots, "requests": N},
           "latency_ms": {"p50": runs["server"]["p50_ms"], "p99": runs["server"]["p99_ms"],
                          "tail_first_p50": runs["server_tail_first"]["p50_ms"], "tail_first_p99": runs["server_tail_first"]["p99_ms"],
                          "one_at_a_time_p50": runs["serial_fused"]["p50_ms"], "one_at_a_time_p99": runs["serial_fused"]["p99_ms"]},
           "closed_loop_vs_one_at_a_time": round(runs["server_closed"]["queries_per_s"] / runs["serial_fused_closed"]["queries_per_s"], 3),
           "runs": runs, "equal_results": equal, "arena_estimates_mib": {k: v >> 20 for k, v in est.items()},
           "peak_pool_in_use_mib": srv.peak_in_use >> 20, "cpu_baseline": None,
           "cpu_baseline_note": "the reference has no serving loop; its per-request primitives are timed in the query workloads"}
    free_tables(B, (dim, fact))
    return out


def free_node(B, z):
    """hpaFreeResident of every HBM payload under a node the .bin reader made."""
    from horsepower_amd.runtime import cell
    from horsepower_amd.vnode import H_A, H_G
    v = z.contents
    if v.typ in (H_A, H_G):
        for k in range(2 if v.typ == H_A else v.len):
            free_node(B, cell(z, k))
    elif v.len > 1 and B.lib.dll.hpaIsDevicePtr(v.g):
        B.lib.rc("hpaFreeResident", z)


def bench_ingest(a, B=None, n_dim=0):
    """Table snapshot ingest: the lineitem table of the serving workload (8 columns) written in
    the reference's .bin format (serializeV), then read back into HBM (readSerializeV with the
    payloads streamed through pinned staging), timed over several reads of the file from the page
    cache; the oracle's host reader of the same file is the CPU baseline."""
    import ctypes as C
    import tempfile
    import numpy as np
    from horsepower_amd.backend import GpuBackend
    from horsepower_amd.datagen import SERVE_FACT
    from horsepower_amd.vnode import H_A, H_G, H_Q, V, V0
    if B is None:
        B = GpuBackend()
    n_dim = n_dim or (a.rows // 4 if a.rows else 15_000_000)
    dim, fact = B.gen_serve_tables(n_dim, 1004)
    nf = fact["fk"].contents.len
    cols = [fact[k] for k in SERVE_FACT]
    names = B.host.vec(np.array([B.symbol("l_" + k) for k in SERVE_FACT], np.int64), H_Q)
    vals = B.list(cols)
    cells = (V0 * 2)()
    cells[0], cells[1] = names.contents, vals.contents

    class A0(C.Structure):
        _fields_ = [("row", C.c_int64), ("col", C.c_int64), ("fkey", C.c_void_p)]
    a0 = A0(nf, len(cols), None)
    t = C.pointer(V0())
    t.contents.typ, t.contents.len = H_A, 2
    t.contents.g = C.addressof(cells)
    t.contents.p = C.addressof(a0)
    fd, path = tempfile.mkstemp(suffix=".bin", prefix="hpa_lineitem_")
    os.close(fd)
    try:
        t0 = time.time()
        B.lib.rc("hpaWriteSerialized", t, path.encode())
        t_write = time.time() - t0
        fbytes = os.path.getsize(path)
        free_tables(B, (dim, fact))
        reps = max(min(a.steps, 5), 2)
        times, moved = [], C.c_int64(0)
        for r in range(reps + 1):
            z = B.node()
            B.sync()
            t0 = time.perf_counter()
            B.lib.rc("hpaReadSerialized", z, path.encode(), C.byref(moved))
            B.sync()
            dt = time.perf_counter() - t0
            if r:                                   # the first read fills the page cache
                times.append(dt)
            free_node(B, z)
        best, med = min(times), float(np.median(times))
        cpu = None
        if not a.no_cpu:
            import oracle
            orc = oracle.backend(1)
            m = orc.mark()
            t0 = time.perf_counter()
            rc = orc.lib.dll.orc_readSerializeV(orc.node(), path.encode())
            dtc = time.perf_counter() - t0
            orc.release(m)
            if rc == 0:
                cpu = {"value": round(fbytes / dtc / 1e9, 3), "unit": "GB/s (file bytes)", "cores": 1, "kind": "port",
                       "sample": f"the whole {fbytes} B file, one read into host memory (readSerializeV restated)"}
    finally:
        os.remove(path)
    print(f"[bench] ingest {nf} rows, {fbytes / 1e9:.2f} GB file: read {med * 1e3:.1f} ms median "
          f"({fbytes / med / 1e9:.1f} GB/s), write {t_write:.2f}s", file=sys.stderr, flush=True)
    return {"metric": "rows/s + % HBM roofline per query; batched queries/s & p99 latency vs plain",
            "value": round(nf / med, 1), "unit": "rows/s", "n_gpus": 1, "steps": reps, "warmup": 1,
            "ms_per_step": round(med * 1e3, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "bytes", "data": "synthetic (the serving lineitem table, generated in HBM, written to a file)",
            "config": {"workload": f"readSerializeV of an 8-column lineitem table, {nf} rows, into HBM", "rows": nf,
                       "file_bytes": fbytes},
            "file_GBs": round(fbytes / med / 1e9, 2), "best_file_GBs": round(fbytes / best / 1e9, 2),
            "payload_bytes": moved.value, "write_s": round(t_write, 3), "cpu_baseline": cpu,
            "bound": "the file read (page cache) and the host-to-device link, not HBM"}


def bench_copy(a):
    import numpy as np
    from horsepower_amd import datagen as D
    from horsepower_amd.backend import GpuBackend
    from horsepower_amd.vnode import H_E, expr
    n = a.rows or 60_000_000
    B = GpuBackend()
    x = B.vec(D.cents(5, 1, n, 0, 10_000_000))
    e = expr([("COL", 0)])
    B.sync()
    mark = B.mark()
    for _ in range(a.warmup):
        B.fused_map(H_E, [x], e)
        B.release(mark)
    kern = profile_kernels(B.lib, lambda: (B.fused_map(H_E, [x], e), B.release(mark)), a.steps)
    k = kern["fused_map"]
    avg_ms = k["ms"] / k["count"]
    out = {"metric": "device copy (traffic-counter reference)", "value": round(16.0 * n / (avg_ms * 1e-3) / 1e9, 1),
           "unit": "GB/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(avg_ms, 4),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic", "config": {"workload": f"copy of {n} f64 (read {8 * n} B, write {8 * n} B per launch)",
                                            "rows": n}}
    return out


def cpu_baseline(a, W):
    """Oracle (C + OpenMP, all host cores) running the interpreter sequence on a sample."""
    import oracle
    from horsepower_amd import queries as Q
    threads = os.cpu_count() or 1
    ob = oracle.backend(threads)
    n = a.cpu_rows or W["cpu_default"]
    cols = W["gen"](n)
    tabs = tuple(Q.load_table(ob, c) for c in cols)
    rows = sum(len(next(iter(c.values()))) for c in cols) if not W["n"] else n
    plan = W["plans"]["plain"]
    mark = ob.mark()
    plan(ob, *tabs)                       # first run discarded (src/horseir/v1/compile-static/main.c:20-35)
    ob.release(mark)
    times = []
    budget = 20.0
    t_start = time.perf_counter()
    while len(times) < 3 and (time.perf_counter() - t_start) < budget:
        t0 = time.perf_counter()
        plan(ob, *tabs)
        times.append(time.perf_counter() - t0)
        ob.release(mark)
    mean = sum(times) / len(times)
    return {"value": rows / mean, "unit": "rows/s", "cores": ob.threads(), "kind": "port",
            "sample": f"{rows} rows of the same workload, mean of {len(times)} runs after one discarded",
            "ms_per_query": round(mean * 1e3, 2)}


if __name__ == "__main__":
    main()
