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

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload q1|q6|q3|batch] [--rows N]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--workload", default="q1", choices=["q1", "q6", "q3", "batch"])
    p.add_argument("--rows", type=int, default=0)
    p.add_argument("--plan", default="best", choices=["best", "fused", "plain"])
    p.add_argument("--cpu-rows", type=int, default=0)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--profile-steps", type=int, default=3)
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
    return dict(n=None, nd=nd, gen=lambda k: D.gen_q3(k), plans=dict(best=Q.q3_plain, fused=Q.q3_plain, plain=Q.q3_plain),
                desc=f"q3 join+group+top-10, {nd} orders", cpu_default=1_500_000)


def run_steps(B, plan, tables, steps, mark):
    for _ in range(steps):
        plan(B, *tables)
        B.release(mark)


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {a.gpus}); this build runs on a single "
              "MI355X and has no multi-GPU path", file=sys.stderr)
        sys.exit(2)
    import ctypes as C
    import numpy as np
    from horsepower_amd import queries as Q
    from horsepower_amd.backend import GpuBackend

    if a.workload == "batch":
        return bench_batch(a)
    W = workload(a.workload, a.rows)
    t0 = time.time()
    cols = W["gen"](W["n"] if W["n"] else W["nd"])
    print(f"[bench] generated {a.workload} in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    B = GpuBackend()
    lib = B.lib
    tables = tuple(Q.load_table(B, c) for c in cols)
    rows = W["n"] if W["n"] else sum(len(next(iter(c.values()))) for c in cols)
    del cols
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
    value = rows * a.steps / dt
    print(f"[bench] timed {a.steps} steps: {ms:.3f} ms/step", file=sys.stderr, flush=True)

    # kernel timing pass (events on every named launch)
    lib.dll.hpaProfileReset()
    lib.dll.hpaProfileEnable(1)
    run_steps(B, plan, tables, a.profile_steps, mark)
    B.sync()
    lib.dll.hpaProfileEnable(0)
    buf = C.create_string_buffer(1 << 16)
    lib.dll.hpaProfileList(buf, 1 << 16)
    kern = {}
    for line in buf.value.decode().splitlines():
        name, cnt, tms, byt = line.split()
        kern[name] = dict(count=int(cnt), ms=float(tms), bytes=float(byt))
    dom = max(kern, key=lambda k: kern[k]["ms"]) if kern else None
    roof = None
    if dom:
        k = kern[dom]
        avg_ms = k["ms"] / k["count"]
        per_launch = k["bytes"] / k["count"]
        achieved = per_launch / (avg_ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS,
                "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": None,
                "bytes_per_launch": per_launch, "avg_launch_us": round(avg_ms * 1e3, 2)}
    per_kernel = {k: {"launches_per_step": v["count"] / a.profile_steps,
                      "us_per_step": round(v["ms"] * 1e3 / a.profile_steps, 2),
                      "GBs": round(v["bytes"] / (v["ms"] * 1e-3) / 1e9, 1) if v["ms"] > 0 else None}
                  for k, v in sorted(kern.items(), key=lambda kv: -kv[1]["ms"])}

    # the interpreter's per-primitive sequence on the GPU, for comparison
    plain = None
    if a.plan != "plain":
        pl = W["plans"]["plain"]
        run_steps(B, pl, tables, 1, mark)
        B.sync()
        t1 = time.perf_counter()
        ps = max(2, a.steps // 4)
        run_steps(B, pl, tables, ps, mark)
        B.sync()
        pdt = (time.perf_counter() - t1) / ps
        plain = {"ms_per_step": round(pdt * 1e3, 3), "rows_per_s": rows / pdt}

    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(a, W)

    out = {
        "metric": "rows/s + % HBM roofline per query; batched queries/s & p99 latency vs plain",
        "value": value, "unit": "rows/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded counter-based generator)",
        "config": {"workload": W["desc"], "rows": rows, "plan": a.plan},
        "roofline": roof, "cpu_baseline": cpu, "plain_gpu": plain, "kernels": per_kernel,
    }
    print(json.dumps(out), flush=True)


def profile_kernels(lib, fn, steps):
    """Per-scope device time and algorithmic bytes from the library's HIP-event profiler."""
    import ctypes as C
    lib.dll.hpaProfileReset()
    lib.dll.hpaProfileEnable(1)
    for _ in range(steps):
        fn()
    lib.sync()
    lib.dll.hpaProfileEnable(0)
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


def bench_batch(a):
    import numpy as np
    from horsepower_amd import datagen as D, queries as Q
    from horsepower_amd.backend import Batch, GpuBackend
    nq = 64
    ns, gs, thr = D.batch_shapes(nq)
    t0 = time.time()
    cols = [D.gen_batch_query(q, int(ns[q]), int(gs[q])) for q in range(nq)]
    rows = int(ns.sum())
    print(f"[bench] generated {nq} queries, {rows} rows in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
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
        while k < nq and time.perf_counter() - t0 < 25.0:
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
        "value": nq / (dt / a.steps), "unit": "queries/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded counter-based generator)",
        "config": {"workload": f"batched group-by: {nq} queries in one persistent launch, {rows} rows, "
                               "row counts log-uniform 1e4..1e7, key ranges log-uniform 4..1e6",
                   "rows": rows, "queries": nq, "tuned": cfg},
        "rows_per_s": rows / (dt / a.steps),
        "latency_ms": {"p50": round(pct(lat, 50) * 1e3, 3), "p99": round(pct(lat, 99) * 1e3, 3)},
        "plain_gpu": {"queries_per_s": nq / sum(plain_t), "ms_per_batch": round(sum(plain_t) * 1e3, 3),
                      "query_p50_ms": round(pct(plain_t, 50) * 1e3, 3), "query_p99_ms": round(pct(plain_t, 99) * 1e3, 3)},
        "fused_gpu": {"queries_per_s": nq / sum(fused_t), "ms_per_batch": round(sum(fused_t) * 1e3, 3)},
        "equal_results": equal, "roofline": roof, "cpu_baseline": cpu,
        "kernels": {k: {"launches_per_step": v["count"] / a.profile_steps,
                        "us_per_step": round(v["ms"] * 1e3 / a.profile_steps, 2)} for k, v in kern.items()},
        "autotune_ms": [round(x, 4) for x in tune],
    }
    print(json.dumps(out), flush=True)
    b.close()


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
