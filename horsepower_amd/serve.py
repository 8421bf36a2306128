"""Serving: continuous batching of mixed queries over device-resident tables.

Requests (the Q6 filter-aggregate, Q1 group-by and Q3 join/group/top-k shapes, each with its
own predicate constants) arrive over time.  The server keeps the tables and the compiled fused
kernels resident (kernel constants are launch arguments, so new constants never recompile) and
runs admitted requests concurrently, each on its own context: a HIP stream plus an arena slice
carved from one device pool.

* Admission.  A request is admitted when a slot (context) is free and its intermediate-memory
  estimate fits in the pool (first fit).  Smaller requests may pass a larger one that does not
  fit yet, but not once the head of the queue has waited longer than `max_skip_s`.
* Continuous batching.  Requests that arrive while others run are admitted as slots and memory
  free up.  Waiting Q6 requests are merged (up to 8) into one shared scan
  (hpaFusedReduceMany): one pass over the fact columns answers all of them.  Waiting Q1
  requests with up to four distinct ship-date cutoffs are merged into one bucketed group-by
  pass (hpaFusedGroupAggBuckets, queries.q1_many).
* Overlap.  Each slot's worker thread issues its request's primitives on its own stream;
  ctypes releases the GIL inside every library call, so several requests are on the device
  at once.
* Estimates.  Each query kind is run once per extreme of its constants in calibration and its
  arena peak recorded; a request that still outgrows its slice (HPA_E_MEMORY) is requeued with
  twice the estimate.

`serve_serial` runs the same request stream one request at a time on the default context with
a chosen plan family (the per-query GPU path the server is compared with).
"""
import collections
import ctypes as C
import threading
import time

import numpy as np

from . import datagen as D
from . import queries as Q
from .vnode import HorseError

KINDS = ("q6", "q1", "q3")
E_MEMORY = 100
MiB = 1 << 20


class Request:
    __slots__ = ("rid", "kind", "params", "arrival", "admit", "done", "result", "est", "tries")

    def __init__(self, rid, kind, params, arrival):
        self.rid, self.kind, self.params, self.arrival = rid, kind, params, arrival
        self.admit = self.done = None
        self.result = None
        self.est = 0
        self.tries = 0

    def latency(self):
        return self.done - self.arrival


def make_requests(n, rate=None, seed=2024, mix=(0.45, 0.35, 0.20)):
    """n requests with seeded random constants; Poisson arrivals at `rate` per second
    (rate None: all arrive at time 0)."""
    u = D.rand_u64(seed, 41, n).astype(np.float64) / 2.0 ** 64
    gaps = -np.log1p(-u) / rate if rate else np.zeros(n)
    arrival = np.cumsum(gaps) - (gaps[0] if rate else 0.0)
    pick = D.rand_u64(seed, 42, n).astype(np.float64) / 2.0 ** 64
    lo = D.uniform(seed, 43, n, 0, 2100)
    dk = D.uniform(seed, 44, n, 2, 7)
    qmax = D.uniform(seed, 45, n, 20, 30)
    ship_max = D.uniform(seed, 46, n, 2300, 2500)
    day = D.uniform(seed, 47, n, 1000, 1300)
    seg = D.uniform(seed, 48, n, 0, 4)
    out = []
    c1, c2 = mix[0], mix[0] + mix[1]
    for i in range(n):
        if pick[i] < c1:
            p = dict(lo=int(lo[i]), hi=int(lo[i]) + 365, dlo=int(dk[i]) / 100.0,
                     dhi=(int(dk[i]) + 2) / 100.0, qmax=int(qmax[i]))
            out.append(Request(i, "q6", p, float(arrival[i])))
        elif pick[i] < c2:
            out.append(Request(i, "q1", dict(ship_max=int(ship_max[i])), float(arrival[i])))
        else:
            out.append(Request(i, "q3", dict(day=int(day[i]), segment=int(seg[i])), float(arrival[i])))
    return out


PLANS = {
    "fused": {"q6": lambda B, T, p: Q.q6_fused(B, T["fact"], **p),
              "q1": lambda B, T, p: Q.q1_chf(B, T["fact"], **p),
              "q3": lambda B, T, p: Q.q3_fused(B, T["dim"], T["fact"], **p)},
    "plain": {"q6": lambda B, T, p: Q.q6_plain(B, T["fact"], **p),
              "q1": lambda B, T, p: Q.q1_plain(B, T["fact"], **p),
              "q3": lambda B, T, p: Q.q3_plain(B, T["dim"], T["fact"], **p)},
}

# constants at the extremes of make_requests' ranges, for calibration
CALIBRATION = [
    ("q6", dict(lo=0, hi=365, dlo=0.02, dhi=0.04, qmax=30)),
    ("q6", dict(lo=2100, hi=2465, dlo=0.07, dhi=0.09, qmax=20)),
    ("q1", dict(ship_max=2300)), ("q1", dict(ship_max=2500)),
    ("q3", dict(day=1000, segment=0)), ("q3", dict(day=1300, segment=4)),
]


class FirstFit:
    """First-fit allocator of [0, size) in 1 MiB units (the device pool's slices)."""

    def __init__(self, size):
        self.size = size
        self.free = [(0, size)]

    def alloc(self, n):
        n = -(-n // MiB) * MiB
        for i, (o, l) in enumerate(self.free):
            if l >= n:
                self.free[i] = (o + n, l - n)
                if l == n:
                    del self.free[i]
                return o, n
        return None

    def release(self, o, n):
        self.free.append((o, n))
        self.free.sort()
        merged = []
        for a, l in self.free:
            if merged and merged[-1][0] + merged[-1][1] == a:
                merged[-1] = (merged[-1][0], merged[-1][1] + l)
            else:
                merged.append((a, l))
        self.free = merged

    def in_use(self):
        return self.size - sum(l for _, l in self.free)


class Server:
    def __init__(self, B, tables, slots=4, pool_bytes=64 << 30, merge=True, max_skip_s=0.05,
                 heavy=("q3",), heavy_slots=1, heavy_share=None, heavy_fifo=False, late_share=0):
        self.B, self.lib, self.T = B, B.lib, tables
        self.slots, self.merge, self.max_skip_s = slots, merge, max_skip_s
        self.base = self.lib.dll.hpaDeviceAlloc(pool_bytes)
        if not self.base:
            raise HorseError(E_MEMORY, "hpaDeviceAlloc", self.lib.detail())
        self.pool_bytes = pool_bytes
        # heavy requests (whole-table joins and groups) run at most heavy_slots at a time, in
        # arrival order: two of them sharing the device would each take about twice as long,
        # while short requests keep filling the other slots
        self.heavy, self.heavy_slots = tuple(heavy), heavy_slots
        # while a heavy request runs, at most heavy_share short requests run beside it (None: any
        # number): fewer neighbours take less of the memory bandwidth the heavy request needs
        self.heavy_share = heavy_share
        # heavy_fifo: a short request does not run beside a heavy one that arrived before it (as
        # in arrival order, it waits for that one to finish); shorts that arrived earlier still do
        self.heavy_fifo = heavy_fifo
        # with heavy_fifo: how many such later short jobs may still run beside the heavy ones
        self.late_share = late_share
        self.ctx = [self.lib.dll.hpaCtxCreate(-1) for _ in range(slots)]
        if not all(self.ctx):
            raise HorseError(E_MEMORY, "hpaCtxCreate", self.lib.detail())
        self.est = {}
        self.peak_in_use = 0

    def close(self):
        for c in self.ctx:
            self.lib.dll.hpaCtxDestroy(c)
        self.ctx = []
        if self.base:
            self.lib.dll.hpaDeviceFree(self.base)
            self.base = None

    # -- running one job on a slot --------------------------------------------------------------
    def _bind(self, slot, off, size):
        ctx = self.ctx[slot]
        self.lib.dll.hpaCtxSetCurrent(ctx)
        self.lib.rc("hpaCtxSetArena", ctx, C.c_void_p(self.base + off), size)

    def _execute(self, job):
        B, T = self.B, self.T
        if job[0].kind == "q1" and len(job) > 1:
            return Q.q1_many(B, T["fact"], [r.params for r in job])
        if len(job) > 1 or (job[0].kind == "q6" and self.merge):
            return Q.q6_many(B, T["fact"], [r.params for r in job])
        r = job[0]
        return [PLANS["fused"][r.kind](B, T, r.params)]

    def calibrate(self, cases=CALIBRATION, margin=1.25):
        """Arena peak of each query kind at the extreme constants; warms every fused kernel
        (including the shared-scan kernels for 1..8 merged requests)."""
        ctx = self.ctx[0]
        try:
            self._bind(0, 0, self.pool_bytes)
            for kind, p in cases:
                self.lib.rc("hpaCtxSetArena", ctx, C.c_void_p(self.base), self.pool_bytes)
                PLANS["fused"][kind](self.B, self.T, p)
                peak = self.lib.dll.hpaCtxArenaPeak(ctx)
                self.est[kind] = max(self.est.get(kind, 0), peak)
            for k in range(1, 9):
                self.lib.rc("hpaCtxSetArena", ctx, C.c_void_p(self.base), self.pool_bytes)
                Q.q6_many(self.B, self.T["fact"], [cases[0][1]] * k)
                self.est["q6"] = max(self.est["q6"], self.lib.dll.hpaCtxArenaPeak(ctx))
            if self.merge:
                # the shared Q1 passes for two to four cutoffs (their kernels compiled here too)
                for k in range(2, Q.Q1_MANY_MAX + 1):
                    self.lib.rc("hpaCtxSetArena", ctx, C.c_void_p(self.base), self.pool_bytes)
                    Q.q1_many(self.B, self.T["fact"], [dict(ship_max=2300 + 50 * j) for j in range(k)])
                    self.est["q1"] = max(self.est["q1"], self.lib.dll.hpaCtxArenaPeak(ctx))
        finally:
            self.lib.dll.hpaSync()
            self.lib.dll.hpaCtxSetCurrent(None)
        self.est = {k: int(v * margin) + 16 * MiB for k, v in self.est.items()}
        return dict(self.est)

    # -- the request loop -------------------------------------------------------------------------
    def serve(self, requests, realtime=True):
        """Run the requests (sorted by arrival); returns the same Request objects with
        admit/done times and results."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        for r in reqs:
            r.est = max(r.est, self.est.get(r.kind, 256 * MiB))
        pool = FirstFit(self.pool_bytes)
        cv = threading.Condition()
        free_slots = list(range(self.slots))
        pending = collections.deque()
        state = dict(done=0, error=None, heavy=0, short=0, late=0)
        heavy_running = []                                   # arrival times of running heavy requests
        N = len(reqs)
        nxt = 0
        t0 = time.perf_counter()

        def worker(slot, job, off, size, late):
            err = None
            try:
                self._bind(slot, off, size)
                res = self._execute(job)
                t = time.perf_counter() - t0
                for r, v in zip(job, res):
                    r.result, r.done = v, t
            except HorseError as e:
                err = e
            finally:
                self.lib.dll.hpaSync()                # the slice is reused only once its stream is idle
                self.lib.dll.hpaCtxSetCurrent(None)
            with cv:
                pool.release(off, size)
                free_slots.append(slot)
                if job[0].kind in self.heavy:
                    state["heavy"] -= 1
                    heavy_running.remove(job[0].arrival)
                else:
                    state["short"] -= 1
                    state["late"] -= late
                if err is None:
                    state["done"] += len(job)
                elif err.code == E_MEMORY:
                    for r in reversed(job):          # requeue at the front with a larger slice
                        r.est *= 2
                        r.tries += 1
                        if r.tries > 4:
                            state["error"] = err
                        pending.appendleft(r)
                else:
                    state["error"] = err
                cv.notify()

        threads = []
        with cv:
            while state["done"] < N and state["error"] is None:
                now = time.perf_counter() - t0
                while nxt < N and (not realtime or reqs[nxt].arrival <= now):
                    pending.append(reqs[nxt])
                    nxt += 1
                while free_slots and pending:
                    pick = None
                    head = None                              # oldest request the lane allows
                    for i, r in enumerate(pending):
                        if r.kind in self.heavy and state["heavy"] >= self.heavy_slots:
                            continue
                        if (r.kind not in self.heavy and self.heavy_share is not None and state["heavy"] > 0
                                and state["short"] >= self.heavy_share):
                            continue
                        if (r.kind not in self.heavy and self.heavy_fifo and heavy_running and min(heavy_running) < r.arrival
                                and state["late"] >= self.late_share):
                            continue
                        if head is not None and now - head.arrival > self.max_skip_s:
                            break                            # it waits for memory: no overtaking
                        got = pool.alloc(r.est)
                        if got is not None:
                            pick = (i, got)
                            break
                        if head is None:
                            head = r
                    if pick is None:
                        break
                    i, (off, size) = pick
                    r = pending[i]
                    del pending[i]
                    job = [r]
                    if self.merge and r.kind in ("q6", "q1"):
                        # Q6 requests share one scan (up to 8); Q1 requests one bucketed pass
                        # (up to Q1_MANY_MAX distinct cutoffs)
                        cap = 8 if r.kind == "q6" else Q.Q1_MANY_MAX
                        cuts = {r.params.get("ship_max")}
                        for q in list(pending):
                            if len(job) == cap:
                                break
                            if (self.heavy_fifo and heavy_running and min(heavy_running) < q.arrival
                                    and not (min(heavy_running) < r.arrival)):
                                continue                     # an on-time job takes no late request along
                            if q.kind != r.kind:
                                continue
                            if r.kind == "q1":
                                if len(cuts | {q.params.get("ship_max")}) > Q.Q1_MANY_MAX:
                                    continue
                                cuts.add(q.params.get("ship_max"))
                            pending.remove(q)
                            job.append(q)
                    for q in job:
                        q.admit = now
                    late = int(r.kind not in self.heavy and self.heavy_fifo and bool(heavy_running)
                               and min(heavy_running) < r.arrival)
                    if r.kind in self.heavy:
                        state["heavy"] += 1
                        heavy_running.append(r.arrival)
                    else:
                        state["short"] += 1
                        state["late"] += late
                    self.peak_in_use = max(self.peak_in_use, pool.in_use())
                    slot = free_slots.pop()
                    th = threading.Thread(target=worker, args=(slot, job, off, size, late), daemon=True)
                    threads.append(th)
                    th.start()
                if state["done"] >= N:
                    break
                wait = None
                if nxt < N and realtime:
                    wait = max(0.0, reqs[nxt].arrival - (time.perf_counter() - t0))
                cv.wait(timeout=min(wait, 0.05) if wait is not None else 0.05)
        for th in threads:
            th.join()
        if state["error"] is not None:
            raise state["error"]
        return reqs


def serve_serial(B, tables, requests, plans="fused", realtime=True):
    """The per-query GPU path: one request at a time, in arrival order, on the caller's
    context; each request's arena is released after it completes."""
    reqs = sorted(requests, key=lambda r: r.arrival)
    mark = B.mark()
    t0 = time.perf_counter()
    for r in reqs:
        now = time.perf_counter() - t0
        if realtime and r.arrival > now:
            time.sleep(r.arrival - now)
        r.admit = time.perf_counter() - t0
        r.result = PLANS[plans][r.kind](B, tables, r.params)
        r.done = time.perf_counter() - t0
        B.release(mark)
    return reqs


def summarize(reqs):
    lat = np.array([r.latency() for r in reqs])
    makespan = max(r.done for r in reqs) - min(r.arrival for r in reqs)
    by = {k: float(np.percentile([r.latency() for r in reqs if r.kind == k], 50)) * 1e3
          for k in KINDS if any(r.kind == k for r in reqs)}
    return {"requests": len(reqs), "queries_per_s": len(reqs) / makespan if makespan > 0 else None,
            "makespan_s": round(makespan, 4),
            "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
            "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3),
            "p50_ms_by_kind": {k: round(v, 3) for k, v in by.items()}}


def same_results(a, b, rtol=1e-12):
    """Request-by-request comparison of two runs' results (integers exact, FP64 within rtol)."""
    bad = []
    for x, y in zip(sorted(a, key=lambda r: r.rid), sorted(b, key=lambda r: r.rid)):
        for k in x.result:
            u, v = np.asarray(x.result[k]), np.asarray(y.result[k])
            if u.shape != v.shape:
                bad.append((x.rid, k))
            elif u.dtype.kind == "f":
                if not np.allclose(u, v, rtol=rtol, atol=0):
                    bad.append((x.rid, k))
            elif not np.array_equal(u, v):
                bad.append((x.rid, k))
    return bad
