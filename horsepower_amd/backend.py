"""A uniform driver over a pfn* primitive library, used by the query plans.

``Backend`` is written against the reference ABI only: it builds value nodes, calls primitives
by name and reads results back as numpy arrays.  ``GpuBackend`` drives libhorsepower_amd (HIP,
payloads resident in HBM).  Any other library exporting the same ABI can be driven by
subclassing ``Backend`` and supplying node allocation and payload reads.
"""
import ctypes as C

import numpy as np

from .runtime import Host, cell, load
from .vnode import H_B, H_E, H_G, H_I, H_L, H_N, NP_OF, V, V0, Expr

_DTYPES = {}


def _dtype_of(typ):
    dt = _DTYPES.get(typ)
    if dt is None:
        dt = _DTYPES[typ] = np.dtype(NP_OF[typ])
    return dt

FUNC_ADDR_NAMES = ("pfnSum", "pfnAvg", "pfnMin", "pfnMax", "pfnLen", "pfnIndex")


class Backend:
    name = "abstract"

    def __init__(self, lib):
        self.lib = lib
        self.host = Host(lib, self.node)
        self.fa = {n: lib.addr(n) for n in FUNC_ADDR_NAMES}

    # -- to be provided -----------------------------------------------------------------------
    def node(self):
        raise NotImplementedError

    def read_raw(self, z, nbytes, dst):
        raise NotImplementedError

    def symbol(self, name):
        raise NotImplementedError

    def resident(self, z):
        return z

    # -- construction ----------------------------------------------------------------------------
    def vec(self, arr, typ=None):
        return self.resident(self.host.vec(arr, typ))

    def scalar(self, v, typ):
        return self.host.scalar(v, typ)

    def sym(self, name):
        return self.host.sym(self.symbol(name))

    def list(self, items):
        return self.host.list(items)

    def strs(self, items):
        return self.resident(self.host.strs(items))

    def to_host(self, z):
        return z

    def read_strs(self, z):
        """String (H_S) node -> list of bytes."""
        self.to_host(z)
        return [C.string_at(int(p)) for p in self.read(z)]

    def bools(self, flags):
        return self.host.vec(np.array(flags, dtype=np.uint8), H_B)

    def call(self, name, *args):
        return self.lib.call(name, *args)

    def new(self, name, *args):
        z = self.node()
        self.lib.call(name, z, *args)
        return z

    # -- primitives by name ----------------------------------------------------------------------
    def __getattr__(self, name):
        if name.startswith("pfn"):
            return lambda *a: self.new(name, *a)
        raise AttributeError(name)

    def each(self, x, fname):
        z = self.node()
        self.lib.call("pfnEach", z, x, self.fa[fname])
        return z

    def each_right(self, x, y, fname):
        z = self.node()
        self.lib.call("pfnEachRight", z, x, y, self.fa[fname])
        return z

    # -- reading -------------------------------------------------------------------------------
    def read(self, z):
        """Vector/scalar node -> numpy array."""
        v = z.contents
        if v.typ == H_G:
            raise TypeError("list node: use read_list / read_groups")
        dt = np.dtype(NP_OF[v.typ])
        n = v.len
        out = np.empty(n, dtype=dt)
        if n == 0:
            return out
        if n == 1 and v.g == C.addressof(v) + 24:
            C.memmove(out.ctypes.data, v.g, dt.itemsize)
            return out
        self.read_raw(z, n * dt.itemsize, out.ctypes.data)
        return out

    def read_scalar(self, z):
        return self.read(z)[0]

    def read_many(self, zs):
        return [self.read(z) for z in zs]

    def read_list(self, z):
        v = z.contents
        if v.typ not in (H_G, H_N):
            raise TypeError("not a list")
        if v.typ == H_G and v.u.p:                          # flat list
            return self.read_flat(z)
        n = v.len if v.typ == H_G else 2
        return [self.read_any(cell(z, k)) for k in range(n)]

    def read_any(self, z):
        return self.read_list(z) if z.contents.typ in (H_G, H_N) else self.read(z)

    def read_flat(self, z):
        """Flat list -> list of arrays split by the group sizes."""
        v = z.contents
        counts = np.empty(v.len, dtype=np.int64)
        if v.len:
            self.read_raw(z, 8 * v.len, counts.ctypes.data)
        inner = C.cast(C.c_void_p(v.u.p), V)
        flat = self.read(inner)
        return np.split(flat, np.cumsum(counts)[:-1]) if v.len else []

    def read_groups(self, z):
        """Group dict -> (keys, counts, flat) with flat concatenating the group index lists."""
        keys = self.read(cell(z, 0))
        vals = cell(z, 1)
        v = vals.contents
        if v.u.p:
            counts = np.empty(v.len, dtype=np.int64)
            if v.len:
                self.read_raw(vals, 8 * v.len, counts.ctypes.data)
            flat = self.read(C.cast(C.c_void_p(v.u.p), V))
        else:
            parts = [self.read(cell(vals, k)) for k in range(v.len)]
            counts = np.array([len(p) for p in parts], dtype=np.int64)
            flat = np.concatenate(parts) if parts else np.zeros(0, np.int64)
        return keys, counts, flat


class GpuBackend(Backend):
    """libhorsepower_amd on the current HIP device."""
    name = "gpu"

    def __init__(self, lib=None):
        super().__init__(lib or load())
        self.dev = self.lib
        self._aggs = {}
        self._fga_args = {}

    def node(self):
        return self.lib.node()

    def resident(self, z):
        self.lib.rc("hpaUpload", z)
        return z

    def read_raw(self, z, nbytes, dst):
        self.lib.rc("hpaReadPayload", z, C.c_void_p(dst), nbytes)

    def to_host(self, z):
        self.lib.rc("hpaToHost", z)
        return z

    def read_many(self, zs):
        """Vector nodes -> numpy arrays with one device synchronisation.  The arrays are views of
        one host buffer (16-byte aligned slices), so the call costs one allocation."""
        n = len(zs)
        cs = [z.contents for z in zs]
        dts = [_dtype_of(v.typ) for v in cs]
        sizes = [v.len * dt.itemsize for v, dt in zip(cs, dts)]
        offs, tot = [], 0
        for sz in sizes:
            offs.append(tot)
            tot += (sz + 15) & ~15
        buf = np.empty(max(tot, 16), dtype=np.uint8)
        base = buf.__array_interface__["data"][0]
        dst = (C.c_void_p * n)(*[base + o for o in offs])
        nb = (C.c_int64 * n)(*sizes)
        self.lib.rc("hpaReadMany", n, (V * n)(*zs), dst, nb)
        return [buf[o:o + sz].view(dt) for o, sz, dt in zip(offs, sizes, dts)]

    def symbol(self, name):
        return self.lib.symbol(name)

    def mark(self):
        return self.lib.mark()

    def release(self, m):
        self.lib.release(m)

    def sync(self):
        self.lib.sync()

    # fused entry points (GPU only)
    def fused_compress(self, ys, xs, cond):
        n = len(ys)
        zs = [self.node() for _ in range(n)]
        za = (V * n)(*zs)
        ya = (V * n)(*ys)
        xa = (V * len(xs))(*xs)
        self.lib.rc("hpaFusedCompress", n, za, ya, len(xs), xa, C.byref(cond))
        return zs

    def compress_many(self, mask, ys):
        n = len(ys)
        zs = [self.node() for _ in range(n)]
        self.lib.rc("hpaCompressMany", n, (V * n)(*zs), mask, (V * n)(*ys))
        return zs

    def compress_map(self, mask, ys, maps, xs):
        """Compress of plain columns ys and of FE maps [(typ, expr)] over columns xs, one pass."""
        n, nm = len(ys), len(maps)
        zs = [self.node() for _ in range(n)]
        zm = [self.node() for _ in range(nm)]
        typs = (C.c_int32 * max(nm, 1))(*[t for t, _ in maps])
        exprs = (Expr * max(nm, 1))(*[e for _, e in maps])
        self.lib.rc("hpaCompressMap", n, (V * max(n, 1))(*zs), mask, (V * max(n, 1))(*ys), nm,
                    (V * max(nm, 1))(*zm), typs, exprs, len(xs), (V * max(len(xs), 1))(*xs))
        return zs, zm

    def fused_compress_map(self, cond, ys, maps, xs):
        """hpaFusedCompressMap: compress of plain columns ys and of FE maps [(typ, expr)] over
        columns xs by the condition cond (an expression over xs), one count and one write pass."""
        n, nm = len(ys), len(maps)
        zs = [self.node() for _ in range(n)]
        zm = [self.node() for _ in range(nm)]
        typs = (C.c_int32 * max(nm, 1))(*[t for t, _ in maps])
        exprs = (Expr * max(nm, 1))(*[e for _, e in maps])
        self.lib.rc("hpaFusedCompressMap", n, (V * max(n, 1))(*zs), (V * max(n, 1))(*ys), nm, (V * max(nm, 1))(*zm),
                    typs, exprs, len(xs), (V * max(len(xs), 1))(*xs), C.byref(cond))
        return zs, zm

    def index_many(self, xs, y):
        """Peephole index: [index(x, y) for x in xs] in one launch."""
        n = len(xs)
        zs = [self.node() for _ in range(n)]
        self.lib.rc("hpaIndexMany", n, (V * n)(*zs), (V * n)(*xs), y)
        return zs

    def fused_reduce(self, kind, xs, cond, body):
        z = self.node()
        xa = (V * len(xs))(*xs)
        self.lib.rc("hpaFusedReduce", z, kind, len(xs), xa,
                    C.byref(cond) if cond is not None else None, C.byref(body))
        return z

    def fused_reduce_many(self, kind, xs, conds, body):
        """Shared scan: one reduction of `body` per condition, one pass over the columns."""
        n = len(conds)
        zs = [self.node() for _ in range(n)]
        self.lib.rc("hpaFusedReduceMany", n, (V * n)(*zs), kind, len(xs), (V * len(xs))(*xs),
                    (Expr * n)(*conds), C.byref(body))
        return zs

    def gen_serve_tables(self, n_dim, seed):
        """Serving tables generated in HBM: (dim, fact) dicts of resident column nodes."""
        from .datagen import SERVE_DIM, SERVE_FACT
        dim = [C.pointer(V0()) for _ in SERVE_DIM]         # Python-owned nodes: outlive arena resets
        fact = [C.pointer(V0()) for _ in SERVE_FACT]
        self.lib.rc("hpaGenServeTables", n_dim, seed, (V * 3)(*dim), (V * 8)(*fact))
        return dict(zip(SERVE_DIM, dim)), dict(zip(SERVE_FACT, fact))

    def fused_map(self, typ, xs, e):
        z = self.node()
        xa = (V * len(xs))(*xs)
        self.lib.rc("hpaFusedMap", z, typ, len(xs), xa, C.byref(e))
        return z

    def fused_group_agg(self, keys, xs, cond, spec):
        """One-pass filter + group + aggregates.  Returns (first, [key vectors], [aggregates]) or
        None when the key cardinality per workgroup exceeds the fused kernel's slots."""
        from .vnode import HorseError, aggs
        nk, na = len(keys), len(spec)
        zk = [self.node() for _ in range(nk)]
        za = [self.node() for _ in range(na)]
        zf = self.node()
        hit = self._aggs.get(id(spec))               # descriptor arrays of a spec seen before
        if hit is None or hit[0] is not spec:
            hit = (spec, aggs(spec))
            self._aggs[id(spec)] = hit
        ag = hit[1]
        rc = self.lib.dll.hpaFusedGroupAgg(nk, (V * nk)(*keys), len(xs), (V * len(xs))(*xs),
                                           C.byref(cond) if cond is not None else None, na,
                                           C.cast(ag, C.c_void_p), zf, (V * nk)(*zk), (V * na)(*za))
        if rc == 99:
            return None
        if rc != 0:
            raise HorseError(rc, "hpaFusedGroupAgg", self.lib.detail())
        return zf, zk, za

    def fused_group_agg_host(self, keys, xs, cond, spec, cap=64):
        """fused_group_agg with its results read back in the same synchronisation
        (hpaFusedGroupAggHost): ([key arrays], [aggregate arrays]) as numpy arrays, or None when
        the key cardinality per workgroup exceeds the fused kernel's slots.  More than `cap`
        groups are read from the result nodes instead."""
        from .vnode import HorseError, aggs
        nk, na = len(keys), len(spec)
        # the call's argument arrays are built once per (columns, spec, cap) and reused: the
        # per-call foreign-function work is one node allocation and the call itself
        ck = (tuple(id(k) for k in keys), tuple(id(x) for x in xs), id(spec), id(cond), cap)
        hit = self._fga_args.get(ck)
        if hit is None or hit[0] is not spec or hit[1] is not cond:
            buf_bytes = cap * (nk + 8 * na) + 16
            offs = [k * cap for k in range(nk)] + [((nk * cap + 7) & ~7) + 8 * cap * a for a in range(na)]
            hit = (spec, cond, list(keys), list(xs), aggs(spec), (V * nk)(*keys), (V * len(xs))(*xs),
                   (V * (nk + na))(), buf_bytes, offs, C.c_int64(0), C.c_int64(cap),
                   C.byref(cond) if cond is not None else None)
            if len(self._fga_args) >= 256:
                self._fga_args.clear()
            self._fga_args[ck] = hit
        (_, _, _, _, agg_arr, karr, xarr, zarr, buf_bytes, offs, ng, capv, condp) = hit
        self.lib.rc("hpaAllocNodes", nk + na, zarr)
        buf = np.empty(buf_bytes, dtype=np.uint8)
        base = buf.ctypes.data
        dst = (C.c_void_p * (nk + na))(*[base + o for o in offs])
        rc = self.lib.dll.hpaFusedGroupAggHost(nk, karr, len(xs), xarr, condp, na, C.cast(agg_arr, C.c_void_p), None,
                                               zarr, C.cast(C.addressof(zarr) + nk * C.sizeof(V), C.POINTER(V)),
                                               capv, dst, None, C.cast(C.addressof(dst) + nk * C.sizeof(C.c_void_p), C.POINTER(C.c_void_p)),
                                               C.byref(ng))
        zk, za = list(zarr[:nk]), list(zarr[nk:])
        if rc == 99:
            return None
        if rc != 0:
            raise HorseError(rc, "hpaFusedGroupAggHost", self.lib.detail())
        g = ng.value
        if g > cap:
            out = self.read_many(zk + za)
            return out[:nk], out[nk:]
        out = []
        for z, o in zip(zk + za, offs):
            dt = _dtype_of(z.contents.typ)
            out.append(buf[o:o + g * dt.itemsize].view(dt))
        return out[:nk], out[nk:]

    def fused_group_agg_buckets(self, keys, xs, cond, bucket, bbits, spec, cap=64):
        """hpaFusedGroupAggBuckets: groups (bucket, keys) with the bucket an expression over xs
        below 2^bbits; returns ([key arrays], bucket array, [aggregate arrays]) in (bucket, keys)
        order, read back with one synchronisation (at most cap groups)."""
        from .vnode import HorseError, aggs
        nk, na = len(keys), len(spec)
        zarr = (V * (nk + na))()
        self.lib.rc("hpaAllocNodes", nk + na, zarr)
        buf = np.empty(cap * (nk + 1 + 8 * na) + 16, dtype=np.uint8)
        base = buf.ctypes.data
        offs = [k * cap for k in range(nk + 1)] + [(((nk + 1) * cap + 7) & ~7) + 8 * cap * a for a in range(na)]
        ng = C.c_int64(0)
        agg_arr = aggs(spec)
        rc = self.lib.dll.hpaFusedGroupAggBuckets(
            nk, (V * nk)(*keys), len(xs), (V * len(xs))(*xs), C.byref(cond) if cond is not None else None, C.byref(bucket), bbits,
            na, C.cast(agg_arr, C.c_void_p), zarr, C.cast(C.addressof(zarr) + nk * C.sizeof(V), C.POINTER(V)), C.c_int64(cap),
            (C.c_void_p * nk)(*[base + o for o in offs[:nk]]), C.c_void_p(base + offs[nk]),
            (C.c_void_p * na)(*[base + o for o in offs[nk + 1:]]), C.byref(ng))
        if rc == 99:
            return None
        if rc != 0:
            raise HorseError(rc, "hpaFusedGroupAggBuckets", self.lib.detail())
        g = ng.value
        if g > cap:
            raise HorseError(99, "hpaFusedGroupAggBuckets", f"{g} groups above the {cap} read back")
        zk, za = list(zarr[:nk]), list(zarr[nk:])
        ks = [buf[o:o + g].view(_dtype_of(z.contents.typ)) for z, o in zip(zk, offs[:nk])]
        bk = buf[offs[nk]:offs[nk] + g].copy()
        ag = [buf[o:o + g * 8].view(_dtype_of(z.contents.typ)) for z, o in zip(za, offs[nk + 1:])]
        return ks, bk, ag

    def group_reduce(self, kind, x, groups):
        z = self.node()
        self.lib.rc("hpaGroupReduce", z, kind, x, groups)
        return z


class Batch:
    """A resident plan over many filter -> group -> sum queries (hpaBatch*), run in one
    persistent launch.  tables: dicts with "key" (I), "val" (E), "flt" (I) column nodes;
    rows with flt < thr[q] take part in query q."""

    def __init__(self, B, tables, thr):
        from .vnode import OP, HorseError
        self.B, self.lib, self.nq = B, B.lib, len(tables)
        nq = self.nq
        cols = [(V * nq)(*[t[k] for t in tables]) for k in ("key", "val", "flt")]
        ops = (C.c_int32 * nq)(*([OP["LT"]] * nq))
        consts = (C.c_int32 * nq)(*[int(x) for x in thr])
        self.h = self.lib.dll.hpaBatchCreate(nq, cols[0], cols[1], cols[2], ops, consts)
        if not self.h:
            raise HorseError(-1, "hpaBatchCreate", self.lib.detail())

    def run(self, graph=True):
        self.lib.rc("hpaBatchRun", self.h, 1 if graph else 0)

    def results(self, fn="hpaBatchFetch"):
        """[{key, count, sum}] per query: numpy views of the plan's pinned staging, filled with
        one synchronisation (valid until the next call)."""
        nq = self.nq
        ng = (C.c_int64 * nq)()
        k, c, s = (C.c_void_p * nq)(), (C.c_void_p * nq)(), (C.c_void_p * nq)()
        self.lib.rc(fn, self.h, ng, k, c, s)
        n = np.frombuffer(ng, dtype=np.int64)
        off = np.zeros(nq + 1, dtype=np.int64)
        np.cumsum(n, out=off[1:])
        tot = int(off[-1])

        def column(ptrs, dt):
            # the queries' results lie one after another in one array: one view over all of it,
            # sliced per query (a view per query costs several microseconds each)
            p = np.frombuffer(ptrs, dtype=np.uint64)
            if tot == 0:
                return [np.zeros(0, dt)] * nq
            size = np.dtype(dt).itemsize
            if not np.array_equal(p, p[0] + off[:-1].astype(np.uint64) * size):
                return [np.zeros(0, dt) if n[q] == 0 else
                        np.frombuffer((C.c_char * (int(n[q]) * size)).from_address(int(p[q])), dtype=dt) for q in range(nq)]
            whole = np.frombuffer((C.c_char * (tot * size)).from_address(int(p[0])), dtype=dt)
            return [whole[off[q]:off[q + 1]] for q in range(nq)]
        ks, cs, ss = column(k, np.int32), column(c, np.int64), column(s, np.float64)
        return [dict(key=ks[q], count=cs[q], sum=ss[q]) for q in range(nq)]

    def fetch_begin(self):
        """Start copying the last run's results to the host; the next run may start at once."""
        self.lib.rc("hpaBatchFetchBegin", self.h)

    def fetch_end(self):
        """The results of the oldest begun fetch (valid until two more fetch_begin calls)."""
        return self.results("hpaBatchFetchEnd")

    def result_nodes(self, q):
        """Query q's results as value nodes (device payloads) for further primitives."""
        z = [self.B.node() for _ in range(3)]
        self.lib.rc("hpaBatchResult", self.h, q, *z)
        return z

    def config(self):
        v = [C.c_int32() for _ in range(4)]
        self.lib.rc("hpaBatchGetConfig", self.h, *[C.byref(x) for x in v])
        return dict(block=v[0].value, tile_rows=v[1].value, lds_keys=v[2].value, part_keys=v[3].value)

    def set_config(self, block, tile_rows, lds_keys, part_keys=2048):
        self.lib.rc("hpaBatchSetConfig", self.h, block, tile_rows, lds_keys, part_keys)

    def autotune(self, iters=3):
        ms = (C.c_float * 64)()
        n = C.c_int32()
        self.lib.rc("hpaBatchAutotune", self.h, iters, ms, 64, C.byref(n))
        return list(ms[:n.value])

    def close(self):
        if self.h:
            self.lib.dll.hpaBatchDestroy(self.h)
            self.h = None
