"""Loader of libhorsepower_amd.so and the device-side value-node helpers.

The HIP library is the only compute path: ``load()`` raises when the built library is
missing, and nothing here falls back to host computation.
"""
import ctypes as C
import os

import numpy as np

from .vnode import (V0, V, H_B, H_S, H_C, H_G, H_L, H_N, H_Q, NP_OF, TYP_OF, Expr, HorseError,
                    PrimitiveLib)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhorsepower_amd.so")

_lib = None


def load(init=True, device=0, arena_bytes=0):
    """Load (once) and initialise the HIP primitive library."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run `python -m horsepower_amd.build`")
        _lib = Device(C.CDLL(LIB_PATH))
    if init:
        _lib.init(device, arena_bytes)
    return _lib


def _bind(dll, name, res, args):
    f = getattr(dll, name)
    f.restype = res
    f.argtypes = args
    return f


class Device(PrimitiveLib):
    """The HIP library: pfn* primitives plus runtime, residency and fused entry points."""

    def __init__(self, dll):
        super().__init__(dll, "")
        i32, i64, vp = C.c_int32, C.c_int64, C.c_void_p
        self._init = _bind(dll, "hpaInit", i32, [i32, i64])
        _bind(dll, "hpaShutdown", None, [])
        _bind(dll, "hpaErrorString", C.c_char_p, [i32])
        _bind(dll, "hpaLastErrorDetail", C.c_char_p, [])
        _bind(dll, "hpaAllocNode", V, [])
        _bind(dll, "hpaAllocNodes", i32, [i32, C.POINTER(V)])
        _bind(dll, "hpaGetHeapOffset", i64, [])
        _bind(dll, "hpaSetHeapOffset", i64, [i64])
        _bind(dll, "hpaInitV", i32, [V, i32, i64])
        _bind(dll, "hpaUpload", i32, [V])
        _bind(dll, "hpaFreeResident", i32, [V])
        _bind(dll, "hpaToHost", i32, [V])
        _bind(dll, "hpaReadPayload", i32, [V, vp, i64])
        _bind(dll, "hpaIsDevicePtr", i32, [vp])
        _bind(dll, "hpaReadMany", i32, [i32, C.POINTER(V), C.POINTER(vp), C.POINTER(i64)])
        _bind(dll, "hpaSync", i32, [])
        _bind(dll, "hpaCtxCreate", vp, [i64])
        _bind(dll, "hpaCtxDestroy", None, [vp])
        _bind(dll, "hpaDeviceAlloc", vp, [i64])
        _bind(dll, "hpaDeviceFree", i32, [vp])
        _bind(dll, "hpaCtxSetArena", i32, [vp, vp, i64])
        _bind(dll, "hpaCtxDefault", vp, [])
        _bind(dll, "hpaCtxSetCurrent", None, [vp])
        _bind(dll, "hpaCtxCurrent", vp, [])
        _bind(dll, "hpaCtxStream", vp, [vp])
        _bind(dll, "hpaCtxArenaBytes", i64, [vp])
        _bind(dll, "hpaCtxArenaPeak", i64, [vp])
        _bind(dll, "hpaCtxResetPeak", None, [vp])
        _bind(dll, "hpaProfileEnable", None, [i32])
        _bind(dll, "hpaProfileReset", None, [])
        _bind(dll, "hpaProfileStats", i32, [C.c_char_p, C.POINTER(i64), C.POINTER(C.c_double)])
        _bind(dll, "hpaProfileList", i32, [C.c_char_p, i64])
        _bind(dll, "hpaProfileBegin", None, [C.c_char_p, C.c_double])
        _bind(dll, "hpaProfileEnd", None, [])
        _bind(dll, "hpaSymbol", i64, [C.c_char_p])
        _bind(dll, "hpaSymbolName", C.c_char_p, [i64])
        _bind(dll, "hpaImportSymbols", i32, [i64, C.POINTER(C.c_char_p)])
        _bind(dll, "hpaReadSerialized", i32, [V, C.c_char_p, C.POINTER(i64)])
        _bind(dll, "hpaInitTableFromBin", i32, [C.c_char_p, C.c_char_p, C.POINTER(V)])
        _bind(dll, "hpaWriteSerialized", i32, [V, C.c_char_p])
        _bind(dll, "hpaFusedMap", i32, [V, i32, i32, C.POINTER(V), C.POINTER(Expr)])
        _bind(dll, "hpaFusedReduce", i32, [V, i32, i32, C.POINTER(V), C.POINTER(Expr), C.POINTER(Expr)])
        _bind(dll, "hpaFusedReduceMany", i32, [i32, C.POINTER(V), i32, i32, C.POINTER(V), C.POINTER(Expr), C.POINTER(Expr)])
        _bind(dll, "hpaFusedCompress", i32, [i32, C.POINTER(V), C.POINTER(V), i32, C.POINTER(V), C.POINTER(Expr)])
        _bind(dll, "hpaCompressMany", i32, [i32, C.POINTER(V), V, C.POINTER(V)])
        _bind(dll, "hpaCompressMap", i32, [i32, C.POINTER(V), V, C.POINTER(V), i32, C.POINTER(V),
                                           C.POINTER(C.c_int32), C.POINTER(Expr), i32, C.POINTER(V)])
        _bind(dll, "hpaFusedCompressMap", i32, [i32, C.POINTER(V), C.POINTER(V), i32, C.POINTER(V), C.POINTER(i32),
                                              C.POINTER(Expr), i32, C.POINTER(V), C.POINTER(Expr)])
        _bind(dll, "hpaIndexMany", i32, [i32, C.POINTER(V), C.POINTER(V), V])
        _bind(dll, "hpaGroupReduce", i32, [V, i32, V, V])
        _bind(dll, "hpaFusedGroupAgg", i32, [i32, C.POINTER(V), i32, C.POINTER(V), C.POINTER(Expr), i32,
                                             vp, V, C.POINTER(V), C.POINTER(V)])
        _bind(dll, "hpaFusedGroupAggHost", i32, [i32, C.POINTER(V), i32, C.POINTER(V), C.POINTER(Expr), i32,
                                                 vp, V, C.POINTER(V), C.POINTER(V), i64, C.POINTER(vp),
                                                 C.POINTER(i64), C.POINTER(vp), C.POINTER(i64)])
        _bind(dll, "hpaFusedGroupAggBuckets", i32, [i32, C.POINTER(V), i32, C.POINTER(V), C.POINTER(Expr), C.POINTER(Expr), i32,
                                                    i32, vp, C.POINTER(V), C.POINTER(V), i64, C.POINTER(vp), vp,
                                                    C.POINTER(vp), C.POINTER(i64)])
        _bind(dll, "hpaJitCompileCheck", i32, [i32, C.POINTER(i32), C.POINTER(Expr), i32, vp])
        _bind(dll, "hpaBatchCreate", vp, [i32, C.POINTER(V), C.POINTER(V), C.POINTER(V),
                                          C.POINTER(i32), C.POINTER(i32)])
        _bind(dll, "hpaBatchDestroy", None, [vp])
        _bind(dll, "hpaGenUniform", i32, [V, i32, i64, i64, C.c_uint64, C.c_uint64, i64, i64])
        _bind(dll, "hpaGenServeTables", i32, [i64, C.c_uint64, C.POINTER(V), C.POINTER(V)])
        _bind(dll, "hpaBatchRun", i32, [vp, i32])
        _bind(dll, "hpaBatchResult", i32, [vp, i32, V, V, V])
        _bind(dll, "hpaBatchFetch", i32, [vp, C.POINTER(i64), C.POINTER(vp), C.POINTER(vp), C.POINTER(vp)])
        _bind(dll, "hpaBatchFetchBegin", i32, [vp])
        _bind(dll, "hpaBatchFetchEnd", i32, [vp, C.POINTER(i64), C.POINTER(vp), C.POINTER(vp), C.POINTER(vp)])
        _bind(dll, "hpaBatchSetConfig", i32, [vp, i32, i32, i32, i32])
        _bind(dll, "hpaBatchGetConfig", i32, [vp, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32), C.POINTER(i32)])
        _bind(dll, "hpaBatchAutotune", i32, [vp, i32, C.POINTER(C.c_float), i32, C.POINTER(i32)])
        self.initialised = False
        self._keep = []

    def init(self, device=0, arena_bytes=0):
        if not self.initialised:
            rc = self._init(device, arena_bytes)
            if rc != 0:
                raise HorseError(rc, "hpaInit", self.detail())
            self.initialised = True

    def detail(self):
        d = self.dll.hpaLastErrorDetail()
        return d.decode() if d else ""

    def call(self, name, *args):
        rc = self.fn(name)(*args)
        if rc != 0:
            raise HorseError(rc, name, self.detail())
        return rc

    def rc(self, name, *args):
        rc = getattr(self.dll, name)(*args)
        if rc != 0:
            raise HorseError(rc, name, self.detail())
        return rc

    # -- nodes -------------------------------------------------------------------------------
    def node(self):
        return self.dll.hpaAllocNode()

    def mark(self):
        return self.dll.hpaGetHeapOffset()

    def release(self, mark):
        self.dll.hpaSetHeapOffset(mark)

    def symbol(self, name):
        return self.dll.hpaSymbol(name.encode())

    def sync(self):
        self.rc("hpaSync")


class Host:
    """Host-side construction and reading of value nodes for a library binding (`lib`)."""

    def __init__(self, lib, node_fn):
        self.lib = lib
        self._node = node_fn
        self._keep = []

    def node(self):
        return self._node()

    def vec(self, arr, typ=None):
        """A node whose payload is the numpy array (kept alive by this object)."""
        a = np.ascontiguousarray(arr)
        if typ is None:
            typ = TYP_OF[a.dtype]
        a = a.astype(NP_OF[typ], copy=False)
        z = self.node()
        z.contents.typ = typ
        z.contents.len = a.shape[0]
        if a.shape[0] <= 1 and typ not in (H_G, H_N):
            raw = np.zeros(8, dtype=np.uint8)
            if a.shape[0] == 1:
                raw[: a.itemsize] = a.view(np.uint8)
            z.contents.u.i64 = int(raw.view(np.int64)[0])
            z.contents.g = C.addressof(z.contents) + 24
        else:
            if typ == H_C:                                 # room for the NUL getTypeSize counts
                b = np.zeros(a.shape[0] + 1, np.uint8)
                b[:-1] = a
                a = b
            self._keep.append(a)
            z.contents.g = a.ctypes.data
        return z

    def scalar(self, v, typ):
        return self.vec(np.array([v], dtype=NP_OF[typ]), typ)

    def strs(self, items):
        """A string (H_S) node: char* values into one NUL-separated buffer kept alive here."""
        bs = [s.encode() if isinstance(s, str) else bytes(s) for s in items]
        buf = np.frombuffer(b"".join(b + b"\0" for b in bs) or b"\0", dtype=np.uint8).copy()
        offs = np.cumsum([0] + [len(b) + 1 for b in bs[:-1]]).astype(np.uint64) if bs else np.zeros(0, np.uint64)
        self._keep.append(buf)
        return self.vec(offs + np.uint64(buf.ctypes.data), H_S)

    def sym(self, symid):
        z = self.node()
        z.contents.typ = H_Q
        z.contents.len = 1
        z.contents.u.i64 = symid
        z.contents.g = C.addressof(z.contents) + 24
        return z

    def list(self, items):
        z = self.node()
        arr = (V * len(items))(*items)
        self.lib.call("pfnList", z, len(items), arr)
        return z


def cell(z, k):
    """k-th cell of a list/dict node (cells are host V0 arrays)."""
    return C.cast(C.c_void_p(z.contents.g + 32 * k), V)
