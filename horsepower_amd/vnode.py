"""ctypes mirror of the HorseIR value node and a typed primitive-library binding.

The node layout is the reference's V0 (src/horseir/v3/backend/common.h:31-55): 32 bytes,
``len`` (i64) at 0, ``g`` (payload pointer) at 8, ``typ`` (i32) at 16, an 8-byte union at 24.
``PrimitiveLib`` binds any shared library exporting the ``pfn*`` signatures
(src/horseir/v3/backend/h_primitive.h:72-211), optionally under a name prefix.
"""
import ctypes as C

import numpy as np

# type codes: src/horseir/v3/backend/h_type.h:1-23
H_B, H_J, H_H, H_I, H_L, H_F, H_E, H_X, H_C, H_Q, H_S = range(11)
H_M, H_D, H_Z, H_U, H_W, H_T, H_G, H_N, H_Y = range(11, 20)
H_A, H_K, H_V = 90, 91, 93

NP_OF = {
    H_B: np.uint8, H_J: np.int8, H_H: np.int16, H_I: np.int32, H_L: np.int64,
    H_F: np.float32, H_E: np.float64, H_C: np.uint8, H_Q: np.int64, H_M: np.int32,
    H_D: np.int32, H_Z: np.int64, H_U: np.int32, H_W: np.int32, H_T: np.int32,
    H_S: np.uint64,                                   # char* values (host addresses)
}
TYP_OF = {np.dtype(np.uint8): H_B, np.dtype(np.bool_): H_B, np.dtype(np.int8): H_J,
          np.dtype(np.int16): H_H, np.dtype(np.int32): H_I, np.dtype(np.int64): H_L,
          np.dtype(np.float32): H_F, np.dtype(np.float64): H_E}

# error codes: src/horseir/v3/backend/h_system.h:128-142
E_DOMAIN, E_GENERAL, E_INDEX, E_DIV_ZERO, E_LENGTH, E_TYPE = 1, 2, 3, 4, 5, 6
E_MATCH, E_RAZE_LIST, E_NOT_IMPL, E_MEMORY, E_DEVICE = 9, 11, 99, 100, 101

OP = dict(COL=1, CONST_I=2, CONST_F=3, ADD=4, SUB=5, MUL=6, DIV=7, LT=8, LEQ=9, GT=10,
          GEQ=11, EQ=12, NEQ=13, AND=14, OR=15, NOT=16, NEG=17)


class _U(C.Union):
    _fields_ = [("b", C.c_uint8), ("i32", C.c_int32), ("i64", C.c_int64), ("f32", C.c_float),
                ("f64", C.c_double), ("p", C.c_void_p)]


class V0(C.Structure):
    _anonymous_ = ("u",)
    _fields_ = [("len", C.c_int64), ("g", C.c_void_p), ("typ", C.c_int32), ("_pad", C.c_int32),
                ("u", _U)]


assert C.sizeof(V0) == 32 and V0.typ.offset == 16 and V0.u.offset == 24

V = C.POINTER(V0)


class Ins(C.Structure):
    _fields_ = [("op", C.c_int32), ("arg", C.c_int32), ("imm_i", C.c_int64), ("imm_f", C.c_double)]


class Expr(C.Structure):
    _fields_ = [("n", C.c_int32), ("_pad", C.c_int32), ("code", C.POINTER(Ins))]


class Agg(C.Structure):
    _fields_ = [("kind", C.c_int32), ("_pad", C.c_int32), ("body", C.POINTER(Expr))]


AGG_KIND = dict(sum=0, avg=1, min=2, max=3, len=4)


def aggs(spec):
    """[(kind, expr-or-None), ...] -> ctypes array of Agg (expressions kept alive)."""
    arr = (Agg * len(spec))()
    keep = []
    for k, (kind, e) in enumerate(spec):
        arr[k].kind = AGG_KIND[kind]
        if e is not None:
            keep.append(e)
            arr[k].body = C.pointer(e)
    arr._keep = keep
    return arr


def expr(code):
    """Build a postfix expression: items are ('COL', k) / ('I', v) / ('F', v) / 'ADD' ..."""
    ins = (Ins * len(code))()
    for k, it in enumerate(code):
        if isinstance(it, tuple):
            tag, v = it
            if tag == "COL":
                ins[k] = Ins(OP["COL"], int(v), 0, 0.0)
            elif tag == "I":
                ins[k] = Ins(OP["CONST_I"], 0, int(v), 0.0)
            elif tag == "F":
                ins[k] = Ins(OP["CONST_F"], 0, 0, float(v))
            else:
                raise ValueError(tag)
        else:
            ins[k] = Ins(OP[it], 0, 0, 0.0)
    e = Expr(len(code), 0, ins)
    e._keep = ins
    return e


class HorseError(RuntimeError):
    def __init__(self, code, where, detail=""):
        super().__init__(f"{where} failed with code {code} {detail}".strip())
        self.code = code


_SIG = {
    1: [V, V], 2: [V, V, V], 3: [V, V, V, V],
}
_FN1 = C.CFUNCTYPE(C.c_int32, V, V)
_FN2 = C.CFUNCTYPE(C.c_int32, V, V, V)

_ARITY = {
    "pfnIndex": 2, "pfnLen": 1, "pfnWhere": 1, "pfnSum": 1, "pfnAvg": 1, "pfnMin": 1, "pfnMax": 1,
    "pfnRaze": 1, "pfnGroup": 1, "pfnGroupTrie": 1, "pfnGroupBasic": 1, "pfnLt": 2, "pfnLeq": 2,
    "pfnGt": 2, "pfnGeq": 2, "pfnEq": 2, "pfnNeq": 2, "pfnPlus": 2, "pfnMinus": 2, "pfnMul": 2,
    "pfnDiv": 2, "pfnAnd": 2, "pfnNand": 2, "pfnOr": 2, "pfnNor": 2, "pfnXor": 2,
    "pfnCompress": 2, "pfnIndexOf": 2, "pfnOrderBy": 2, "pfnMember": 2, "pfnKeys": 1,
    "pfnValues": 1, "pfnJoinIndex": 3, "pfnEnum": 2, "pfnFetch": 1, "pfnColumnValue": 2,
    "pfnLoadTable": 1, "pfnAddFKey": 3, "pfnLike": 2, "pfnSubString": 2,
}


class PrimitiveLib:
    """Binding of a library exporting the pfn* ABI (optionally prefixed)."""

    def __init__(self, cdll, prefix=""):
        self.dll = cdll
        self.prefix = prefix
        for name, ar in _ARITY.items():
            f = getattr(cdll, prefix + name)
            f.argtypes = [V] * (ar + 1)
            f.restype = C.c_int32
        for name in ("pfnCompare", "pfnArith", "pfnReduce"):
            f = getattr(cdll, prefix + name)
            f.argtypes = [V, V, V, C.c_int32] if name != "pfnReduce" else [V, V, C.c_int32]
            f.restype = C.c_int32
        f = getattr(cdll, prefix + "pfnList")
        f.argtypes = [V, C.c_int64, C.POINTER(V)]
        f.restype = C.c_int32
        for name in ("pfnEach",):
            f = getattr(cdll, prefix + name)
            f.argtypes = [V, V, C.c_void_p]
            f.restype = C.c_int32
        f = getattr(cdll, prefix + "pfnEachRight")
        f.argtypes = [V, V, V, C.c_void_p]
        f.restype = C.c_int32

    def fn(self, name):
        return getattr(self.dll, self.prefix + name)

    def addr(self, name):
        return C.cast(getattr(self.dll, self.prefix + name), C.c_void_p).value

    def call(self, name, *args):
        rc = self.fn(name)(*args)
        if rc != 0:
            raise HorseError(rc, name)
        return rc
