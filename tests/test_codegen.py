"""CPU checks of the fused-loop code generator: the HIP C++ emitted for the benchmark's fused
shapes compiles for gfx950 with hipRTC (no device is touched)."""
import ctypes as C
import os

import pytest

from horsepower_amd import queries as Q
from horsepower_amd import runtime
from horsepower_amd.vnode import H_E, H_I, aggs, expr


@pytest.fixture(scope="module")
def dll():
    if not os.path.exists(runtime.LIB_PATH):
        pytest.fail("libhorsepower_amd.so is not built")
    d = C.CDLL(runtime.LIB_PATH)
    d.hpaJitCompileCheck.restype = C.c_int32
    d.hpaLastErrorDetail.restype = C.c_char_p
    return d


def check(dll, typs, cond, spec):
    t = (C.c_int32 * len(typs))(*typs)
    a = aggs(spec)
    rc = dll.hpaJitCompileCheck(len(typs), t, C.byref(cond) if cond else None, len(spec), C.cast(a, C.c_void_p))
    assert rc == 0, dll.hpaLastErrorDetail()


def test_q1_one_pass_kernel_compiles(dll):
    cond, spec = Q.q1_chf_spec()
    check(dll, [H_I, H_I, H_E, H_E, H_E], cond, spec)


def test_minmax_and_division_compile(dll):
    cond = expr([("COL", 0), ("F", 0.5), "GT", ("COL", 1), ("I", 3), "NEQ", "OR"])
    spec = [("min", expr([("COL", 0)])), ("max", expr([("COL", 1), ("I", 2), "DIV"])),
            ("sum", expr([("COL", 1), "NEG"])), ("len", None)]
    check(dll, [H_E, H_I], cond, spec)


def test_bad_expression_rejected(dll):
    t = (C.c_int32 * 1)(H_E)
    bad = expr([("COL", 0), "ADD"])
    a = aggs([("sum", bad)])
    assert dll.hpaJitCompileCheck(1, t, None, 1, C.cast(a, C.c_void_p)) != 0
