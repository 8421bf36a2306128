"""hpaSumNoAlloc -- pfnSum of an E vector with no allocation (its partials live in the library's
static device scratch) -- against pfnSum on the same vectors, two calls in a row with the second
smaller, through the C-ABI."""
import ctypes as C
import math

import numpy as np
import pytest

from horsepower_amd.vnode import H_E, V, V0

pytestmark = pytest.mark.gpu


def _bind(dll):
    dll.hpaSumNoAlloc.argtypes = [V, C.c_void_p]
    dll.hpaSumNoAlloc.restype = C.c_int32
    dll.hpaDeviceAlloc.argtypes = [C.c_int64]
    dll.hpaDeviceAlloc.restype = C.c_void_p
    dll.hpaDeviceFree.argtypes = [C.c_void_p]
    dll.hpaDeviceFree.restype = C.c_int32


def _read_f64(gpu, dptr):
    node = V0()
    node.len, node.typ, node.g = 1, H_E, dptr            # a view of the device double
    out = np.zeros(1, np.float64)
    assert gpu.lib.dll.hpaReadPayload(C.pointer(node), out.ctypes.data_as(C.c_void_p), 8) == 0
    return float(out[0])


def test_sum_noalloc_matches_pfnsum(gpu):
    dll = gpu.lib.dll
    _bind(dll)
    out = dll.hpaDeviceAlloc(256)
    assert out
    try:
        r = np.random.default_rng(19)
        for n in (6_000_000, 100_003, 1000, 1, 0):         # each call smaller than the one before
            print("noalloc sum", n, flush=True)
            x = np.round(r.uniform(-1e5, 1e5, n), 2)
            xv = gpu.vec(x, H_E)
            assert dll.hpaSumNoAlloc(xv, out) == 0, gpu.lib.detail()
            got = _read_f64(gpu, out)
            want = float(gpu.read(gpu.pfnSum(xv))[0])
            assert got == want, (n, got, want)                # the same kernels, bit for bit
            exact = math.fsum(x)
            assert abs(got - exact) <= 1e-12 * max(abs(exact), 1e-300), (n, got, exact)
        # errors: a payload outside HBM, another type
        host = gpu.host.vec(np.arange(10, dtype=np.float64))
        assert dll.hpaSumNoAlloc(host, out) == 1
        assert dll.hpaSumNoAlloc(gpu.vec(np.arange(10, dtype=np.int64)), out) == 6
    finally:
        dll.hpaDeviceFree(out)
