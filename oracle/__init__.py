"""TEST INFRASTRUCTURE ONLY: ctypes driver of the CPU oracle (oracle/horse_oracle.c).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this package, as the
checker; the product never imports it.
"""
import ctypes as C
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle.so")

_dll = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


def _heap_reservation():
    """Bytes of address space for the oracle's bump heap: half the host's memory (at least
    4 GiB), or HPA_ORACLE_HEAP_GB.  Pages are committed only when touched, and a heap that
    fills up makes every primitive return HPA_E_MEMORY."""
    gb = os.environ.get("HPA_ORACLE_HEAP_GB")
    if gb:
        return int(float(gb) * (1 << 30))
    try:
        phys = os.sysconf("SC_PHYS_PAGES") * os.sysconf("SC_PAGE_SIZE")
    except (ValueError, OSError):
        phys = 8 << 30
    return max(4 << 30, phys // 2)


def heap_capacity():
    return load().orc_heap_capacity()


def load(threads=0, heap_bytes=None):
    global _dll
    if _dll is None:
        if not os.path.exists(LIB):
            build()
        _dll = C.CDLL(LIB)
        _dll.orc_init.argtypes = [C.c_int64, C.c_int32]
        _dll.orc_init.restype = C.c_int32
        _dll.orc_heap_capacity.restype = C.c_int64
        _dll.orc_set_heap_top.argtypes = [C.c_int64]
        _dll.orc_set_heap_top.restype = C.c_int64
        if _dll.orc_init(heap_bytes or _heap_reservation(), threads) != 0:
            raise RuntimeError("oracle heap allocation failed")
    elif threads:
        _dll.orc_init(0, threads)
    return _dll


def backend(threads=0):
    """A horsepower_amd.backend.Backend over the oracle's orc_pfn* functions."""
    from horsepower_amd.backend import Backend
    from horsepower_amd.vnode import V, PrimitiveLib

    dll = load(threads)
    dll.orc_alloc_node.restype = V
    dll.orc_get_heap.restype = C.c_int64
    dll.orc_set_heap.argtypes = [C.c_int64]
    dll.orc_set_heap.restype = C.c_int64
    dll.orc_symbol.argtypes = [C.c_char_p]
    dll.orc_symbol.restype = C.c_int64
    dll.orc_threads.restype = C.c_int32
    dll.orc_symbol_name.argtypes = [C.c_int64]
    dll.orc_symbol_name.restype = C.c_char_p
    dll.orc_serializeV.argtypes = [V, C.c_char_p]
    dll.orc_serializeV.restype = C.c_int32
    dll.orc_readSerializeV.argtypes = [V, C.c_char_p]
    dll.orc_readSerializeV.restype = C.c_int32
    dll.orc_sum_neumaier.argtypes = [C.c_void_p, C.c_int64]
    dll.orc_sum_neumaier.restype = C.c_double
    dll.orc_sum_neumaier_idx.argtypes = [C.c_void_p, C.c_void_p, C.c_int64]
    dll.orc_sum_neumaier_idx.restype = C.c_double
    dll.orc_q6_fused.argtypes = [C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                 C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_int32]
    dll.orc_q6_fused.restype = C.c_double

    class OracleBackend(Backend):
        name = "oracle"

        def node(self):
            return dll.orc_alloc_node()

        def read_raw(self, z, nbytes, dst):
            C.memmove(dst, z.contents.g, nbytes)

        def symbol(self, name):
            return dll.orc_symbol(name.encode())

        def mark(self):
            return dll.orc_get_heap()

        def release(self, m):
            dll.orc_set_heap(m)

        def sync(self):
            pass

        def threads(self):
            return dll.orc_threads()

    return OracleBackend(PrimitiveLib(dll, "orc_"))


def neumaier(x):
    import numpy as np
    dll = load()
    a = np.ascontiguousarray(x, dtype=np.float64)
    return dll.orc_sum_neumaier(C.c_void_p(a.ctypes.data), a.shape[0])
