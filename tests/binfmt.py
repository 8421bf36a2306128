"""The .bin snapshot format of serializeV (backend/h_io.cc:630-696), restated in Python for the
tests: `image(node_spec)` packs the bytes a file must hold, with a mask of the bytes whose value
the format leaves free (the pointer and padding fields of each 32-byte V0 header, its union, and
a table's foreign-key pointer).  A node spec is (type code, payload), where the payload is a
numpy array for basic types, a list of bytes for symbols (Q) and strings (S), a list of specs for
a list (G), and (row, col, keys spec, values spec) for a table (A)."""
import struct

import numpy as np

H_I, H_L, H_E, H_C, H_Q, H_S, H_G, H_A = 3, 4, 6, 8, 9, 10, 17, 90


def _header(typ, n):
    b = struct.pack("<qQiiq", n, 0, typ, 0, 0)
    m = b"\x00" * 8 + b"\xff" * 8 + b"\x00" * 4 + b"\xff" * 12
    return b, m


def image(spec):
    typ, pay = spec
    out, mask = [], []

    def add(b, m=None):
        out.append(b)
        mask.append(m if m is not None else b"\x00" * len(b))

    if typ == H_A:
        row, col, keys, vals = pay
        add(*_header(typ, 2))
        add(struct.pack("<qqQ", row, col, 0), b"\x00" * 16 + b"\xff" * 8)
        for sub in (keys, vals):
            b, m = image(sub)
            add(b, m)
    elif typ == H_G:
        add(*_header(typ, len(pay)))
        for sub in pay:
            b, m = image(sub)
            add(b, m)
    elif typ in (H_Q, H_S):
        add(*_header(typ, len(pay)))
        add(b"".join(struct.pack("<q", len(s)) for s in pay))
        add(b"".join(s + b"\x00" for s in pay))
    else:
        a = np.ascontiguousarray(pay)
        add(*_header(typ, a.shape[0]))
        data = a.tobytes()
        if typ == H_C and a.shape[0] > 1:
            # getTypeSize adds one byte for char vectors, which nothing sets
            add(data + b"\x00", b"\x00" * len(data) + b"\xff")
        else:
            add(data + (b"\x00" if typ == H_C else b""))
    return b"".join(out), b"".join(mask)


def same(file_bytes, spec):
    want, mask = image(spec)
    if len(file_bytes) != len(want):
        return False
    f = np.frombuffer(file_bytes, np.uint8)
    w = np.frombuffer(want, np.uint8)
    k = np.frombuffer(mask, np.uint8) == 0
    return bool(np.array_equal(f[k], w[k]))


def build(B, spec, keep):
    """The node of a spec on backend B (columns resident where B keeps them)."""
    import ctypes as C
    from horsepower_amd.vnode import V0
    typ, pay = spec
    if typ == H_A:
        row, col, keys, vals = pay
        kn, vn = build(B, keys, keep), build(B, vals, keep)
        cells = (V0 * 2)()
        cells[0], cells[1] = kn.contents, vn.contents

        class A0(C.Structure):
            _fields_ = [("row", C.c_int64), ("col", C.c_int64), ("fkey", C.c_void_p)]
        a = A0(row, col, None)
        t = C.pointer(V0())
        t.contents.typ, t.contents.len = H_A, 2
        t.contents.g = C.addressof(cells)
        t.contents.p = C.addressof(a)
        keep += [kn, vn, cells, a, t]
        return t
    if typ == H_G:
        items = [build(B, sub, keep) for sub in pay]
        keep += items
        return B.list(items)
    if typ == H_Q:
        return B.vec(np.array([B.symbol(s.decode()) for s in pay], np.int64), H_Q)
    if typ == H_S:
        return B.strs(pay)
    return B.vec(np.ascontiguousarray(pay), typ)


def read_back(B, z, spec):
    """True when node z holds what spec says (names and strings by value)."""
    from horsepower_amd.runtime import cell
    typ, pay = spec
    v = z.contents
    if v.typ != typ:
        return False
    if typ == H_A:
        return read_back(B, cell(z, 0), pay[2]) and read_back(B, cell(z, 1), pay[3])
    if typ == H_G:
        return v.len == len(pay) and all(read_back(B, cell(z, k), s) for k, s in enumerate(pay))
    if typ == H_Q:
        lib = B.lib.dll
        name = lib.hpaSymbolName if B.name == "gpu" else lib.orc_symbol_name
        return [name(int(q)) for q in B.read(z)] == list(pay)
    if typ == H_S:
        return B.read_strs(z) == list(pay)
    got = B.read(z)
    return got.dtype == np.asarray(pay).dtype and np.array_equal(got, np.asarray(pay))
