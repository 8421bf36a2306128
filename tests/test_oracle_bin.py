"""The oracle's serializeV / readSerializeV (the .bin snapshot format, backend/h_io.cc:630-776),
pinned on the CPU: the file the oracle writes for a table of every payload kind equals, byte for
byte outside the fields the format leaves free, an image packed independently from the format's
description (tests/binfmt.py); the oracle reads its own file back to the same values."""
import numpy as np
import pytest

import binfmt as F


def table_spec(n, seed):
    r = np.random.default_rng(seed)
    cols = [
        (F.H_I, r.integers(-10**6, 10**6, n).astype(np.int32)),
        (F.H_E, r.random(n)),
        (F.H_C, r.integers(65, 91, n).astype(np.uint8)),
        (F.H_Q, [b"sym%d" % k for k in r.integers(0, 7, n)]),
        (F.H_S, [b"x" * int(k) for k in r.integers(0, 5, n)]),
        (F.H_L, np.array([7], np.int64)),                   # one value: kept in the node
        (F.H_E, np.zeros(0, np.float64)),
        (F.H_C, np.array([66], np.uint8)),
        (F.H_G, [(F.H_I, np.arange(3, dtype=np.int32)), (F.H_S, [b"in", b"list"])]),
    ]
    names = [b"c%d" % k for k in range(len(cols))]
    return (F.H_A, (n, len(cols), (F.H_Q, names), (F.H_G, cols)))


@pytest.mark.parametrize("n", [2, 5, 1000])
def test_oracle_writer_matches_format(orc, tmp_path, n):
    spec = table_spec(n, n)
    keep = []
    t = F.build(orc, spec, keep)
    path = tmp_path / "t.bin"
    assert orc.lib.dll.orc_serializeV(t, str(path).encode()) == 0
    assert F.same(path.read_bytes(), spec)
    z = orc.node()
    assert orc.lib.dll.orc_readSerializeV(z, str(path).encode()) == 0
    assert F.read_back(orc, z, spec)


def test_oracle_reader_rejects_short_file(orc, tmp_path):
    spec = table_spec(10, 1)
    keep = []
    path = tmp_path / "t.bin"
    assert orc.lib.dll.orc_serializeV(F.build(orc, spec, keep), str(path).encode()) == 0
    data = path.read_bytes()
    (tmp_path / "short.bin").write_bytes(data[:-3])
    assert orc.lib.dll.orc_readSerializeV(orc.node(), str(tmp_path / "short.bin").encode()) == 2
