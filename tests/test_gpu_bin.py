"""Table snapshots (.bin, serializeV / readSerializeV) on the GPU: files the oracle writes (its
writer is pinned byte for byte against the format in tests/test_oracle_bin.py) are read into HBM
and hold the same values; the GPU's writer produces the same bytes and the oracle reads them back;
initTableFromBin registers the table for load_table / column_value.  Payloads larger than the
32 MB staging buffers stream in several chunks."""
import ctypes as C

import numpy as np
import pytest

import binfmt as F
from test_oracle_bin import table_spec
from horsepower_amd.vnode import H_Q, V, HorseError

pytestmark = pytest.mark.gpu


def write_oracle(orc, spec, path):
    keep = []
    assert orc.lib.dll.orc_serializeV(F.build(orc, spec, keep), str(path).encode()) == 0


@pytest.mark.parametrize("n", [2, 1000, 300_000])
def test_read_oracle_file(gpu, orc, tmp_path, n):
    spec = table_spec(n, n + 1)
    path = tmp_path / "t.bin"
    write_oracle(orc, spec, path)
    z = gpu.node()
    moved = C.c_int64(0)
    gpu.lib.rc("hpaReadSerialized", z, str(path).encode(), C.byref(moved))
    assert F.read_back(gpu, z, spec)
    assert moved.value > 0 or n <= 2
    # the GPU writes the same bytes; the oracle reads them back
    back = tmp_path / "g.bin"
    gpu.lib.rc("hpaWriteSerialized", z, str(back).encode())
    assert F.same(back.read_bytes(), spec)
    zo = orc.node()
    assert orc.lib.dll.orc_readSerializeV(zo, str(back).encode()) == 0
    assert F.read_back(orc, zo, spec)


def test_large_columns_stream(gpu, orc, tmp_path):
    """A 10M-row lineitem-like table: 80 MB f64 and 40 MB i32 payloads, three staging chunks."""
    r = np.random.default_rng(4)
    n = 10_000_000
    price = r.integers(90000, 10500000, n) / 100.0
    ship = r.integers(0, 2500, n).astype(np.int32)
    flag = r.integers(65, 70, n).astype(np.uint8)
    spec = (F.H_A, (n, 3, (F.H_Q, [b"l_extendedprice", b"l_shipdate", b"l_returnflag"]),
                    (F.H_G, [(F.H_E, price), (F.H_I, ship), (F.H_C, flag)])))
    path = tmp_path / "lineitem.bin"
    write_oracle(orc, spec, path)
    z = gpu.node()
    moved = C.c_int64(0)
    gpu.lib.rc("hpaReadSerialized", z, str(path).encode(), C.byref(moved))
    assert moved.value == n * (8 + 4 + 1) + 1
    assert F.read_back(gpu, z, spec)
    # the columns are in HBM and feed the primitives directly
    from horsepower_amd.runtime import cell
    cols = cell(z, 1)
    s = gpu.read_scalar(gpu.pfnSum(cell(cols, 0)))
    assert abs(s - np.sum(price)) <= 1e-12 * abs(s)


def test_init_table_from_bin(gpu, orc, tmp_path):
    spec = table_spec(5000, 9)
    path = tmp_path / "orders.bin"
    write_oracle(orc, spec, path)
    tz = V()
    gpu.lib.rc("hpaInitTableFromBin", b"orders_bin", str(path).encode(), C.byref(tz))
    sym = lambda s: gpu.host.vec(np.array([gpu.symbol(s)], np.int64), H_Q)
    t = gpu.pfnLoadTable(sym("orders_bin"))
    c0 = gpu.read(gpu.pfnColumnValue(t, sym("c0")))
    assert np.array_equal(c0, spec[1][3][1][0][1])


def test_truncated_file(gpu, orc, tmp_path):
    spec = table_spec(1000, 3)
    path = tmp_path / "t.bin"
    write_oracle(orc, spec, path)
    (tmp_path / "short.bin").write_bytes(path.read_bytes()[:-5])
    with pytest.raises(HorseError) as ei:
        gpu.lib.rc("hpaReadSerialized", gpu.node(), str(tmp_path / "short.bin").encode(), None)
    assert ei.value.code == 2
