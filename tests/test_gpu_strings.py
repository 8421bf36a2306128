"""Symbol and string primitives on the GPU against the oracle (itself pinned against Python's
`re` and closed forms in tests/test_oracle_strings.py): like over string, symbol and char
values, symbol order compares, sub_string, member of strings, and the string mirrors behind them
(uploaded columns, host payloads staged for one call, device-written sub_string results read
back and reused after an arena rewind).  Everything is compared bit for bit."""
import numpy as np
import pytest

from horsepower_amd.vnode import H_I, H_Q, HorseError

pytestmark = pytest.mark.gpu


def rand_bytes(r, alphabet, lo, hi):
    return r.choice(np.frombuffer(alphabet, np.uint8), r.integers(lo, hi + 1)).astype(np.uint8).tobytes()


@pytest.mark.parametrize("staged", [False, True])
def test_like_random_vs_oracle(gpu, orc, staged):
    r = np.random.default_rng(21 + staged)
    subj = [rand_bytes(r, b"ab.\\\n", 0, 9) for _ in range(3000)] + [b"", b"\n", b"a\n", b"ab\n"]
    xg = gpu.host.strs(subj) if staged else gpu.strs(subj)
    xo = orc.strs(subj)
    m = gpu.mark()
    for _ in range(60):
        pat = rand_bytes(r, b"ab%_.\\\n", 1, 7)
        g = gpu.read(gpu.pfnLike(xg, gpu.strs([pat])))
        o = orc.read(orc.pfnLike(xo, orc.strs([pat])))
        assert np.array_equal(g, o), pat
    gpu.release(m)


WORDS = [b"special", b"requests", b"packages", b"deposits", b"furiously", b"ironic", b"blithely",
         b"Customer", b"Complaints", b"green", b"forest", b"PROMO", b"BRASS"]


def comments(r, n):
    k = r.integers(2, 9, n)
    w = r.integers(0, len(WORDS), int(k.sum()))
    out, p = [], 0
    for c in k:
        out.append(b" ".join(WORDS[j] for j in w[p:p + c]))
        p += c
    return out


def test_like_comment_column(gpu, orc):
    """A 300K-row comment column (q13's like(o_comment, "%special%requests%"), q16's, ...)."""
    r = np.random.default_rng(13)
    col = comments(r, 300_000)
    xg, xo = gpu.strs(col), orc.strs(col)
    for pat in (b"%special%requests%", b"%Customer%Complaints%", b"forest%", b"%BRASS", b"%green%", b"_r%"):
        g = gpu.read(gpu.pfnLike(xg, gpu.strs([pat])))
        o = orc.read(orc.pfnLike(xo, orc.strs([pat])))
        assert np.array_equal(g, o), pat
        assert 0 < g.sum() < g.size


def test_like_symbols(gpu, orc):
    names = [f"{a} {b} {c}" for a in ("PROMO", "STANDARD", "MEDIUM", "ECONOMY") for b in ("BRUSHED", "POLISHED", "PLATED")
             for c in ("TIN", "BRASS", "COPPER", "STEEL")]
    r = np.random.default_rng(7)
    pick = r.integers(0, len(names), 1_000_000)
    out = {}
    for B in (gpu, orc):
        ids = np.array([B.symbol(s) for s in names], np.int64)
        q = B.vec(ids[pick], H_Q)
        out[B.name] = [B.read(B.pfnLike(q, B.sym(p))) for p in ("PROMO%", "%BRASS", "MEDIUM POLISHED%", "%_T%")]
        out[B.name].append(B.read(B.pfnLike(q, B.strs([b"%PLATED%"]))))
    for g, o in zip(out["gpu"], out["oracle"]):
        assert np.array_equal(g, o)
    cv = np.frombuffer(b"forest green\0tail", np.uint8)
    for pat in (b"%green", b"forest%", b"%tail"):
        assert gpu.read(gpu.pfnLike(gpu.vec(cv, 8), gpu.strs([pat]))).tolist() == \
            orc.read(orc.pfnLike(orc.vec(cv, 8), orc.strs([pat]))).tolist()


def test_like_errors(gpu):
    x = gpu.strs([b"a", b"bb", b"ccc"])
    for bad, code in ((b"a*", 99), (b"(a)", 99), (b"a|b", 99), (b"", 7)):
        with pytest.raises(HorseError) as ei:
            gpu.pfnLike(x, gpu.strs([bad]))
        assert ei.value.code == code, bad
    with pytest.raises(HorseError) as ei:
        gpu.pfnLike(gpu.vec(np.arange(3, dtype=np.int32)), gpu.strs([b"a"]))
    assert ei.value.code == 1


def test_symbol_order_compares(gpu, orc):
    r = np.random.default_rng(17)
    words = sorted({rand_bytes(r, b"abcAB z", 1, 6).decode() for _ in range(500)})
    a = r.integers(0, len(words), 200_000)
    b = r.integers(0, len(words), 200_000)
    res = {}
    for B in (gpu, orc):
        ids = np.array([B.symbol(w) for w in words], np.int64)
        x, y = B.vec(ids[a], H_Q), B.vec(ids[b], H_Q)
        res[B.name] = [B.read(getattr(B, op)(x, y)) for op in ("pfnLt", "pfnLeq", "pfnGt", "pfnGeq", "pfnEq", "pfnNeq")]
        res[B.name] += [B.read(B.pfnGeq(x, B.sym(words[len(words) // 2]))), B.read(B.pfnLt(B.sym(words[3]), y))]
    for g, o in zip(res["gpu"], res["oracle"]):
        assert np.array_equal(g, o)


def phones(r, n):
    cc = r.integers(10, 35, n)
    return [b"%d-%03d-%03d-%04d" % (c, r.integers(100, 999), r.integers(100, 999), r.integers(0, 9999)) for c in cc]


@pytest.mark.parametrize("start,seg", [(1, 2), (4, 3), (14, 4), (1, 0)])
def test_sub_string_vs_oracle(gpu, orc, start, seg):
    r = np.random.default_rng(start * 10 + seg)
    col = phones(r, 100_000) + [b"", b"7"]
    y = np.array([start, seg], np.int32)
    g = gpu.read_strs(gpu.pfnSubString(gpu.strs(col), gpu.vec(y, H_I)))
    o = orc.read_strs(orc.pfnSubString(orc.strs(col), orc.vec(y, H_I)))
    assert g == o


def test_q22_shape(gpu, orc):
    """q22's statements (tests/hyper/online-hir/q22.hir:8-12): sub_string of the phone column,
    member of the country codes, then a like over the device-made strings; repeated after an
    arena rewind, with the strings read back between."""
    r = np.random.default_rng(22)
    col = phones(r, 400_000)
    codes = [b"13", b"31", b"23", b"29", b"30", b"18", b"17"]
    xo = orc.strs(col)
    so = orc.pfnSubString(xo, orc.vec(np.array([1, 2], np.int32), H_I))
    mo = orc.read(orc.pfnMember(so, orc.strs(codes)))
    lo = orc.read(orc.pfnLike(so, orc.strs([b"1_"])))
    xg = gpu.strs(col)
    for rep in range(2):
        m = gpu.mark()
        sg = gpu.pfnSubString(xg, gpu.vec(np.array([1, 2], np.int32), H_I))
        assert np.array_equal(gpu.read(gpu.pfnMember(sg, gpu.strs(codes))), mo), rep
        assert np.array_equal(gpu.read(gpu.pfnLike(sg, gpu.strs([b"1_"]))), lo), rep
        assert gpu.read_strs(sg)[:1000] == orc.read_strs(so)[:1000]
        gpu.release(m)


def test_member_strings_large_set(gpu, orc):
    r = np.random.default_rng(9)
    pool = [rand_bytes(r, b"abcdefgh", 1, 8) for _ in range(120_000)]
    xs = pool[:60_000]
    ys = [pool[i] for i in r.integers(0, len(pool), 500_000)]
    g = gpu.read(gpu.pfnMember(gpu.strs(ys), gpu.strs(xs)))
    o = orc.read(orc.pfnMember(orc.strs(ys), orc.strs(xs)))
    assert np.array_equal(g, o) and 0 < g.sum() < g.size


def test_import_symbols(gpu):
    """hpaImportSymbols: the caller's table gives the ids; names the library had interned move
    after it; symbol values made with the caller's ids then compare and match by name."""
    import ctypes as C
    dll = gpu.lib.dll
    before = gpu.symbol("eq")
    names = [b"zeta", b"alpha", b"eq", b"mid"]
    arr = (C.c_char_p * len(names))(*names)
    assert dll.hpaImportSymbols(len(names), arr) == 0
    assert [dll.hpaSymbolName(i) for i in range(4)] == names
    assert gpu.symbol("eq") == 2 and gpu.symbol("alpha") == 1
    assert before >= 0
    q = gpu.vec(np.array([0, 1, 2, 3, 1], np.int64), H_Q)
    assert gpu.read(gpu.pfnLt(q, gpu.sym("mid"))).tolist() == [0, 1, 1, 0, 1]
    assert gpu.read(gpu.pfnLike(q, gpu.strs([b"%a"]))).tolist() == [1, 1, 0, 0, 1]
