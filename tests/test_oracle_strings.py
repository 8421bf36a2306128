"""The oracle's symbol and string primitives pinned on the CPU, each by an independent rule:

  like        the reference's own recipe run by another regular-expression engine: the pattern
              translated as genLikeString does (h_system.cc:319-334) and matched anchored with
              Python's `re` (whose defaults agree with PCRE2's: . stops at a newline, $ matches at
              the end or before a final newline), over random subjects and patterns built from
              the bytes that matter (%, _, ., \\, newline, literals)
  symbols     order compares against Python's sort of the names' bytes (strcmp order)
  sub_string  a closed form of the reference's loop (h_primitive.cc:2476-2479)
  member      Python set membership of the strings
"""
import re

import numpy as np
import pytest

from horsepower_amd.vnode import H_I, H_Q, HorseError


def like_regex(pat):
    """genLikeString, restated in Python: ^ + % -> .*, _ -> ., . and \\ escaped + $."""
    out = b"^"
    for ch in pat:
        c = bytes([ch])
        out += b".*" if c == b"%" else b"." if c == b"_" else b"\\" + c if c in (b".", b"\\") else c
    return out + b"$"


def rand_bytes(r, alphabet, lo, hi):
    return r.choice(np.frombuffer(alphabet, np.uint8), r.integers(lo, hi + 1)).astype(np.uint8).tobytes()


def test_like_vs_re(orc):
    r = np.random.default_rng(11)
    subj = [rand_bytes(r, b"ab.\\\n", 0, 7) for _ in range(400)]
    subj += [b"", b"\n", b"a\n", b"ab\n", b"a.b", b"a\\b"]
    x = orc.strs(subj)
    for k in range(250):
        pat = rand_bytes(r, b"ab%_.\\\n", 1, 6)
        got = orc.read(orc.pfnLike(x, orc.strs([pat])))
        rx = re.compile(like_regex(pat))
        want = [1 if rx.match(s) else 0 for s in subj]
        assert got.tolist() == want, pat


def test_like_symbols_chars_and_errors(orc):
    names = ["PROMO BRUSHED TIN", "STANDARD POLISHED BRASS", "PROMO", "MEDIUM POLISHED COPPER", "promo x"]
    q = np.array([orc.symbol(s) for s in names] * 3, np.int64)
    got = orc.read(orc.pfnLike(orc.vec(q, H_Q), orc.sym("PROMO%")))
    assert got.tolist() == [1, 0, 1, 0, 0] * 3
    got = orc.read(orc.pfnLike(orc.vec(q, H_Q), orc.strs([b"%BRASS"])))
    assert got.tolist() == [0, 1, 0, 0, 0] * 3
    # a char vector is one string: up to its first NUL
    cv = orc.vec(np.frombuffer(b"forest green\0xyz", np.uint8), 8)
    assert orc.read(orc.pfnLike(cv, orc.strs([b"%green"]))).tolist() == [1]
    x = orc.strs([b"a", b"b"])
    for bad, code in ((b"a*", 99), (b"(a)", 99), (b"", 7)):
        with pytest.raises(HorseError) as ei:
            orc.pfnLike(x, orc.strs([bad]))
        assert ei.value.code == code, bad


def test_symbol_order_compares(orc):
    r = np.random.default_rng(5)
    words = sorted({rand_bytes(r, b"abcAB ", 1, 5).decode() for _ in range(60)})
    ids = np.array([orc.symbol(w) for w in words], np.int64)
    a = r.integers(0, len(words), 500)
    b = r.integers(0, len(words), 500)
    key = {w: w.encode() for w in words}
    x, y = orc.vec(ids[a], H_Q), orc.vec(ids[b], H_Q)
    for op, f in (("pfnLt", lambda p, q: p < q), ("pfnLeq", lambda p, q: p <= q), ("pfnGt", lambda p, q: p > q),
                  ("pfnGeq", lambda p, q: p >= q), ("pfnEq", lambda p, q: p == q), ("pfnNeq", lambda p, q: p != q)):
        got = orc.read(getattr(orc, op)(x, y))
        want = [int(f(key[words[i]], key[words[j]])) for i, j in zip(a, b)]
        assert got.tolist() == want, op
    got = orc.read(orc.pfnLt(x, orc.sym(words[10])))
    assert got.tolist() == [int(key[words[i]] < key[words[10]]) for i in a]


def substr_closed(s, start, seg):
    room = len(s) - start                       # the reference counts strlen - start
    n = max(min(seg, room), 0)
    return s[start - 1:start - 1 + n] + b" " * (seg - n)


@pytest.mark.parametrize("start,seg", [(1, 2), (2, 3), (1, 0), (4, 5), (9, 2)])
def test_sub_string(orc, start, seg):
    subj = [b"13-555-0100", b"31", b"a", b"", b"hello world", b"12345"]
    got = orc.read_strs(orc.pfnSubString(orc.strs(subj), orc.vec(np.array([start, seg], np.int32), H_I)))
    assert got == [substr_closed(s, start, seg) for s in subj]
    with pytest.raises(HorseError) as ei:
        orc.pfnSubString(orc.strs(subj), orc.vec(np.array([0, 2], np.int32), H_I))
    assert ei.value.code == 1


def test_member_strings(orc):
    codes = [b"13", b"31", b"23", b"29", b"30", b"18", b"17"]
    r = np.random.default_rng(3)
    y = [rand_bytes(r, b"123890", 2, 2) for _ in range(300)]
    got = orc.read(orc.pfnMember(orc.strs(y), orc.strs(codes)))
    assert got.tolist() == [int(s in set(codes)) for s in y]
