"""Writes tests/golden/kat.json: known-answer vectors taken from the reference's own files.

Sources (read as text, never executed):
  docs/mkdocs/docs/horseir/builtin/{compress,where,indexing,member,order,group,indexof,
  reduce,len,lt,eq,unique}.md  -- the "Examples" blocks
  tests/unit/pass/t33.hir (expected value in a comment), t34.hir (each/sum/raze)
The docs of group, index_of and the two-key order disagree with the implementation
(src/horseir/v3/backend/h_libs.cc:1295-1329, h_primitive.cc:1525-1538, h_libs.cc:604-617);
for those the fixture keeps the doc's text in "doc_expected" and the implementation's
answer, derived by hand from the cited code, in "expected".

Run in a container that holds the reference:  python tests/golden/make_golden.py
"""
import json
import os
import re

REF = "/root/reference"
DOCS = os.path.join(REF, "docs/mkdocs/docs/horseir/builtin")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "kat.json")

LIT = re.compile(r"\(?([-0-9., ]+)\)?:(bool|i8|i16|i32|i64|f32|f64)")


def parse_lit(s):
    m = LIT.fullmatch(s.strip())
    if not m:
        raise ValueError(s)
    vals, typ = m.group(1), m.group(2)
    xs = [v for v in re.split(r"[ ,]+", vals.strip()) if v]
    conv = float if typ.startswith("f") else int
    return {"type": typ, "values": [conv(v) for v in xs]}


def example_lines(fname):
    text = open(os.path.join(DOCS, fname)).read()
    blocks = re.findall(r"```[a-z-]*\n(.*?)```", text.split("Examples", 1)[1], re.S)
    lines = []
    for b in blocks:
        lines += [ln.strip() for ln in b.splitlines() if ln.strip()]
    return lines


def call_args(line):
    inner = line[line.index("(") + 1: line.rindex(")")]
    parts, depth, cur = [], 0, ""
    for ch in inner:
        if ch == "(":
            depth += 1
        if ch == ")":
            depth -= 1
        if ch == "," and depth == 0:
            parts.append(cur)
            cur = ""
        else:
            cur += ch
    parts.append(cur)
    return [parse_lit(p) for p in parts]


def main():
    kat = []

    def add(name, src, fn, line, result, expected=None, note=None):
        e = {"name": name, "source": src, "fn": fn, "args": call_args(line), "doc_expected": result}
        e["expected"] = expected if expected is not None else parse_lit(result)
        if note:
            e["note"] = note
        kat.append(e)

    for fn, fname in (("compress", "compress.md"), ("where", "where.md"), ("index", "indexing.md"),
                      ("member", "member.md"), ("min", "reduce.md"), ("len", "len.md"),
                      ("eq", "eq.md"), ("unique", "unique.md")):
        lines = example_lines(fname)
        call, res = lines[0], lines[1]
        src = f"docs/mkdocs/docs/horseir/builtin/{fname}"
        if fn == "unique":
            add("doc_unique", src, fn, call, res)
            continue
        add(f"doc_{fn}", src, fn, call, res)

    lt = example_lines("lt.md")
    add("doc_lt", "docs/mkdocs/docs/horseir/builtin/lt.md", "lt", lt[0], "(1,0,0):bool",
        note="doc prints '1 0 0:bool'")

    od = example_lines("order.md")
    add("doc_order_1", "docs/mkdocs/docs/horseir/builtin/order.md", "order", od[0], od[1])
    # second example: x = @list(a, b); @order(x, (1,0):bool)
    a = parse_lit(re.search(r"@list\((\([^)]*\):i32),", od[2]).group(1))
    b = parse_lit(re.search(r",(\([^)]*\):i32)\);", od[2]).group(1))
    kat.append({
        "name": "doc_order_2", "source": "docs/mkdocs/docs/horseir/builtin/order.md", "fn": "order_list",
        "args": [a, b, {"type": "bool", "values": [1, 0]}], "doc_expected": od[4],
        "expected": {"type": "i64", "values": [2, 4, 1, 0, 5, 3]},
        "note": "column 0 stable (merge sort) 7,12,12,32,45,78; the run of 12s at rows 1,4 has "
                "column-1 values 34,78 which are non-decreasing, so the descending flag reverses "
                "the run (lib_order_by_sorted, h_libs.cc:604-617)"})

    gd = example_lines("group.md")
    kat.append({
        "name": "doc_group", "source": "docs/mkdocs/docs/horseir/builtin/group.md", "fn": "group",
        "args": [parse_lit(re.search(r"group\((.*)\)$", gd[0]).group(1))], "doc_expected": gd[1],
        "expected": {"keys": [2, 1, 0], "groups": [[2], [1, 3], [0]]},
        "note": "i32 keys are radix-sorted on unsigned bytes: 2 < 3 < 0xffffffff (-1); key = first row "
                "of each group (lib_group_by_normal_int, h_libs.cc:1295-1329)"})

    io = example_lines("indexof.md")
    kat.append({
        "name": "doc_index_of", "source": "docs/mkdocs/docs/horseir/builtin/indexof.md", "fn": "index_of",
        "args": call_args(io[0]), "doc_expected": io[1],
        "expected": {"type": "i64", "values": [1, 2, 2]},
        "note": "pfnIndexOf(z,x,y) has len(z)=len(y) (h_primitive.cc:1534): positions of 4,9,16 in "
                "(3,4); misses give len(x)=2; (3,4) is ordered so searchOrdered answers"})

    t33 = open(os.path.join(REF, "tests/unit/pass/t33.hir")).read()
    exp = re.search(r"expect: (\([^)]*\):i32)", t33).group(1)
    kat.append({"name": "t33_plus", "source": "tests/unit/pass/t33.hir", "fn": "plus_chain",
                "args": [parse_lit("1:i32"), parse_lit("(1,2,3):i32"), parse_lit("10:i32")],
                "expected": parse_lit(exp)})
    kat.append({"name": "t34_each_sum", "source": "tests/unit/pass/t34.hir", "fn": "each_sum_raze",
                "args": [parse_lit("(1,2,3):i32"), parse_lit("(4,5,6):i32")],
                "expected": {"type": "i64", "values": [6, 15]},
                "note": "sum of i32 is i64 (h_primitive.cc:771, the file's own TODO)"})

    with open(OUT, "w") as f:
        json.dump(kat, f, indent=1)
    print(f"wrote {len(kat)} vectors to {OUT}")


if __name__ == "__main__":
    main()
