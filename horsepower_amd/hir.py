"""HorseIR programs on the library: a parser for the statement form the compilers emit, the
interpreter's execution (one primitive per statement, translator/interpreter.cc:42-60,609-614),
and the fusion the optimiser applies before it emits code (optimizer/elementwise.cc:239-281 FE,
autofusion.cc:770-935 FA, pattern.cc:180-219 FP), here emitting the library's fused descriptors
(hpa_expr postfix code) instead of OpenMP C text.

Fusion, over the straight-line statement list:
  * an element-wise statement (compares, +, -, *, and, or) whose value has exactly one use, by
    another element-wise statement or a sink below, is inlined into its user's expression tree
    (division stays a statement: pfnDiv's zero-divisor check, E_DIV_ZERO, is the primitive's);
  * FA reduction: sum / avg / min / max of compress(mask, body) (the compress used once), or of an
    element-wise tree, becomes one hpaFusedReduce with the mask's and the body's trees;
  * FA compress: the compresses of one mask run as one hpaCompressMany; a compressed
    element-wise tree (used once) is evaluated inside the compress (hpaCompressMap); a mask that
    is an element-wise statement used only by those compresses becomes their condition, counted
    without building the mask (hpaFusedCompress / hpaFusedCompressMap);
  * FE: an element-wise tree of two or more operations that is not absorbed by a sink becomes
    one hpaFusedMap;
  * FP: raze(each(f, each_right(index, v, groups))) with f in sum / avg / min / max / len becomes
    hpaGroupReduce.
Trees whose column types the descriptors do not take (byte-sized arithmetic, outputs other than
B/I/L/E, one-row operands) run as the plain statements.  Integer arithmetic in a tree is 64-bit,
as the descriptors define it: it agrees with the statements' own types wherever no intermediate
overflows its type.  The fused forms compute what the optimiser's emitted loops compute: float compares
are exact (the interpreter's compares allow an epsilon) and sums are compensated.
"""
import re

import numpy as np

from .vnode import (H_B, H_C, H_D, H_E, H_F, H_H, H_I, H_J, H_L, H_M, H_Q, H_S, H_T, H_U, H_W, H_Z,
                    expr)

TYPES = {"bool": H_B, "i8": H_J, "i16": H_H, "i32": H_I, "i64": H_L, "f32": H_F, "f64": H_E,
         "char": H_C, "sym": H_Q, "str": H_S, "m": H_M, "d": H_D, "z": H_Z, "u": H_U, "w": H_W, "t": H_T}

# element-wise builtins and their expression opcodes
EW = {"lt": "LT", "leq": "LEQ", "gt": "GT", "geq": "GEQ", "eq": "EQ", "neq": "NEQ",
      "plus": "ADD", "minus": "SUB", "mul": "MUL", "and": "AND", "or": "OR"}
CMP = {"lt", "leq", "gt", "geq", "eq", "neq"}
LOGIC = {"and", "or"}
REDUCE = {"sum": 0, "avg": 1, "min": 2, "max": 3}
GROUP_FN = {"sum": 0, "avg": 1, "min": 2, "max": 3, "len": 4}
PFN = {"lt": "pfnLt", "leq": "pfnLeq", "gt": "pfnGt", "geq": "pfnGeq", "eq": "pfnEq", "neq": "pfnNeq",
       "plus": "pfnPlus", "minus": "pfnMinus", "mul": "pfnMul", "div": "pfnDiv", "and": "pfnAnd",
       "or": "pfnOr", "compress": "pfnCompress", "sum": "pfnSum", "avg": "pfnAvg", "min": "pfnMin",
       "max": "pfnMax", "len": "pfnLen", "index": "pfnIndex", "group": "pfnGroup", "keys": "pfnKeys",
       "values": "pfnValues", "raze": "pfnRaze", "order": "pfnOrderBy", "join_index": "pfnJoinIndex",
       "member": "pfnMember", "index_of": "pfnIndexOf", "where": "pfnWhere", "like": "pfnLike",
       "sub_string": "pfnSubString", "load_table": "pfnLoadTable", "column_value": "pfnColumnValue",
       "fetch": "pfnFetch", "enum": "pfnEnum"}
FN_ARG = {"sum": "pfnSum", "avg": "pfnAvg", "min": "pfnMin", "max": "pfnMax", "len": "pfnLen",
          "index": "pfnIndex"}


# ---------------------------------------------------------------------------------- parsing
class Var:
    def __init__(self, name):
        self.name = name

    def __repr__(self):
        return self.name


class Lit:
    def __init__(self, vals, typ, vector):
        self.vals, self.typ, self.vector = vals, typ, vector

    def __repr__(self):
        return f"{self.vals}:{self.typ}"


class Fn:
    def __init__(self, name):
        self.name = name


class Stmt:
    def __init__(self, target, fn, args):
        self.target, self.fn, self.args = target, fn, args

    def __repr__(self):
        return f"{self.target} = @{self.fn}({', '.join(map(repr, self.args))})"


_TOKEN = re.compile(r"""\s*(?:(?P<str>"(?:[^"\\]|\\.)*")|(?P<sym>`[A-Za-z_][\w.]*)|(?P<num>-?\d+(?:\.\d*)?(?:[eE][-+]?\d+)?)|
                    (?P<name>[A-Za-z_][\w.]*)|(?P<at>@)|(?P<p>[(),:;=?<>]))""", re.X)


def _tokens(text):
    text = re.sub(r"//[^\n]*", "", text)
    pos, out = 0, []
    while pos < len(text):
        m = _TOKEN.match(text, pos)
        if not m or m.end() == pos:
            if text[pos:].strip() == "":
                break
            raise SyntaxError(f"HorseIR: cannot read {text[pos:pos + 20]!r}")
        pos = m.end()
        kind = m.lastgroup
        out.append((kind, m.group(kind)))
    return out


class _Parser:
    def __init__(self, text):
        self.t = _tokens(text)
        self.k = 0

    def peek(self, n=0):
        return self.t[self.k + n] if self.k + n < len(self.t) else (None, None)

    def take(self, want=None):
        tok = self.peek()
        if want is not None and tok[1] != want:
            raise SyntaxError(f"HorseIR: expected {want!r}, got {tok[1]!r}")
        self.k += 1
        return tok

    def typ(self):
        kind, v = self.take()
        if v == "?":
            return None
        if v == "enum" and self.peek()[1] == "<":      # enum<i32,i32>: the element types do not matter here
            while self.take()[1] != ">":
                pass
            return "enum"
        return v

    def literal_tail(self, vals):
        self.take(":")
        t = self.typ()
        return t

    def arg(self):
        kind, v = self.peek()
        if kind == "at":
            self.take()
            return Fn(self.take()[1])
        if kind == "sym":
            self.take()
            name = v[1:]
            if self.peek()[1] == ":":
                self.take(); self.typ()
            return Lit([name], "sym", False)
        if kind == "str":
            self.take()
            s = bytes(v[1:-1], "utf-8").decode("unicode_escape").encode("latin-1")
            t = self.literal_tail([s])
            return Lit([s], t or "str", False)
        if kind == "num":
            self.take()
            t = self.literal_tail(None)
            return Lit([v], t, False)
        if v == "(":
            self.take()
            vals = []
            while True:
                kind, v = self.take()
                if kind == "str":
                    vals.append(bytes(v[1:-1], "utf-8").decode("unicode_escape").encode("latin-1"))
                elif kind == "sym":
                    vals.append(v[1:])
                else:
                    vals.append(v)
                if self.take()[1] == ")":
                    break
            t = self.literal_tail(vals)
            return Lit(vals, t, True)
        if kind == "name":
            self.take()
            if v == "check_cast":                         # check_cast(x, type): x
                self.take("(")
                inner = self.call_or_arg()
                self.take(",")
                self.typ()
                self.take(")")
                return inner
            return Var(v)
        raise SyntaxError(f"HorseIR: unexpected {v!r}")

    def call_or_arg(self):
        if self.peek()[0] == "at":
            self.take()
            fn = self.take()[1]
            self.take("(")
            args = []
            if self.peek()[1] != ")":
                args.append(self.arg())
                while self.peek()[1] == ",":
                    self.take()
                    args.append(self.arg())
            self.take(")")
            return ("call", fn, args)
        return self.arg()

    def program(self):
        stmts, ret = [], []
        while self.peek()[0] is not None:
            kind, v = self.peek()
            if v in ("module", "import", "def", "{", "}") or kind == "p" and v in ";":
                self.take()
                continue
            if v == "return":
                self.take()
                ret.append(self.take()[1])
                while self.peek()[1] == ",":
                    self.take(); ret.append(self.take()[1])
                self.take(";")
                continue
            name = self.take()[1]
            if self.peek()[1] == ":":
                self.take(); self.typ()
            self.take("=")
            rhs = self.call_or_arg()
            self.take(";")
            if isinstance(rhs, tuple):
                stmts.append(Stmt(name, rhs[1], rhs[2]))
            elif isinstance(rhs, Var):
                stmts.append(Stmt(name, "copy", [rhs]))
            else:
                stmts.append(Stmt(name, "literal", [rhs]))
        return stmts, ret


def parse(text):
    """HorseIR statement text -> (statements, returned names).  Takes the straight-line form of
    a method body: `t:type = @fn(args);` with variables, literals (`730:i32`, `0.05:f64`,
    `(1,2):i32`, `` `eq:sym ``, `"x":str`), function arguments (`@sum`) and check_cast."""
    text = re.sub(r"\bmodule\s+\w+\s*\{|\bimport\s+[\w.*]+\s*;|\bdef\s+\w+\s*\([^)]*\)\s*:\s*[\w?]+\s*\{", " ", text)
    text = text.replace("}", " ")
    return _Parser(text).program()


# ---------------------------------------------------------------------------------- planning
class Plan:
    """The statements with fusion decided: `ops` runs in order; each op is a statement to run
    as it is ('stmt', s), or a fused form over trees of inlined statements."""

    def __init__(self, stmts, ret, ops):
        self.stmts, self.ret, self.ops = stmts, ret, ops

    def kinds(self):
        return [op[0] for op in self.ops]


def _uses(stmts, ret):
    u = {}
    for s in stmts:
        for a in s.args:
            if isinstance(a, Var):
                u[a.name] = u.get(a.name, 0) + 1
    for r in ret:
        u[r] = u.get(r, 0) + 1
    return u


def plan(stmts, ret, fuse=True):
    if not fuse:
        return Plan(stmts, ret, [("stmt", s) for s in stmts])
    defs = {s.target: s for s in stmts}
    uses = _uses(stmts, ret)
    ew = {s.target for s in stmts if s.fn in EW and len(s.args) == 2
          and all(isinstance(a, Var) or (isinstance(a, Lit) and not a.vector) for a in s.args)}
    # users of each statement's value
    user = {}
    for s in stmts:
        for a in s.args:
            if isinstance(a, Var):
                user.setdefault(a.name, []).append(s)
    # compresses whose only use is a reduction go into that FA reduction
    to_reduce = {s.args[0].name for s in stmts if s.fn in REDUCE and len(s.args) == 1 and isinstance(s.args[0], Var)
                 and s.args[0].name in defs and defs[s.args[0].name].fn == "compress" and uses.get(s.args[0].name) == 1}

    def inlinable(name):
        if name not in ew or uses.get(name, 0) != 1 or name not in user:
            return False                               # several uses, or returned
        u = user[name][0]
        if u.target in ew or u.fn in REDUCE:
            return True
        if u.fn == "compress":
            # both operands of a compress that feeds an FA reduction; otherwise only the value
            # compressed (an FE map inside the compress) -- the mask must exist
            return u.target in to_reduce or (isinstance(u.args[1], Var) and u.args[1].name == name)
        return False

    absorbed = set()

    def tree(arg):
        """expression tree of an argument: ('op', fn, l, r) | ('var', name) | ('lit', Lit)"""
        if isinstance(arg, Lit):
            return ("lit", arg)
        if inlinable(arg.name):
            s = defs[arg.name]
            absorbed.add(arg.name)
            return ("op", s.fn, tree(s.args[0]), tree(s.args[1]))
        return ("var", arg.name)

    ops = []
    done = set()
    mask_groups = {}
    for s in stmts:
        if s.fn == "compress" and len(s.args) == 2 and isinstance(s.args[0], Var) and s.target not in to_reduce:
            mask_groups.setdefault(s.args[0].name, []).append(s)
    # masks used by nothing but their group's compresses: the compress evaluates the mask's tree
    # as its condition in its count pass (FA compress scan of one condition), no mask is built
    cond_masks = {m for m, g in mask_groups.items() if m in ew and uses.get(m, 0) == len(g)
                  and all(isinstance(a, Var) and a.name == m for a in (x.args[0] for x in g))}
    for s in stmts:
        if s.target in done:
            continue
        # FP: raze(each(f, each_right(index, v, g)))
        if s.fn == "raze" and isinstance(s.args[0], Var):
            e1 = defs.get(s.args[0].name)
            if e1 is not None and e1.fn == "each" and uses.get(e1.target) == 1 and isinstance(e1.args[0], Fn) \
                    and e1.args[0].name in GROUP_FN and isinstance(e1.args[1], Var):
                e2 = defs.get(e1.args[1].name)
                if e2 is not None and e2.fn == "each_right" and uses.get(e2.target) == 1 and \
                        isinstance(e2.args[0], Fn) and e2.args[0].name == "index":
                    ops = [op for op in ops if not (op[0] == "stmt" and op[1].target in (e1.target, e2.target))]
                    done.update((e1.target, e2.target))
                    ops.append(("group_reduce", s.target, GROUP_FN[e1.args[0].name], e2.args[1], e2.args[2]))
                    done.add(s.target)
                    continue
        # FA reduction
        if s.fn in REDUCE and len(s.args) == 1 and isinstance(s.args[0], Var):
            a = s.args[0].name
            d = defs.get(a)
            if a in to_reduce and isinstance(d.args[0], Var):
                absorbed.add(a)
                ops.append(("reduce", s.target, REDUCE[s.fn], tree(d.args[0]), tree(d.args[1]), s))
                done.update((s.target, a))
                continue
            if inlinable(a):
                ops.append(("reduce", s.target, REDUCE[s.fn], None, tree(s.args[0]), s))
                done.add(s.target)
                continue
        # a mask that becomes its compresses' condition is not computed here
        if s.target in cond_masks:
            absorbed.add(s.target)
            done.add(s.target)
            continue
        # FA compress of one mask
        if s.fn == "compress" and isinstance(s.args[0], Var) and s.args[0].name in mask_groups:
            mname = s.args[0].name
            group = [g for g in mask_groups.pop(mname) if g.target not in done]
            if len(group) >= 2 or mname in cond_masks or \
                    (len(group) == 1 and isinstance(group[0].args[1], Var) and inlinable(group[0].args[1].name)):
                plain_ys, maps = [], []
                for g in group:
                    y = g.args[1]
                    if isinstance(y, Var) and inlinable(y.name) and len(maps) < 2:
                        maps.append((g.target, tree(y)))
                    else:
                        plain_ys.append((g.target, y))
                    done.add(g.target)
                cond = None
                if mname in cond_masks:
                    d = defs[mname]
                    cond = ("op", d.fn, tree(d.args[0]), tree(d.args[1]))
                ops.append(("compress", mname, plain_ys, maps, cond))
                continue
        if s.target in ew and not inlinable(s.target):
            t = ("op", s.fn, tree(s.args[0]), tree(s.args[1]))
            if t[2][0] == "op" or t[3][0] == "op":
                ops.append(("map", s.target, t, s))
            else:
                ops.append(("stmt", s))
            done.add(s.target)
            continue
        ops.append(("stmt", s))
        done.add(s.target)
    ops = [op for op in ops if not (op[0] == "stmt" and op[1].target in absorbed)]
    return Plan(stmts, ret, ops)


# ---------------------------------------------------------------------------------- execution
_INT_RANK = {H_B: 0, H_J: 1, H_H: 2, H_I: 3, H_L: 4, H_F: 5, H_E: 6}


class _Code:
    """A tree's postfix code over the distinct columns it reads."""

    def __init__(self, env):
        self.env, self.cols, self.idx, self.code = env, [], {}, []

    def col(self, name):
        if name not in self.idx:
            self.idx[name] = len(self.cols)
            self.cols.append(self.env[name])
        return self.idx[name]


def _tree_type(t, env):
    """The plain statements' result type of a tree, or None when the descriptors cannot compute
    it as they do."""
    if t[0] == "lit":
        return TYPES.get(t[1].typ)
    if t[0] == "var":
        return env[t[1]].contents.typ
    fn, a, b = t[1], _tree_type(t[2], env), _tree_type(t[3], env)
    if a is None or b is None:
        return None
    if fn in CMP:
        num = lambda x: x in _INT_RANK
        if num(a) and num(b):
            return H_B
        return H_B if a == b and a in (H_C, H_Q, H_D, H_M, H_Z, H_U, H_W, H_T) and fn in ("eq", "neq") else None
    if fn in LOGIC:
        return H_B if a == H_B and b == H_B else None
    if a not in _INT_RANK or b not in _INT_RANK or H_B in (a, b):
        return None
    t_out = a if _INT_RANK[a] >= _INT_RANK[b] else b
    return t_out if t_out in (H_I, H_L, H_E) else None


def _emit(t, code):
    if t[0] == "lit":
        lit = t[1]
        v = lit.vals[0]
        if TYPES.get(lit.typ) in (H_F, H_E):
            code.code.append(("F", float(v)))
        else:
            code.code.append(("I", int(float(v)) if "." in str(v) else int(v)))
        return
    if t[0] == "var":
        code.code.append(("COL", code.col(t[1])))
        return
    _emit(t[2], code)
    _emit(t[3], code)
    code.code.append(EW[t[1]])


def _lit_node(B, lit):
    t = TYPES.get(lit.typ)
    if lit.typ == "sym":
        if lit.vector:
            return B.host.vec(np.array([B.symbol(v) for v in lit.vals], np.int64), H_Q)
        return B.sym(lit.vals[0])
    if t == H_S:
        return B.strs(lit.vals)
    if t is None:
        raise TypeError(f"HorseIR literal of type {lit.typ}")
    from .vnode import NP_OF
    dt = NP_OF[t]
    arr = np.array([float(v) if np.dtype(dt).kind == "f" else int(v) for v in lit.vals], dtype=dt)
    return B.host.vec(arr, t) if (lit.vector and len(arr) > 1) else B.scalar(arr[0], t)


def _arg_node(B, env, a):
    if isinstance(a, Var):
        if a.name not in env:
            raise NameError(f"HorseIR: {a.name} used before it is set")
        return env[a.name]
    if isinstance(a, Lit):
        return _lit_node(B, a)
    raise TypeError("HorseIR: a function where a value is expected")


def _run_stmt(B, env, s):
    if s.fn == "literal":
        env[s.target] = _lit_node(B, s.args[0])
        return
    if s.fn == "copy":
        env[s.target] = env[s.args[0].name]
        return
    if s.fn == "list":
        env[s.target] = B.list([_arg_node(B, env, a) for a in s.args])
        return
    if s.fn in ("each", "each_right"):
        f = s.args[0]
        if not isinstance(f, Fn) or f.name not in FN_ARG:
            raise NotImplementedError(f"HorseIR: @{s.fn} of {getattr(f, 'name', f)}")
        if s.fn == "each":
            env[s.target] = B.each(_arg_node(B, env, s.args[1]), FN_ARG[f.name])
        else:
            env[s.target] = B.each_right(_arg_node(B, env, s.args[1]), _arg_node(B, env, s.args[2]), FN_ARG[f.name])
        return
    if s.fn not in PFN:
        raise NotImplementedError(f"HorseIR: @{s.fn}")
    env[s.target] = getattr(B, PFN[s.fn])(*[_arg_node(B, env, a) for a in s.args])


def _run_tree_plain(B, env, t):
    """A tree as the plain statements it came from (for trees the descriptors do not take)."""
    if t[0] == "lit":
        return _lit_node(B, t[1])
    if t[0] == "var":
        return env[t[1]]
    return getattr(B, PFN[t[1]])(_run_tree_plain(B, env, t[2]), _run_tree_plain(B, env, t[3]))


def _leaf_lens(t, env, out):
    if t[0] == "var":
        out.add(env[t[1]].contents.len)
    elif t[0] == "op":
        _leaf_lens(t[2], env, out)
        _leaf_lens(t[3], env, out)
    return out


def _types_ok(t, env):
    if _tree_type(t, env) is None:
        return False
    return t[0] != "op" or (_types_ok(t[2], env) and _types_ok(t[3], env))


def _fusable(t, env, *more):
    """The descriptors compute the tree(s) as the statements do: every type is one they take,
    and the columns read are vectors of one length (a one-row value stays a statement)."""
    lens = set()
    for x in (t,) + more:
        if x is None:
            continue
        if not _types_ok(x, env):
            return False
        _leaf_lens(x, env, lens)
    return len(lens) <= 1 and all(n > 1 for n in lens)


def run(program, B, env=None, fuse=True):
    """Execute a HorseIR program (text or parse() result) on backend B.  env: names bound
    before the first statement (columns, tables).  fuse=True applies the fusion above (GPU
    backend only); fuse=False runs one primitive per statement, as the interpreter does.
    Returns (values of the returned names, the plan)."""
    stmts, ret = parse(program) if isinstance(program, str) else program
    p = plan(stmts, ret, fuse=fuse)
    env = dict(env or {})
    for op in p.ops:
        kind = op[0]
        if kind == "stmt":
            _run_stmt(B, env, op[1])
        elif kind == "map":
            _, target, t, s = op
            if _fusable(t, env):
                code = _Code(env)
                _emit(t, code)
                env[target] = B.fused_map(_tree_type(t, env), code.cols, expr(code.code))
            else:
                env[target] = _run_tree_plain(B, env, t)
        elif kind == "reduce":
            _, target, rk, cond, body, s = op
            if _fusable(body, env, cond) and body[0] != "lit":
                code = _Code(env)
                ce = None
                if cond is not None:
                    _emit(cond, code)
                    ce = expr(code.code)
                    code.code = []
                _emit(body, code)
                env[target] = B.fused_reduce(rk, code.cols, ce, expr(code.code))
            else:
                val = _run_tree_plain(B, env, body)
                if cond is not None:
                    val = B.pfnCompress(_run_tree_plain(B, env, cond), val)
                env[target] = getattr(B, PFN[s.fn])(val)
        elif kind == "compress":
            _, mask, plain_ys, maps, cond = op
            if cond is not None:
                # the rows: the length of the condition's leaves (it has a column leaf, being a
                # vector mask's tree); the condition runs fused when its columns allow it
                lens = _leaf_lens(cond, env, set())
                n = next(iter(lens)) if len(lens) == 1 else None
                if n is None or not _fusable(cond, env):
                    env[mask] = _run_tree_plain(B, env, cond)
                    cond = None
            if cond is None:
                n = env[mask].contents.len
            ok_maps = [(tg, t) for tg, t in maps
                       if t[0] == "op" and _fusable(t, env) and _leaf_lens(t, env, set()) <= {n}]
            bad = [(tg, t) for tg, t in maps if (tg, t) not in ok_maps]
            for tg, t in bad:
                plain_ys.append((tg, ("tree", t)))
            ys = []
            for tg, y in plain_ys:
                ys.append(_run_tree_plain(B, env, y[1]) if isinstance(y, tuple) else _arg_node(B, env, y))
            if cond is not None:
                code = _Code(env)
                _emit(cond, code)
                ce = expr(code.code)
                exprs = []
                for tg, t in ok_maps:
                    code.code = []
                    _emit(t, code)
                    exprs.append((_tree_type(t, env), expr(code.code)))
                if ok_maps:
                    zs, zm = B.fused_compress_map(ce, ys, exprs, code.cols)
                    for (tg, _), z in zip(ok_maps, zm):
                        env[tg] = z
                else:
                    zs = B.fused_compress(ys, code.cols, ce)
                for (tg, _), z in zip(plain_ys, zs):
                    env[tg] = z
                continue
            m = env[mask]
            if ok_maps:
                code = _Code(env)
                exprs = []
                for tg, t in ok_maps:
                    code.code = []
                    _emit(t, code)
                    exprs.append((_tree_type(t, env), expr(code.code)))
                zs, zm = B.compress_map(m, ys, exprs, code.cols)
                for (tg, _), z in zip(ok_maps, zm):
                    env[tg] = z
            elif len(ys) > 1:
                zs = B.compress_many(m, ys)
            else:
                zs = [B.pfnCompress(m, ys[0])]
            for (tg, _), z in zip(plain_ys, zs):
                env[tg] = z
        elif kind == "group_reduce":
            _, target, gk, v, g = op
            env[target] = B.group_reduce(gk, _arg_node(B, env, v), _arg_node(B, env, g))
        else:
            raise AssertionError(kind)
    return [env[r] for r in p.ret], p
