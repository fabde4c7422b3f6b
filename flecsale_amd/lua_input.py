"""Lua input files of the reference's apps (`-f input.lua`; apps/hydro/
inputs.h:99-186, the shock_box_2d.lua / shock_box_3d.lua files beside the
apps), read without a Lua installation.

The reference loads the file into an embedded Lua state (ristra's, not
vendored) and reads the global table `hydro`: prefix, postfix, output_freq,
CFL, final_time, max_steps, eos = {type = "ideal_gas", gas_constant,
specific_heat} and the initial-conditions function ics(x, y[, z], t) ->
density, {velocity}, pressure.  This module evaluates the subset of Lua such
files are written in: comments, number / string / boolean / nil literals,
table constructors, the full operator set with Lua's precedence (`^` and
`..` right-associative, `and` / `or` returning operands), function literals
and calls, indexing, local and global assignment, if / elseif / else, while,
numeric for, return with several values, and a `math` table.  Numbers are
doubles parsed with the same correctly rounded conversion as strtod, so a
literal means the same value it does to the reference.
"""
import math
import re

__all__ = ["load", "LuaError", "hydro_inputs"]


class LuaError(ValueError):
    pass


# ------------------------------------------------------------------ lexer
_TOKEN = re.compile(r"""
    (?P<ws>\s+)
  | (?P<comment>--\[(?P<eq>=*)\[.*?\](?P=eq)\]|--[^\n]*)
  | (?P<num>0[xX][0-9a-fA-F]+|(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?)
  | (?P<str>"(?:\\.|[^"\\])*"|'(?:\\.|[^'\\])*')
  | (?P<name>[A-Za-z_][A-Za-z_0-9]*)
  | (?P<op>\.\.\.|\.\.|==|~=|<=|>=|//|[-+*/%^\#<>=(){}\[\];:,.])
""", re.S | re.X)

_KEYWORDS = {"and", "break", "do", "else", "elseif", "end", "false", "for", "function", "if", "in", "local",
             "nil", "not", "or", "repeat", "return", "then", "true", "until", "while"}
_ESC = {"n": "\n", "t": "\t", "r": "\r", "\\": "\\", '"': '"', "'": "'", "a": "\a", "b": "\b", "f": "\f",
        "v": "\v", "0": "\0"}


def _tokens(src):
    out, pos = [], 0
    while pos < len(src):
        m = _TOKEN.match(src, pos)
        if not m:
            raise LuaError(f"unexpected character {src[pos]!r} at offset {pos}")
        pos = m.end()
        kind = m.lastgroup
        if kind in ("ws", "comment", "eq"):
            continue
        text = m.group(kind)
        if kind == "num":
            out.append(("num", float(int(text, 16)) if text[:2] in ("0x", "0X") else float(text)))
        elif kind == "str":
            body, s, i = text[1:-1], [], 0
            while i < len(body):
                if body[i] == "\\" and i + 1 < len(body):
                    s.append(_ESC.get(body[i + 1], body[i + 1]))
                    i += 2
                else:
                    s.append(body[i])
                    i += 1
            out.append(("str", "".join(s)))
        elif kind == "name":
            out.append(("kw", text) if text in _KEYWORDS else ("name", text))
        else:
            out.append(("op", text))
    out.append(("eof", None))
    return out


# ----------------------------------------------------------------- parser
# AST nodes are tuples: (kind, ...)
_BINARY = {  # precedence (left, right): Lua 5.3 reference manual, 3.4.8
    "or": (1, 1), "and": (2, 2),
    "<": (3, 3), ">": (3, 3), "<=": (3, 3), ">=": (3, 3), "~=": (3, 3), "==": (3, 3),
    "..": (9, 8), "+": (10, 10), "-": (10, 10),
    "*": (11, 11), "/": (11, 11), "//": (11, 11), "%": (11, 11),
    "^": (14, 13),
}
_UNARY_PRIORITY = 12


class _Parser:
    def __init__(self, src):
        self.t = _tokens(src)
        self.i = 0

    def peek(self, k=0):
        return self.t[self.i + k]

    def next(self):
        tok = self.t[self.i]
        self.i += 1
        return tok

    def check(self, kind, value=None):
        tok = self.peek()
        return tok[0] == kind and (value is None or tok[1] == value)

    def accept(self, kind, value=None):
        if self.check(kind, value):
            return self.next()
        return None

    def expect(self, kind, value=None):
        tok = self.next()
        if tok[0] != kind or (value is not None and tok[1] != value):
            raise LuaError(f"expected {value or kind}, found {tok[1]!r}")
        return tok

    # block: statements until one of the terminators
    def block(self, *ends):
        body = []
        while not (self.check("eof") or any(self.check("kw", e) for e in ends)):
            if self.accept("kw", "return"):
                exprs = [] if (self.check("kw", "end") or self.check("eof") or self.check("op", ";") or
                               any(self.check("kw", e) for e in ends)) else self.exprlist()
                self.accept("op", ";")
                body.append(("return", exprs))
                break
            st = self.statement()
            if st is not None:
                body.append(st)
        return body

    def statement(self):
        if self.accept("op", ";"):
            return None
        if self.accept("kw", "local"):
            if self.accept("kw", "function"):
                name = self.expect("name")[1]
                return ("local", [name], [self.funcbody()])
            names = [self.expect("name")[1]]
            while self.accept("op", ","):
                names.append(self.expect("name")[1])
            exprs = self.exprlist() if self.accept("op", "=") else []
            return ("local", names, exprs)
        if self.accept("kw", "if"):
            arms = [(self.expr(), None)]
            self.expect("kw", "then")
            arms[0] = (arms[0][0], self.block("elseif", "else", "end"))
            other = []
            while True:
                if self.accept("kw", "elseif"):
                    c = self.expr()
                    self.expect("kw", "then")
                    arms.append((c, self.block("elseif", "else", "end")))
                elif self.accept("kw", "else"):
                    other = self.block("end")
                    self.expect("kw", "end")
                    break
                else:
                    self.expect("kw", "end")
                    break
            return ("if", arms, other)
        if self.accept("kw", "while"):
            c = self.expr()
            self.expect("kw", "do")
            b = self.block("end")
            self.expect("kw", "end")
            return ("while", c, b)
        if self.accept("kw", "do"):
            b = self.block("end")
            self.expect("kw", "end")
            return ("do", b)
        if self.accept("kw", "for"):
            var = self.expect("name")[1]
            self.expect("op", "=")
            start = self.expr()
            self.expect("op", ",")
            stop = self.expr()
            step = self.expr() if self.accept("op", ",") else ("const", 1.0)
            self.expect("kw", "do")
            b = self.block("end")
            self.expect("kw", "end")
            return ("for", var, start, stop, step, b)
        if self.accept("kw", "function"):
            target = ("name", self.expect("name")[1])
            while self.accept("op", "."):
                target = ("index", target, ("const", self.expect("name")[1]))
            return ("assign", [target], [self.funcbody()])
        if self.accept("kw", "break"):
            return ("break",)
        # assignment or call
        first = self.suffixed()
        if self.check("op", "=") or self.check("op", ","):
            targets = [first]
            while self.accept("op", ","):
                targets.append(self.suffixed())
            self.expect("op", "=")
            for tg in targets:
                if tg[0] not in ("name", "index"):
                    raise LuaError("cannot assign to this expression")
            return ("assign", targets, self.exprlist())
        if first[0] != "call":
            raise LuaError("syntax error: expression is not a statement")
        return ("callstat", first)

    def funcbody(self):
        self.expect("op", "(")
        params = []
        if not self.check("op", ")"):
            while True:
                if self.accept("op", "..."):
                    params.append("...")
                    break
                params.append(self.expect("name")[1])
                if not self.accept("op", ","):
                    break
        self.expect("op", ")")
        body = self.block("end")
        self.expect("kw", "end")
        return ("function", params, body)

    def exprlist(self):
        out = [self.expr()]
        while self.accept("op", ","):
            out.append(self.expr())
        return out

    def expr(self, limit=0):
        tok = self.peek()
        if (tok[0] == "kw" and tok[1] == "not") or (tok[0] == "op" and tok[1] in ("-", "#")):
            self.next()
            left = ("unop", tok[1], self.expr(_UNARY_PRIORITY))
        else:
            left = self.simple()
        while True:
            tok = self.peek()
            op = tok[1] if tok[0] in ("op", "kw") else None
            if op not in _BINARY or _BINARY[op][0] <= limit:
                return left
            self.next()
            right = self.expr(_BINARY[op][1])
            left = ("binop", op, left, right)

    def simple(self):
        tok = self.peek()
        if tok[0] == "num" or tok[0] == "str":
            self.next()
            return ("const", tok[1])
        if tok[0] == "kw" and tok[1] in ("nil", "true", "false"):
            self.next()
            return ("const", {"nil": None, "true": True, "false": False}[tok[1]])
        if tok[0] == "op" and tok[1] == "{":
            return self.table()
        if tok[0] == "kw" and tok[1] == "function":
            self.next()
            return self.funcbody()
        if tok[0] == "op" and tok[1] == "...":
            self.next()
            return ("vararg",)
        return self.suffixed()

    def primary(self):
        if self.accept("op", "("):
            e = self.expr()
            self.expect("op", ")")
            return ("paren", e)
        return ("name", self.expect("name")[1])

    def suffixed(self):
        e = self.primary()
        while True:
            if self.accept("op", "."):
                e = ("index", e, ("const", self.expect("name")[1]))
            elif self.accept("op", "["):
                k = self.expr()
                self.expect("op", "]")
                e = ("index", e, k)
            elif self.check("op", "("):
                self.next()
                args = [] if self.check("op", ")") else self.exprlist()
                self.expect("op", ")")
                e = ("call", e, args)
            elif self.check("op", "{") or self.check("str"):
                arg = self.table() if self.check("op", "{") else ("const", self.next()[1])
                e = ("call", e, [arg])
            else:
                return e

    def table(self):
        self.expect("op", "{")
        items = []   # (key expr or None for positional, value expr)
        while not self.check("op", "}"):
            if self.accept("op", "["):
                k = self.expr()
                self.expect("op", "]")
                self.expect("op", "=")
                items.append((k, self.expr()))
            elif self.peek()[0] == "name" and self.peek(1) == ("op", "="):
                k = ("const", self.next()[1])
                self.next()
                items.append((k, self.expr()))
            else:
                items.append((None, self.expr()))
            if not (self.accept("op", ",") or self.accept("op", ";")):
                break
        self.expect("op", "}")
        return ("table", items)


# --------------------------------------------------------------- evaluator
class LuaTable(dict):
    """A Lua table: positional entries at keys 1.0, 2.0, ..."""

    def array(self):
        out, i = [], 1.0
        while i in self:
            out.append(self[i])
            i += 1.0
        return out


class LuaFunction:
    def __init__(self, params, body, scope, interp):
        self.params, self.body, self.scope, self.interp = params, body, scope, interp

    def __call__(self, *args):
        frame = {}
        for i, p in enumerate(self.params):
            if p == "...":
                frame["..."] = list(args[i:])
                break
            frame[p] = args[i] if i < len(args) else None
        try:
            self.interp.run(self.body, [frame] + self.scope)
        except _Return as r:
            return r.values
        return []


class _Return(Exception):
    def __init__(self, values):
        self.values = values


class _Break(Exception):
    pass


def _truth(v):
    return v is not None and v is not False


def _num(v):
    if isinstance(v, bool) or not isinstance(v, (int, float, str)):
        raise LuaError(f"attempt to perform arithmetic on a {type(v).__name__} value")
    return float(v)


def _math_table():
    t = LuaTable()
    for name in ("sqrt", "sin", "cos", "tan", "asin", "acos", "exp", "log", "floor", "ceil"):
        t[name] = (lambda f: (lambda *a: [float(f(_num(a[0])))]))(getattr(math, name))
    t["abs"] = lambda *a: [abs(_num(a[0]))]
    t["atan"] = lambda *a: [math.atan2(_num(a[0]), _num(a[1])) if len(a) > 1 else math.atan(_num(a[0]))]
    t["max"] = lambda *a: [max(_num(x) for x in a)]
    t["min"] = lambda *a: [min(_num(x) for x in a)]
    t["pi"] = math.pi
    t["huge"] = math.inf
    return t


class _Interp:
    def __init__(self):
        self.globals = {"math": _math_table(), "print": lambda *a: [], "tonumber": lambda *a: [float(a[0])],
                        "tostring": lambda *a: [str(a[0])]}

    def run(self, block, scope):
        for st in block:
            self.stat(st, scope)

    def lookup(self, name, scope):
        for frame in scope:
            if name in frame:
                return frame
        return None

    def stat(self, st, scope):
        kind = st[0]
        if kind == "local":
            vals = self.exprs(st[2], scope)
            for i, n in enumerate(st[1]):
                scope[0][n] = vals[i] if i < len(vals) else None
        elif kind == "assign":
            vals = self.exprs(st[2], scope)
            for i, tg in enumerate(st[1]):
                v = vals[i] if i < len(vals) else None
                if tg[0] == "name":
                    frame = self.lookup(tg[1], scope)
                    (frame if frame is not None else self.globals)[tg[1]] = v
                else:
                    t = self.eval(tg[1], scope)
                    k = self.eval(tg[2], scope)
                    t[float(k) if isinstance(k, (int, float)) and not isinstance(k, bool) else k] = v
        elif kind == "callstat":
            self.call(st[1], scope)
        elif kind == "if":
            for cond, body in st[1]:
                if _truth(self.eval(cond, scope)):
                    self.run(body, [{}] + scope)
                    return
            self.run(st[2], [{}] + scope)
        elif kind == "while":
            try:
                while _truth(self.eval(st[1], scope)):
                    self.run(st[2], [{}] + scope)
            except _Break:
                pass
        elif kind == "do":
            self.run(st[1], [{}] + scope)
        elif kind == "for":
            v, stop, step = _num(self.eval(st[2], scope)), _num(self.eval(st[3], scope)), _num(self.eval(st[4], scope))
            try:
                while (step > 0 and v <= stop) or (step <= 0 and v >= stop):
                    self.run(st[5], [{st[1]: v}] + scope)
                    v += step
            except _Break:
                pass
        elif kind == "return":
            raise _Return(self.exprs(st[1], scope))
        elif kind == "break":
            raise _Break()

    def exprs(self, exprs, scope):
        out = []
        for i, e in enumerate(exprs):
            if i == len(exprs) - 1 and e[0] in ("call", "vararg"):
                out.extend(self.call(e, scope) if e[0] == "call" else self.lookup("...", scope)["..."])
            else:
                out.append(self.eval(e, scope))
        return out

    def call(self, e, scope):
        f = self.eval(e[1], scope)
        if not callable(f):
            raise LuaError("attempt to call a non-function value")
        r = f(*self.exprs(e[2], scope))
        return list(r) if isinstance(r, (list, tuple)) else [r]

    def eval(self, e, scope):
        kind = e[0]
        if kind == "const":
            return e[1]
        if kind == "name":
            frame = self.lookup(e[1], scope)
            return (frame if frame is not None else self.globals).get(e[1])
        if kind == "paren":
            return self.eval(e[1], scope)
        if kind == "vararg":
            vals = self.lookup("...", scope)["..."]
            return vals[0] if vals else None
        if kind == "index":
            t = self.eval(e[1], scope)
            k = self.eval(e[2], scope)
            if not isinstance(t, dict):
                raise LuaError("attempt to index a non-table value")
            return t.get(float(k) if isinstance(k, (int, float)) and not isinstance(k, bool) else k)
        if kind == "call":
            r = self.call(e, scope)
            return r[0] if r else None
        if kind == "function":
            return LuaFunction(e[1], e[2], scope, self)
        if kind == "table":
            t, n = LuaTable(), 1.0
            for idx, (k, v) in enumerate(e[1]):
                if k is None:
                    if idx == len(e[1]) - 1 and v[0] in ("call", "vararg"):
                        for x in (self.call(v, scope) if v[0] == "call" else self.lookup("...", scope)["..."]):
                            t[n] = x
                            n += 1.0
                    else:
                        t[n] = self.eval(v, scope)
                        n += 1.0
                else:
                    key = self.eval(k, scope)
                    t[float(key) if isinstance(key, (int, float)) and not isinstance(key, bool) else key] = \
                        self.eval(v, scope)
            return t
        if kind == "unop":
            v = self.eval(e[2], scope)
            if e[1] == "not":
                return not _truth(v)
            if e[1] == "-":
                return -_num(v)
            return float(len(v.array()) if isinstance(v, LuaTable) else len(v))
        if kind == "binop":
            op = e[1]
            if op == "and":
                a = self.eval(e[2], scope)
                return self.eval(e[3], scope) if _truth(a) else a
            if op == "or":
                a = self.eval(e[2], scope)
                return a if _truth(a) else self.eval(e[3], scope)
            a, b = self.eval(e[2], scope), self.eval(e[3], scope)
            if op == "==":
                return a == b
            if op == "~=":
                return a != b
            if op == "..":
                fmt = lambda x: ("%.14g" % x) if isinstance(x, float) else str(x)
                return fmt(a) + fmt(b)
            if op in ("<", ">", "<=", ">="):
                if isinstance(a, str) and isinstance(b, str):
                    pass
                else:
                    a, b = _num(a), _num(b)
                return {"<": a < b, ">": a > b, "<=": a <= b, ">=": a >= b}[op]
            a, b = _num(a), _num(b)
            if op == "+":
                return a + b
            if op == "-":
                return a - b
            if op == "*":
                return a * b
            if op == "/":
                if b == 0:
                    return math.copysign(math.inf, a) * math.copysign(1.0, b) if a != 0 else math.nan
                return a / b
            if op == "%":
                return a - math.floor(a / b) * b
            if op == "//":
                return float(math.floor(a / b))
            if op == "^":
                return math.pow(a, b)
        raise LuaError(f"cannot evaluate {kind}")


def load(path_or_source, is_source=False):
    """Runs a Lua chunk; returns its global variables."""
    src = path_or_source if is_source else open(path_or_source).read()
    interp = _Interp()
    body = _Parser(src).block()
    try:
        interp.run(body, [{}])   # the chunk's own locals
    except _Return:
        pass
    return interp.globals


def hydro_inputs(path):
    """The `hydro` table of an apps/hydro input file, as inputs_t::load_lua
    reads it (apps/hydro/inputs.h:103-141): prefix, postfix, output_freq,
    CFL, final_time, max_steps, eos (gamma, cv) and ics(x, t) -> (density,
    velocity, pressure) with x a coordinate sequence."""
    g = load(path)
    h = g.get("hydro")
    if not isinstance(h, dict):
        raise LuaError(f"{path}: no table `hydro`")

    def need(t, key):
        if key not in t or t[key] is None:
            raise LuaError(f"{path}: hydro input `{key}` is missing")
        return t[key]

    eos = need(h, "eos")
    if need(eos, "type") != "ideal_gas":
        raise LuaError(f"Unknown eos type \"{eos['type']}\"")
    ics_f = need(h, "ics")

    def ics(x, t):
        r = ics_f(*[float(v) for v in x], float(t))
        if len(r) < 3:
            raise LuaError("ics must return density, velocity, pressure")
        v = r[1].array() if isinstance(r[1], LuaTable) else [r[1]]
        return float(r[0]), [float(c) for c in v], float(r[2])

    return dict(prefix=str(need(h, "prefix")), postfix=str(h.get("postfix", "")),
                output_freq=int(float(need(h, "output_freq"))), CFL=float(need(h, "CFL")),
                final_time=float(need(h, "final_time")), max_steps=int(float(need(h, "max_steps"))),
                eos=(float(need(eos, "gas_constant")), float(need(eos, "specific_heat"))), ics=ics)
