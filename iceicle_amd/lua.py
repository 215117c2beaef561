"""A Lua 5.4 subset interpreter for input decks (no reference counterpart:
the reference embeds the Lua C library through sol2,
solvers_lua_interface.hpp:402-694, and no Lua runtime exists here).

Covers what configuration scripts use: numbers (integers and floats),
strings, booleans, nil, tables with array and hash parts, closures with
upvalues and varargs, multiple assignment and multiple returns, `local`,
`if`/`elseif`/`else`, numeric and generic `for`, `while`, `repeat`,
`break`, `do` blocks, method calls, the arithmetic, comparison, logical,
concatenation and length operators with Lua's precedence and
associativity, and the standard functions decks call (`math.*`, `print`,
`type`, `tostring`, `tonumber`, `pairs`, `ipairs`, `select`, `error`,
`assert`, `string.format`/`rep`/`sub`/`len`/`upper`/`lower`, `table.insert`/
`remove`/`concat`/`unpack`).  Not covered: goto, coroutines, metatables,
integer overflow wrap-around, string methods through `s:method()`, and the
io/os libraries; using them raises LuaError.

`run(source)` executes a chunk and returns its return values; Lua tables
come back as LuaTable (`to_python` converts them), Lua functions as Python
callables.  Arithmetic and `math.*` accept numpy arrays, so a deck function
of the coordinates can be evaluated over many points in one call
(`vectorize`).
"""
import math
import re

import numpy as np


class LuaError(RuntimeError):
    pass


# ---------------------------------------------------------------- values

class LuaTable:
    """Lua table: integer keys 1..n and any other hashable keys in one dict."""

    def __init__(self, items=None):
        self.d = {} if items is None else dict(items)

    def get(self, k):
        return self.d.get(_key(k))

    def set(self, k, v):
        if k is None:
            raise LuaError("table index is nil")
        k = _key(k)
        if v is None:
            self.d.pop(k, None)
        else:
            self.d[k] = v

    def length(self):
        n = 0
        while (n + 1) in self.d:
            n += 1
        return n

    def array(self):
        return [self.d[i] for i in range(1, self.length() + 1)]

    def __getitem__(self, k):
        return self.get(k)

    def __contains__(self, k):
        return _key(k) in self.d

    def keys(self):
        return self.d.keys()

    def __repr__(self):
        return f"LuaTable({self.d!r})"


def _key(k):
    # float keys with integral values are the same key as the integer
    if isinstance(k, float) and k.is_integer():
        return int(k)
    return k


class LuaFunction:
    """A Lua closure, callable from Python with positional arguments."""

    def __init__(self, interp, params, vararg, body, env, name="?"):
        self.interp, self.params, self.vararg, self.body, self.env, self.name = interp, params, vararg, body, env, name

    def __call__(self, *args):
        r = self.interp.call(self, list(args))
        if not r:
            return None
        return r[0] if len(r) == 1 else tuple(r)

    def call_multi(self, args):
        return self.interp.call(self, list(args))

    def __repr__(self):
        return f"<lua function {self.name}>"


def to_python(v):
    """LuaTable -> list (a sequence 1..n and nothing else) or dict,
    recursively; other values unchanged."""
    if isinstance(v, LuaTable):
        n = v.length()
        if n > 0 and len(v.d) == n:
            return [to_python(x) for x in v.array()]
        if n == 0 and not v.d:
            return {}
        return {k: to_python(x) for k, x in v.d.items()}
    return v


def truthy(v):
    if isinstance(v, np.ndarray):
        raise LuaError("a condition on an array of points (evaluate point by point)")
    return v is not None and v is not False


# ---------------------------------------------------------------- lexer

KEYWORDS = {"and", "break", "do", "else", "elseif", "end", "false", "for", "function", "goto", "if", "in",
            "local", "nil", "not", "or", "repeat", "return", "then", "true", "until", "while"}
SYMBOLS = ["...", "..", "==", "~=", "<=", ">=", "//", "::", "<<", ">>", "+", "-", "*", "/", "%", "^", "#", "&",
           "~", "|", "<", ">", "=", "(", ")", "{", "}", "[", "]", ";", ":", ",", "."]
_NUM = re.compile(r"0[xX][0-9a-fA-F]+(\.[0-9a-fA-F]*)?([pP][+-]?\d+)?|(\d+\.?\d*|\.\d+)([eE][+-]?\d+)?")
_NAME = re.compile(r"[A-Za-z_][A-Za-z_0-9]*")
_ESC = {"n": "\n", "t": "\t", "r": "\r", "a": "\a", "b": "\b", "f": "\f", "v": "\v", "\\": "\\", '"': '"',
        "'": "'", "\n": "\n"}


def _long_bracket(src, i):
    """Level of a long bracket [==[ at i, or -1."""
    if src[i] != "[":
        return -1
    j = i + 1
    while j < len(src) and src[j] == "=":
        j += 1
    return j - i - 1 if j < len(src) and src[j] == "[" else -1


def tokenize(src):
    toks = []
    i, line, n = 0, 1, len(src)
    while i < n:
        c = src[i]
        if c == "\n":
            line += 1
            i += 1
            continue
        if c in " \t\r\f\v":
            i += 1
            continue
        if src.startswith("--", i):
            lv = _long_bracket(src, i + 2) if i + 2 < n else -1
            if lv >= 0:
                close = "]" + "=" * lv + "]"
                j = src.find(close, i)
                if j < 0:
                    raise LuaError(f"line {line}: unfinished long comment")
                line += src.count("\n", i, j)
                i = j + len(close)
            else:
                j = src.find("\n", i)
                i = n if j < 0 else j
            continue
        if c.isdigit() or (c == "." and i + 1 < n and src[i + 1].isdigit()):
            m = _NUM.match(src, i)
            s = m.group(0)
            if s[:2] in ("0x", "0X"):
                v = float.fromhex(s) if ("." in s or "p" in s or "P" in s) else int(s, 16)
            elif "." in s or "e" in s or "E" in s:
                v = float(s)
            else:
                v = int(s)
            toks.append(("number", v, line))
            i = m.end()
            continue
        if c.isalpha() or c == "_":
            m = _NAME.match(src, i)
            w = m.group(0)
            toks.append(("kw" if w in KEYWORDS else "name", w, line))
            i = m.end()
            continue
        if c in "\"'":
            j = i + 1
            out = []
            while True:
                if j >= n or src[j] == "\n":
                    raise LuaError(f"line {line}: unfinished string")
                ch = src[j]
                if ch == c:
                    break
                if ch == "\\":
                    e = src[j + 1]
                    if e in _ESC:
                        out.append(_ESC[e])
                        j += 2
                    elif e.isdigit():
                        m = re.match(r"\d{1,3}", src[j + 1:])
                        out.append(chr(int(m.group(0))))
                        j += 1 + len(m.group(0))
                    elif e == "x":
                        out.append(chr(int(src[j + 2:j + 4], 16)))
                        j += 4
                    elif e == "z":
                        j += 2
                        while j < n and src[j] in " \t\r\n":
                            j += 1
                    else:
                        raise LuaError(f"line {line}: invalid escape \\{e}")
                    continue
                out.append(ch)
                j += 1
            toks.append(("string", "".join(out), line))
            i = j + 1
            continue
        lv = _long_bracket(src, i)
        if lv >= 0:
            close = "]" + "=" * lv + "]"
            j = src.find(close, i)
            if j < 0:
                raise LuaError(f"line {line}: unfinished long string")
            s = src[i + lv + 2:j]
            if s.startswith("\n"):
                s = s[1:]
            toks.append(("string", s, line))
            line += src.count("\n", i, j)
            i = j + len(close)
            continue
        for s in SYMBOLS:
            if src.startswith(s, i):
                toks.append(("op", s, line))
                i += len(s)
                break
        else:
            raise LuaError(f"line {line}: unexpected character {c!r}")
    toks.append(("eof", None, line))
    return toks


# ---------------------------------------------------------------- parser
# AST nodes are tuples whose first entry names the node.

BINPRI = {"or": (1, 1), "and": (2, 2), "<": (3, 3), ">": (3, 3), "<=": (3, 3), ">=": (3, 3), "~=": (3, 3),
          "==": (3, 3), "|": (4, 4), "~": (5, 5), "&": (6, 6), "<<": (7, 7), ">>": (7, 7), "..": (9, 8),
          "+": (10, 10), "-": (10, 10), "*": (11, 11), "/": (11, 11), "//": (11, 11), "%": (11, 11),
          "^": (14, 13)}
UNARY_PRI = 12


class Parser:
    def __init__(self, toks):
        self.t, self.i = toks, 0

    def peek(self, k=0):
        return self.t[self.i + k]

    def next(self):
        tok = self.t[self.i]
        self.i += 1
        return tok

    def check(self, kind, val=None):
        tok = self.peek()
        return tok[0] == kind and (val is None or tok[1] == val)

    def accept(self, kind, val=None):
        if self.check(kind, val):
            return self.next()
        return None

    def expect(self, kind, val=None):
        tok = self.next()
        if tok[0] != kind or (val is not None and tok[1] != val):
            raise LuaError(f"line {tok[2]}: expected {val or kind}, found {tok[1]!r}")
        return tok

    # ---- blocks and statements
    def chunk(self):
        b = self.block()
        self.expect("eof")
        return b

    def block_end(self):
        tok = self.peek()
        return tok[0] == "eof" or (tok[0] == "kw" and tok[1] in ("end", "else", "elseif", "until"))

    def block(self):
        stmts = []
        while not self.block_end():
            if self.check("kw", "return"):
                self.next()
                exprs = [] if (self.block_end() or self.check("op", ";")) else self.exprlist()
                self.accept("op", ";")
                stmts.append(("return", exprs))
                break
            st = self.statement()
            if st is not None:
                stmts.append(st)
        return stmts

    def statement(self):
        tok = self.peek()
        line = tok[2]
        if self.accept("op", ";"):
            return None
        if tok[0] == "kw":
            w = tok[1]
            if w == "if":
                self.next()
                arms = []
                cond = self.expr()
                self.expect("kw", "then")
                arms.append((cond, self.block()))
                orelse = None
                while True:
                    if self.accept("kw", "elseif"):
                        c = self.expr()
                        self.expect("kw", "then")
                        arms.append((c, self.block()))
                    elif self.accept("kw", "else"):
                        orelse = self.block()
                        self.expect("kw", "end")
                        break
                    else:
                        self.expect("kw", "end")
                        break
                return ("if", arms, orelse)
            if w == "while":
                self.next()
                c = self.expr()
                self.expect("kw", "do")
                b = self.block()
                self.expect("kw", "end")
                return ("while", c, b)
            if w == "do":
                self.next()
                b = self.block()
                self.expect("kw", "end")
                return ("do", b)
            if w == "for":
                self.next()
                n1 = self.expect("name")[1]
                if self.accept("op", "="):
                    a = self.expr()
                    self.expect("op", ",")
                    b = self.expr()
                    c = self.expr() if self.accept("op", ",") else None
                    self.expect("kw", "do")
                    body = self.block()
                    self.expect("kw", "end")
                    return ("fornum", n1, a, b, c, body)
                names = [n1]
                while self.accept("op", ","):
                    names.append(self.expect("name")[1])
                self.expect("kw", "in")
                exprs = self.exprlist()
                self.expect("kw", "do")
                body = self.block()
                self.expect("kw", "end")
                return ("forin", names, exprs, body)
            if w == "repeat":
                self.next()
                b = self.block()
                self.expect("kw", "until")
                return ("repeat", b, self.expr())
            if w == "function":
                self.next()
                name = self.expect("name")[1]
                target = ("name", name)
                method = False
                full = name
                while self.check("op", ".") or self.check("op", ":"):
                    sep = self.next()[1]
                    key = self.expect("name")[1]
                    full += sep + key
                    target = ("index", target, ("const", key))
                    if sep == ":":
                        method = True
                        break
                f = self.funcbody(full, method)
                return ("assign", [target], [f])
            if w == "local":
                self.next()
                if self.accept("kw", "function"):
                    name = self.expect("name")[1]
                    return ("localfunc", name, self.funcbody(name, False))
                names = []
                while True:
                    names.append(self.expect("name")[1])
                    if self.accept("op", "<"):  # attribs <const> / <close>
                        self.expect("name")
                        self.expect("op", ">")
                    if not self.accept("op", ","):
                        break
                exprs = self.exprlist() if self.accept("op", "=") else []
                return ("local", names, exprs)
            if w == "break":
                self.next()
                return ("break",)
            if w == "goto":
                raise LuaError(f"line {line}: goto is not supported")
        # expression statement: call or assignment
        e = self.suffixedexp()
        if self.check("op", "=") or self.check("op", ","):
            targets = [e]
            while self.accept("op", ","):
                targets.append(self.suffixedexp())
            self.expect("op", "=")
            for t in targets:
                if t[0] not in ("name", "index"):
                    raise LuaError(f"line {line}: cannot assign to this expression")
            return ("assign", targets, self.exprlist())
        if e[0] not in ("call", "method"):
            raise LuaError(f"line {line}: syntax error near {self.peek()[1]!r}")
        return ("exprstat", e)

    def funcbody(self, name, method):
        self.expect("op", "(")
        params = ["self"] if method else []
        vararg = False
        if not self.check("op", ")"):
            while True:
                if self.accept("op", "..."):
                    vararg = True
                    break
                params.append(self.expect("name")[1])
                if not self.accept("op", ","):
                    break
        self.expect("op", ")")
        body = self.block()
        self.expect("kw", "end")
        return ("function", params, vararg, body, name)

    # ---- expressions
    def exprlist(self):
        es = [self.expr()]
        while self.accept("op", ","):
            es.append(self.expr())
        return es

    def expr(self, limit=0):
        tok = self.peek()
        if (tok[0] == "kw" and tok[1] == "not") or (tok[0] == "op" and tok[1] in ("-", "#", "~")):
            self.next()
            left = ("unop", tok[1], self.expr(UNARY_PRI))
        else:
            left = self.simpleexp()
        while True:
            tok = self.peek()
            op = tok[1] if tok[0] in ("op", "kw") else None
            if op not in BINPRI or BINPRI[op][0] <= limit:
                break
            self.next()
            right = self.expr(BINPRI[op][1])
            left = ("binop", op, left, right)
        return left

    def simpleexp(self):
        tok = self.peek()
        if tok[0] == "number" or tok[0] == "string":
            self.next()
            return ("const", tok[1])
        if tok[0] == "kw":
            if tok[1] == "nil":
                self.next()
                return ("const", None)
            if tok[1] == "true":
                self.next()
                return ("const", True)
            if tok[1] == "false":
                self.next()
                return ("const", False)
            if tok[1] == "function":
                self.next()
                return self.funcbody("anonymous", False)
        if tok[0] == "op" and tok[1] == "...":
            self.next()
            return ("vararg",)
        if tok[0] == "op" and tok[1] == "{":
            return self.table()
        return self.suffixedexp()

    def primaryexp(self):
        tok = self.next()
        if tok[0] == "name":
            return ("name", tok[1])
        if tok[0] == "op" and tok[1] == "(":
            e = self.expr()
            self.expect("op", ")")
            return ("paren", e)
        raise LuaError(f"line {tok[2]}: unexpected symbol {tok[1]!r}")

    def args(self):
        if self.check("string"):
            return [("const", self.next()[1])]
        if self.check("op", "{"):
            return [self.table()]
        self.expect("op", "(")
        a = [] if self.check("op", ")") else self.exprlist()
        self.expect("op", ")")
        return a

    def suffixedexp(self):
        e = self.primaryexp()
        while True:
            if self.accept("op", "."):
                e = ("index", e, ("const", self.expect("name")[1]))
            elif self.accept("op", "["):
                k = self.expr()
                self.expect("op", "]")
                e = ("index", e, k)
            elif self.accept("op", ":"):
                m = self.expect("name")[1]
                e = ("method", e, m, self.args())
            elif self.check("op", "(") or self.check("string") or self.check("op", "{"):
                e = ("call", e, self.args())
            else:
                return e

    def table(self):
        self.expect("op", "{")
        items = []  # ("pos", expr) | ("key", kexpr, vexpr)
        while not self.check("op", "}"):
            if self.check("op", "["):
                self.next()
                k = self.expr()
                self.expect("op", "]")
                self.expect("op", "=")
                items.append(("key", k, self.expr()))
            elif self.check("name") and self.peek(1)[0] == "op" and self.peek(1)[1] == "=":
                k = self.next()[1]
                self.next()
                items.append(("key", ("const", k), self.expr()))
            else:
                items.append(("pos", self.expr()))
            if not (self.accept("op", ",") or self.accept("op", ";")):
                break
        self.expect("op", "}")
        return ("table", items)


# ---------------------------------------------------------------- interpreter

class _Break(Exception):
    pass


class _Return(Exception):
    def __init__(self, values):
        self.values = values


class Scope:
    __slots__ = ("vars", "parent")

    def __init__(self, parent=None):
        self.vars, self.parent = {}, parent

    def lookup(self, name):
        s = self
        while s is not None:
            if name in s.vars:
                return s
            s = s.parent
        return None


def _num(v, what="arithmetic"):
    if isinstance(v, bool) or v is None:
        raise LuaError(f"attempt to perform {what} on a {lua_type(v)} value")
    if isinstance(v, str):
        try:
            return int(v) if re.fullmatch(r"\s*-?\d+\s*", v) else float(v)
        except ValueError:
            raise LuaError(f"attempt to perform {what} on a string value")
    return v


def lua_type(v):
    if v is None:
        return "nil"
    if isinstance(v, bool):
        return "boolean"
    if isinstance(v, (int, float, np.ndarray, np.floating, np.integer)):
        return "number"
    if isinstance(v, str):
        return "string"
    if isinstance(v, LuaTable):
        return "table"
    if callable(v):
        return "function"
    return "userdata"


def tostr(v):
    if v is None:
        return "nil"
    if v is True:
        return "true"
    if v is False:
        return "false"
    if isinstance(v, float):
        if v.is_integer() and abs(v) < 1e16:
            return f"{v:.1f}"
        return f"{v:.14g}"
    return str(v)


def _arith(op, a, b):
    a, b = _num(a), _num(b)
    if op == "+":
        return a + b
    if op == "-":
        return a - b
    if op == "*":
        return a * b
    if op == "/":
        if isinstance(a, np.ndarray) or isinstance(b, np.ndarray):
            return np.true_divide(a, b)
        if b == 0:
            if a == 0 or a != a:
                return math.nan
            return math.copysign(math.inf, a) * math.copysign(1.0, b)
        return a / b
    if op == "^":
        if isinstance(a, np.ndarray) or isinstance(b, np.ndarray):
            with np.errstate(invalid="ignore", divide="ignore"):
                return np.power(np.asarray(a, dtype=float), b)
        try:
            return math.pow(a, b)
        except ValueError:
            return math.nan
        except OverflowError:
            return math.inf
    if op == "//":
        if isinstance(a, int) and isinstance(b, int):
            if b == 0:
                raise LuaError("attempt to perform 'n//0'")
            return a // b
        return np.floor_divide(a, b) if isinstance(a, np.ndarray) or isinstance(b, np.ndarray) \
            else float(math.floor(a / b)) if b != 0 else (math.inf if a > 0 else -math.inf if a < 0 else math.nan)
    if op == "%":
        if isinstance(a, int) and isinstance(b, int):
            if b == 0:
                raise LuaError("attempt to perform 'n%%0'")
            return a % b
        if isinstance(a, np.ndarray) or isinstance(b, np.ndarray):
            return np.mod(a, b)
        return a - math.floor(a / b) * b if b != 0 else math.nan
    raise LuaError(f"unknown operator {op}")


def _raw_equal(a, b):
    if isinstance(a, bool) or isinstance(b, bool) or a is None or b is None:
        return a is b
    na = isinstance(a, (int, float, np.integer, np.floating))
    nb = isinstance(b, (int, float, np.integer, np.floating))
    if na and nb:
        return a == b
    if isinstance(a, str) and isinstance(b, str):
        return a == b
    return a is b  # tables and functions compare by identity


def _compare(op, a, b):
    if op in ("==", "~="):
        if isinstance(a, np.ndarray) or isinstance(b, np.ndarray):
            raise LuaError("a condition on an array of points (evaluate point by point)")
        same = _raw_equal(a, b)
        return same if op == "==" else not same
    if isinstance(a, np.ndarray) or isinstance(b, np.ndarray):
        raise LuaError("a condition on an array of points (evaluate point by point)")
    if isinstance(a, str) and isinstance(b, str):
        pass
    elif lua_type(a) != "number" or lua_type(b) != "number":
        raise LuaError(f"attempt to compare {lua_type(a)} with {lua_type(b)}")
    if op == "<":
        return a < b
    if op == "<=":
        return a <= b
    if op == ">":
        return a > b
    return a >= b


class Interpreter:
    def __init__(self, out=None):
        self.out = out if out is not None else print
        self.globals = Scope()
        self.globals.vars.update(self._stdlib())

    # ---- standard library
    def _stdlib(self):
        def m1(npf, pyf):
            def g(x):
                x = _num(x)
                return npf(x) if isinstance(x, np.ndarray) else pyf(x)
            return g

        def lmax(*a):
            if not a:
                raise LuaError("bad argument #1 to 'max' (number expected, got no value)")
            r = _num(a[0])
            for x in a[1:]:
                x = _num(x)
                r = np.maximum(r, x) if isinstance(r, np.ndarray) or isinstance(x, np.ndarray) else max(r, x)
            return r

        def lmin(*a):
            if not a:
                raise LuaError("bad argument #1 to 'min' (number expected, got no value)")
            r = _num(a[0])
            for x in a[1:]:
                x = _num(x)
                r = np.minimum(r, x) if isinstance(r, np.ndarray) or isinstance(x, np.ndarray) else min(r, x)
            return r

        def latan(y, x=1.0):
            y, x = _num(y), _num(x)
            if isinstance(y, np.ndarray) or isinstance(x, np.ndarray):
                return np.arctan2(y, x)
            return math.atan2(y, x)

        def llog(x, base=None):
            x = _num(x)
            arr = isinstance(x, np.ndarray)
            r = np.log(x) if arr else (math.log(x) if x > 0 else (-math.inf if x == 0 else math.nan))
            if base is not None:
                r = r / (np.log(base) if arr else math.log(base))
            return r

        def lfloor(x):
            x = _num(x)
            if isinstance(x, np.ndarray):
                return np.floor(x)
            return math.floor(x)

        def lceil(x):
            x = _num(x)
            if isinstance(x, np.ndarray):
                return np.ceil(x)
            return math.ceil(x)

        def lsqrt(x):
            x = _num(x)
            if isinstance(x, np.ndarray):
                return np.sqrt(x)
            return math.sqrt(x) if x >= 0 else math.nan

        def lfmod(a, b):
            a, b = _num(a), _num(b)
            if isinstance(a, np.ndarray) or isinstance(b, np.ndarray):
                return np.fmod(a, b)
            return math.fmod(a, b)

        def lmodf(x):
            x = _num(x)
            f, i = math.modf(x)
            return (float(i), f)

        def tointeger(x):
            return int(x) if isinstance(x, (int, float)) and float(x).is_integer() else None

        math_t = LuaTable({
            "pi": math.pi, "huge": math.inf, "maxinteger": 2 ** 63 - 1, "mininteger": -2 ** 63,
            "sin": m1(np.sin, math.sin), "cos": m1(np.cos, math.cos), "tan": m1(np.tan, math.tan),
            "asin": m1(np.arcsin, math.asin), "acos": m1(np.arccos, math.acos), "atan": latan,
            "exp": m1(np.exp, math.exp), "log": llog, "sqrt": lsqrt, "abs": m1(np.abs, abs),
            "floor": lfloor, "ceil": lceil, "max": lmax, "min": lmin, "fmod": lfmod, "modf": lmodf,
            "tointeger": tointeger, "sinh": m1(np.sinh, math.sinh), "cosh": m1(np.cosh, math.cosh),
            "tanh": m1(np.tanh, math.tanh),
            "type": lambda x: ("integer" if isinstance(x, int) and not isinstance(x, bool) else
                               "float" if isinstance(x, (float, np.ndarray)) else None),
        })

        def lprint(*a):
            self.out("\t".join(tostr(x) for x in a))

        def ltonumber(x, base=None):
            if base is not None:
                try:
                    return int(str(x).strip(), int(base))
                except ValueError:
                    return None
            if isinstance(x, (int, float)) and not isinstance(x, bool):
                return x
            try:
                s = str(x).strip()
                if re.fullmatch(r"-?\d+", s):
                    return int(s)
                return float(s)
            except ValueError:
                return None

        def lipairs(t):
            def it(tt, i):
                v = tt.get(i + 1)
                return (None,) if v is None else (i + 1, v)
            return (it, t, 0)

        def lnext(t, k=None):
            keys = list(t.d.keys())
            if k is None:
                idx = 0
            else:
                try:
                    idx = keys.index(_key(k)) + 1
                except ValueError:
                    raise LuaError("invalid key to 'next'")
            if idx >= len(keys):
                return (None,)
            return (keys[idx], t.d[keys[idx]])

        def lpairs(t):
            if not isinstance(t, LuaTable):
                raise LuaError(f"bad argument #1 to 'pairs' (table expected, got {lua_type(t)})")
            return (lnext, t, None)

        def lselect(n, *a):
            if n == "#":
                return len(a)
            n = int(n)
            return tuple(a[n - 1:]) if n > 0 else tuple(a[n:])

        def lerror(msg=None, level=1):
            raise LuaError(tostr(msg))

        def lassert(v, msg="assertion failed!", *rest):
            if not truthy(v):
                raise LuaError(tostr(msg))
            return (v, msg) + rest

        def lformat(fmt, *a):
            out, ai = [], 0
            for m in re.finditer(r"%([-+ #0]*\d*(?:\.\d+)?)([diouxXeEfgGqsc%])|[^%]+", fmt):
                if m.group(2) is None:
                    out.append(m.group(0))
                    continue
                spec, conv = m.group(1), m.group(2)
                if conv == "%":
                    out.append("%")
                    continue
                v = a[ai]
                ai += 1
                if conv in "di":
                    out.append(("%" + spec + "d") % int(v))
                elif conv in "ouxX":
                    out.append(("%" + spec + conv) % int(v))
                elif conv in "eEfgG":
                    out.append(("%" + spec + conv) % float(v))
                elif conv == "c":
                    out.append(chr(int(v)))
                elif conv == "q":
                    out.append('"' + tostr(v).replace("\\", "\\\\").replace('"', '\\"') + '"')
                else:
                    out.append(("%" + spec + "s") % tostr(v))
            return "".join(out)

        def lsub(s, i, j=-1):
            n = len(s)
            i = int(i)
            j = int(j)
            i = n + i + 1 if i < 0 else (1 if i == 0 else i)
            j = n + j + 1 if j < 0 else min(j, n)
            return s[i - 1:j] if i <= j else ""

        string_t = LuaTable({"format": lformat, "rep": lambda s, n, sep="": sep.join([s] * int(n)),
                             "sub": lsub, "len": lambda s: len(s), "upper": lambda s: s.upper(),
                             "lower": lambda s: s.lower()})

        def tinsert(t, *a):
            if len(a) == 1:
                t.set(t.length() + 1, a[0])
            else:
                pos, v = int(a[0]), a[1]
                n = t.length()
                for k in range(n, pos - 1, -1):
                    t.set(k + 1, t.get(k))
                t.set(pos, v)

        def tremove(t, pos=None):
            n = t.length()
            pos = n if pos is None else int(pos)
            v = t.get(pos)
            for k in range(pos, n):
                t.set(k, t.get(k + 1))
            t.set(n, None)
            return v

        def tunpack(t, i=1, j=None):
            j = t.length() if j is None else int(j)
            return tuple(t.get(k) for k in range(int(i), j + 1))

        table_t = LuaTable({"insert": tinsert, "remove": tremove, "unpack": tunpack,
                            "concat": lambda t, sep="", i=1, j=None: sep.join(
                                tostr(v) for v in tunpack(t, i, j))})
        return {"math": math_t, "string": string_t, "table": table_t, "print": lprint, "type": lua_type,
                "tostring": tostr, "tonumber": ltonumber, "ipairs": lipairs, "pairs": lpairs, "next": lnext,
                "select": lselect, "error": lerror, "assert": lassert, "unpack": tunpack}

    # ---- execution
    def run(self, source, name="chunk"):
        ast = Parser(tokenize(source)).chunk()
        f = LuaFunction(self, [], True, ast, self.globals, name)
        return self.call(f, [])

    def call(self, f, args):
        if isinstance(f, LuaFunction):
            scope = Scope(f.env)
            for i, p in enumerate(f.params):
                scope.vars[p] = args[i] if i < len(args) else None
            varargs = args[len(f.params):] if f.vararg else []
            try:
                self.exec_block(f.body, scope, varargs)
            except _Return as r:
                return r.values
            return []
        if callable(f):
            r = f(*args)
            if isinstance(r, tuple):
                return list(r)
            return [] if r is None and getattr(f, "_noret", False) else [r]
        raise LuaError(f"attempt to call a {lua_type(f)} value")

    def exec_block(self, stmts, scope, va):
        for st in stmts:
            self.exec_stmt(st, scope, va)

    def exec_stmt(self, st, scope, va):
        k = st[0]
        if k == "local":
            vals = self.eval_list(st[2], scope, va)
            for i, n in enumerate(st[1]):
                scope.vars[n] = vals[i] if i < len(vals) else None
        elif k == "assign":
            vals = self.eval_list(st[2], scope, va)
            # evaluate every target's table and key before assigning
            places = []
            for t in st[1]:
                if t[0] == "name":
                    places.append(("name", t[1]))
                else:
                    obj = self.eval(t[1], scope, va)
                    places.append(("index", obj, self.eval(t[2], scope, va)))
            for i, pl in enumerate(places):
                v = vals[i] if i < len(vals) else None
                if pl[0] == "name":
                    s = scope.lookup(pl[1]) or self.globals
                    s.vars[pl[1]] = v
                else:
                    if not isinstance(pl[1], LuaTable):
                        raise LuaError(f"attempt to index a {lua_type(pl[1])} value")
                    pl[1].set(pl[2], v)
        elif k == "exprstat":
            self.eval_multi(st[1], scope, va)
        elif k == "localfunc":
            scope.vars[st[1]] = None
            scope.vars[st[1]] = self.eval(st[2], scope, va)
        elif k == "return":
            raise _Return(self.eval_list(st[1], scope, va))
        elif k == "if":
            for cond, body in st[1]:
                if truthy(self.eval(cond, scope, va)):
                    self.exec_block(body, Scope(scope), va)
                    return
            if st[2] is not None:
                self.exec_block(st[2], Scope(scope), va)
        elif k == "while":
            try:
                while truthy(self.eval(st[1], scope, va)):
                    self.exec_block(st[2], Scope(scope), va)
            except _Break:
                pass
        elif k == "repeat":
            try:
                while True:
                    inner = Scope(scope)
                    self.exec_block(st[1], inner, va)
                    if truthy(self.eval(st[2], inner, va)):
                        break
            except _Break:
                pass
        elif k == "do":
            self.exec_block(st[1], Scope(scope), va)
        elif k == "break":
            raise _Break()
        elif k == "fornum":
            a = _num(self.eval(st[2], scope, va), "'for' initial value")
            b = _num(self.eval(st[3], scope, va), "'for' limit")
            c = _num(self.eval(st[4], scope, va), "'for' step") if st[4] is not None else 1
            if c == 0:
                raise LuaError("'for' step is zero")
            i = a
            try:
                while (c > 0 and i <= b) or (c < 0 and i >= b):
                    inner = Scope(scope)
                    inner.vars[st[1]] = i
                    self.exec_block(st[5], inner, va)
                    i = i + c
            except _Break:
                pass
        elif k == "forin":
            vals = self.eval_list(st[2], scope, va)
            f, s, ctl = (vals + [None, None, None])[:3]
            try:
                while True:
                    r = self.call(f, [s, ctl])
                    if not r or r[0] is None:
                        break
                    ctl = r[0]
                    inner = Scope(scope)
                    for i, n in enumerate(st[1]):
                        inner.vars[n] = r[i] if i < len(r) else None
                    self.exec_block(st[3], inner, va)
            except _Break:
                pass
        else:
            raise LuaError(f"unknown statement {k}")

    def eval_list(self, exprs, scope, va):
        out = []
        for i, e in enumerate(exprs):
            if i == len(exprs) - 1 and e[0] in ("call", "method", "vararg"):
                out.extend(self.eval_multi(e, scope, va))
            else:
                out.append(self.eval(e, scope, va))
        return out

    def eval_multi(self, e, scope, va):
        k = e[0]
        if k == "vararg":
            return list(va)
        if k == "call":
            f = self.eval(e[1], scope, va)
            return self.call(f, self.eval_list(e[2], scope, va))
        if k == "method":
            obj = self.eval(e[1], scope, va)
            if not isinstance(obj, LuaTable):
                raise LuaError(f"attempt to index a {lua_type(obj)} value (method {e[2]})")
            return self.call(obj.get(e[2]), [obj] + self.eval_list(e[3], scope, va))
        return [self.eval(e, scope, va)]

    def eval(self, e, scope, va):
        k = e[0]
        if k == "const":
            return e[1]
        if k == "name":
            s = scope.lookup(e[1])
            return s.vars[e[1]] if s is not None else None
        if k == "index":
            obj = self.eval(e[1], scope, va)
            key = self.eval(e[2], scope, va)
            if isinstance(obj, LuaTable):
                return obj.get(key)
            raise LuaError(f"attempt to index a {lua_type(obj)} value (key {tostr(key)!r})")
        if k in ("call", "method", "vararg"):
            r = self.eval_multi(e, scope, va)
            return r[0] if r else None
        if k == "paren":
            return self.eval(e[1], scope, va)
        if k == "binop":
            op = e[1]
            if op == "and":
                a = self.eval(e[2], scope, va)
                return self.eval(e[3], scope, va) if truthy(a) else a
            if op == "or":
                a = self.eval(e[2], scope, va)
                return a if truthy(a) else self.eval(e[3], scope, va)
            a = self.eval(e[2], scope, va)
            b = self.eval(e[3], scope, va)
            if op in ("+", "-", "*", "/", "//", "%", "^"):
                return _arith(op, a, b)
            if op == "..":
                for x in (a, b):
                    if lua_type(x) not in ("string", "number"):
                        raise LuaError(f"attempt to concatenate a {lua_type(x)} value")
                return tostr(a) + tostr(b)
            if op in ("==", "~=", "<", "<=", ">", ">="):
                return _compare(op, a, b)
            raise LuaError(f"operator {op} is not supported")
        if k == "unop":
            op = e[1]
            v = self.eval(e[2], scope, va)
            if op == "-":
                return -_num(v)
            if op == "not":
                return not truthy(v)
            if op == "#":
                if isinstance(v, str):
                    return len(v)
                if isinstance(v, LuaTable):
                    return v.length()
                raise LuaError(f"attempt to get length of a {lua_type(v)} value")
            raise LuaError(f"operator {op} is not supported")
        if k == "table":
            t = LuaTable()
            pos = 1
            items = e[1]
            for idx, it in enumerate(items):
                if it[0] == "pos":
                    if idx == len(items) - 1 and it[1][0] in ("call", "method", "vararg"):
                        for v in self.eval_multi(it[1], scope, va):
                            t.set(pos, v)
                            pos += 1
                    else:
                        t.set(pos, self.eval(it[1], scope, va))
                        pos += 1
                else:
                    t.set(self.eval(it[1], scope, va), self.eval(it[2], scope, va))
            return t
        if k == "function":
            return LuaFunction(self, e[1], e[2], e[3], scope, e[4])
        raise LuaError(f"unknown expression {k}")


def run(source, out=None):
    """Execute a chunk; returns its list of return values."""
    return Interpreter(out).run(source)


def run_file(path, out=None):
    with open(path) as fh:
        return run(fh.read(), out)


def vectorize(f, nout, *coords):
    """Evaluate the Lua function f(x[, y[, z]]) at npts points given as
    coordinate arrays; returns [npts][nout].  One call on whole arrays when
    the function is straight-line arithmetic, else one call per point.  A
    function returning a table returns its entries 1..nout."""
    coords = [np.asarray(c, dtype=np.float64) for c in coords]
    npts = coords[0].size

    def unpack(r):
        if isinstance(r, LuaTable):
            return [r.get(i + 1) for i in range(nout)]
        if isinstance(r, tuple):
            return list(r[:nout])
        return [r]

    try:
        r = unpack(f(*coords))
        cols = [np.broadcast_to(np.asarray(v, dtype=np.float64), (npts,)) for v in r]
        if len(cols) == nout:
            return np.stack(cols, axis=1).copy()
    except (LuaError, TypeError, ValueError):
        pass
    out = np.empty((npts, nout))
    for i in range(npts):
        r = unpack(f(*[float(c[i]) for c in coords]))
        if len(r) != nout or any(v is None for v in r):
            raise LuaError(f"expected {nout} value(s) from the function, got {len(r)}")
        out[i] = r
    return out
