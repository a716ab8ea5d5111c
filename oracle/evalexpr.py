"""Oracle restatement of the expression evaluation of the multigrid front end
— TEST INFRASTRUCTURE ONLY (tests/ use it as the checker of the library's
compiled expression programs; the product never imports it).

The reference evaluates its configuration's expressions with the `evalexpr`
crate, version 7.2.0 (multigrid/Cargo.toml:20), through
multigrid/src/eval/evalexpr.rs: `default_context` (:292-350: math
constants, GAMMA, and the configuration's own `if` and `case`), the globals
evaluated once into it (:22-24, :45-47), and per point a ContextWrapper
(:109-166) whose inputs t, x, y, rho, rhou, rhov, u, v shadow every other
name. The crate is absent here; this is a plain tree-walking interpreter of
its published semantics, written from its documentation:
  - values: Int (i64), Float (f64), Boolean, Empty, Tuple;
  - precedence: `^` 120, unary `-` `!` 110, `* / %` 100, `+ -` 95,
    comparisons 80, `&&` 75, `||` 70, `=` and `op=` 50 (right-associative),
    `,` 40 (tuples), `;` 0 (chaining: the value of the last statement, Empty
    after a trailing `;`); function application binds tightest;
  - Int op Int stays Int (division truncates, % keeps the dividend's sign),
    any Float operand makes the result Float, `^` is always Float;
    `==` compares type and value (Int 1 != Float 1.0), `<` etc. compare
    numbers;
  - math::* functions return Float; min, max keep Int when all arguments
    are Int.
It evaluates one point at a time (pure Python: small grids only).
"""
import math

import numpy as np

__all__ = ["Evaluator", "ExprError", "evaluate", "default_globals"]


class ExprError(ValueError):
    pass


INT, FLOAT, BOOL, EMPTY, TUPLE = "int", "float", "bool", "empty", "tuple"


class V:
    __slots__ = ("kind", "v")

    def __init__(self, kind, v):
        self.kind, self.v = kind, v

    def __repr__(self):
        return f"{self.kind}({self.v!r})"


def _num(a):
    if a.kind not in (INT, FLOAT):
        raise ExprError(f"expected a number, got {a.kind}")
    return float(a.v)


# ---------------------------------------------------------------- parsing
_OPS = ["&&=", "||=", "==", "!=", "<=", ">=", "&&", "||", "+=", "-=", "*=", "/=", "%=", "^=",
        "+", "-", "*", "/", "%", "^", "(", ")", "<", ">", "=", "!", ",", ";"]
_BIN = {"||": 70, "&&": 75, "<": 80, ">": 80, "<=": 80, ">=": 80, "==": 80, "!=": 80, "+": 95, "-": 95,
        "*": 100, "/": 100, "%": 100, "^": 120}
_ASSIGN = {"=", "+=", "-=", "*=", "/=", "%=", "^=", "&&=", "||="}


def _tokens(src):
    out, i = [], 0
    while i < len(src):
        c = src[i]
        if c.isspace():
            i += 1
        elif c.isdigit() or (c == "." and i + 1 < len(src) and src[i + 1].isdigit()):
            j = i
            while j < len(src) and (src[j].isdigit() or src[j] == "."):
                j += 1
            if j < len(src) and src[j] in "eE":
                k = j + 1
                if k < len(src) and src[k] in "+-":
                    k += 1
                if k < len(src) and src[k].isdigit():
                    while k < len(src) and src[k].isdigit():
                        k += 1
                    j = k
            text = src[i:j]
            if any(ch in text for ch in ".eE"):
                out.append(("num", V(FLOAT, float(text))))
            else:
                out.append(("num", V(INT, int(text))))
            i = j
        elif c.isalpha() or c == "_":
            j = i
            while j < len(src) and (src[j].isalnum() or src[j] in "_:"):
                j += 1
            out.append(("id", src[i:j]))
            i = j
        else:
            for o in _OPS:
                if src.startswith(o, i):
                    out.append(("op", o))
                    i += len(o)
                    break
            else:
                raise ExprError(f"unexpected character {c!r}")
    out.append(("end", None))
    return out


class _Parser:
    """Nodes: ("num", V), ("var", name), ("bin", op, a, b), ("neg", a),
    ("not", a), ("assign", name, op, rhs), ("call", name, [args]),
    ("tuple", [items]), ("chain", [stmts]), ("empty",)."""

    def __init__(self, src):
        self.t = _tokens(src)
        self.i = 0

    def peek(self):
        return self.t[self.i]

    def is_op(self, o):
        k, v = self.peek()
        return k == "op" and v == o

    def parse(self):
        n = self.chain()
        if self.peek()[0] != "end":
            raise ExprError(f"unexpected {self.peek()[1]!r}")
        return n

    def starts(self):
        k, v = self.peek()
        return k in ("num", "id") or (k == "op" and v in ("(", "-", "!"))

    def chain(self):
        first = self.tuple() if self.starts() else ("empty",)
        if not self.is_op(";"):
            return first
        stmts = [first]
        while self.is_op(";"):
            self.i += 1
            stmts.append(self.tuple() if self.starts() else ("empty",))
        return ("chain", stmts)

    def tuple(self):
        first = self.assign()
        if not self.is_op(","):
            return first
        items = [first]
        while self.is_op(","):
            self.i += 1
            items.append(self.assign())
        return ("tuple", items)

    def assign(self):
        lhs = self.binary(70)
        k, v = self.peek()
        if k == "op" and v in _ASSIGN:
            if lhs[0] != "var":
                raise ExprError("assignment to a non-variable")
            self.i += 1
            return ("assign", lhs[1], v, self.assign())
        return lhs

    def binary(self, min_prec):
        lhs = self.unary()
        while True:
            k, v = self.peek()
            if k != "op" or v not in _BIN or _BIN[v] < min_prec:
                return lhs
            self.i += 1
            lhs = ("bin", v, lhs, self.binary(_BIN[v] + 1))

    def unary(self):
        if self.is_op("-") or self.is_op("!"):
            op = self.peek()[1]
            self.i += 1
            return ("neg" if op == "-" else "not", self.binary(110))
        return self.primary()

    def primary(self):
        k, v = self.peek()
        if k == "num":
            self.i += 1
            return ("num", v)
        if k == "id":
            self.i += 1
            if v in ("true", "false"):
                return ("num", V(BOOL, v == "true"))
            if self.is_op("("):
                self.i += 1
                args = [] if self.is_op(")") else [self.chain()]
                if not self.is_op(")"):
                    raise ExprError("expected ')'")
                self.i += 1
                if args and args[0][0] == "tuple":
                    args = args[0][1]
                return ("call", v, args)
            return ("var", v)
        if self.is_op("("):
            self.i += 1
            n = self.chain()
            if not self.is_op(")"):
                raise ExprError("expected ')'")
            self.i += 1
            return n
        raise ExprError(f"unexpected {v!r}")


# ---------------------------------------------------------------- evaluation
def _arith(op, a, b):
    if op == "^":
        return V(FLOAT, _F2["math::pow"](_num(a), _num(b)))
    if a.kind == INT and b.kind == INT:
        x, y = a.v, b.v
        if op == "+":
            return V(INT, x + y)
        if op == "-":
            return V(INT, x - y)
        if op == "*":
            return V(INT, x * y)
        if y == 0:
            raise ExprError("integer division by zero")
        q = abs(x) // abs(y)
        q = q if (x >= 0) == (y >= 0) else -q  # truncation toward zero
        return V(INT, q if op == "/" else x - q * y)
    x, y = _num(a), _num(b)
    if op == "+":
        return V(FLOAT, x + y)
    if op == "-":
        return V(FLOAT, x - y)
    if op == "*":
        return V(FLOAT, x * y)
    with np.errstate(all="ignore"):
        if op == "/":
            return V(FLOAT, float(np.float64(x) / np.float64(y)))
        return V(FLOAT, float(np.fmod(np.float64(x), np.float64(y))))


def _compare(op, a, b):
    if op in ("==", "!="):
        eq = a.kind == b.kind and a.v == b.v
        return V(BOOL, eq if op == "==" else not eq)
    x, y = _num(a), _num(b)
    return V(BOOL, {"<": x < y, ">": x > y, "<=": x <= y, ">=": x >= y}[op])


def _ieee(f):
    """f64 semantics (no exceptions: ln(0) = -inf, sqrt(-1) = NaN), as Rust's."""
    def g(*a):
        with np.errstate(all="ignore"):
            return float(f(*[np.float64(x) for x in a]))
    return g


_F1 = {k: _ieee(f) for k, f in {
    "math::ln": np.log, "math::log2": np.log2, "math::log10": np.log10, "math::exp": np.exp, "math::exp2": np.exp2,
    "math::sqrt": np.sqrt, "math::cbrt": np.cbrt, "math::sin": np.sin, "math::cos": np.cos, "math::tan": np.tan,
    "math::asin": np.arcsin, "math::acos": np.arccos, "math::atan": np.arctan, "math::sinh": np.sinh,
    "math::cosh": np.cosh, "math::tanh": np.tanh, "math::asinh": np.arcsinh, "math::acosh": np.arccosh,
    "math::atanh": np.arctanh}.items()}
_F2 = {k: _ieee(f) for k, f in {
    "math::pow": np.power, "math::hypot": np.hypot, "math::atan2": np.arctan2,
    "math::log": lambda u, b: np.log(u) / np.log(b)}.items()}


def _call(name, args):
    if name == "if":
        if len(args) != 3 or args[0].kind != BOOL:
            raise ExprError("if(condition, a, b)")
        return args[1] if args[0].v else args[2]
    if name == "case":
        if len(args) % 2 == 0:
            raise ExprError("case needs an odd number of arguments")
        for k in range(0, len(args) - 1, 2):
            if args[k].kind != BOOL:
                raise ExprError("case conditions are booleans")
            if args[k].v:
                return args[k + 1]
        return args[-1]
    if name in ("min", "max"):
        if not args:
            raise ExprError(name)
        vals = [_num(a) for a in args]
        r = min(vals) if name == "min" else max(vals)
        return V(INT, int(r)) if all(a.kind == INT for a in args) else V(FLOAT, r)
    if name in ("floor", "round", "ceil", "math::abs"):
        if len(args) != 1:
            raise ExprError(name)
        a = args[0]
        if a.kind == INT:
            return V(INT, abs(a.v)) if name == "math::abs" else a
        x = _num(a)
        if name == "floor":
            return V(FLOAT, float(math.floor(x)))
        if name == "ceil":
            return V(FLOAT, float(math.ceil(x)))
        if name == "math::abs":
            return V(FLOAT, abs(x))
        return V(FLOAT, math.copysign(math.floor(abs(x) + 0.5), x))  # half away from zero
    if name in _F1:
        if len(args) != 1:
            raise ExprError(name)
        return V(FLOAT, _F1[name](_num(args[0])))
    if name in _F2:
        if len(args) != 2:
            raise ExprError(name)
        return V(FLOAT, _F2[name](_num(args[0]), _num(args[1])))
    raise ExprError(f"unknown function {name!r}")


_INPUTS = ("t", "x", "y", "rho", "rhou", "rhov", "u", "v")


class _Scope:
    """ContextWrapper: the point's inputs (None where the form does not
    provide one), the expression's own variables, the globals."""

    def __init__(self, globals_, inputs):
        self.g, self.inputs, self.own = globals_, inputs, {}

    def get(self, name):
        if name in _INPUTS:
            v = self.inputs.get(name)
            if v is None:
                raise ExprError(f"{name!r} is not available")
            return V(FLOAT, v)
        if name in self.own:
            return self.own[name]
        if name in self.g:
            return self.g[name]
        raise ExprError(f"unknown variable {name!r}")


def _eval(n, sc):
    k = n[0]
    if k == "num":
        return n[1]
    if k == "empty":
        return V(EMPTY, None)
    if k == "var":
        return sc.get(n[1])
    if k == "neg":
        a = _eval(n[1], sc)
        if a.kind == INT:
            return V(INT, -a.v)
        return V(FLOAT, -_num(a))
    if k == "not":
        a = _eval(n[1], sc)
        if a.kind != BOOL:
            raise ExprError("! needs a boolean")
        return V(BOOL, not a.v)
    if k == "bin":
        op = n[1]
        a, b = _eval(n[2], sc), _eval(n[3], sc)
        if op in ("&&", "||"):
            if a.kind != BOOL or b.kind != BOOL:
                raise ExprError(f"{op} needs booleans")
            return V(BOOL, (a.v and b.v) if op == "&&" else (a.v or b.v))
        if op in ("==", "!=", "<", ">", "<=", ">="):
            return _compare(op, a, b)
        return _arith(op, a, b)
    if k == "assign":
        _, name, op, rhs = n
        val = _eval(rhs, sc)
        if op != "=":
            cur = sc.get(name)
            base = op[:-1]
            if base in ("&&", "||"):
                val = _eval(("bin", base, ("num", cur), ("num", val)), sc)
            else:
                val = _arith(base, cur, val)
        sc.own[name] = val
        return V(EMPTY, None)
    if k == "call":
        args = [_eval(a, sc) for a in n[2]]
        return _call(n[1], args)
    if k == "tuple":
        return V(TUPLE, [_eval(a, sc) for a in n[1]])
    if k == "chain":
        r = V(EMPTY, None)
        for s_ in n[1]:
            r = _eval(s_, sc)
        return r
    raise ExprError(f"bad node {k}")


def default_globals(gamma):
    """default_context (evalexpr.rs:292-301): the math constants and GAMMA."""
    return {"PI": V(FLOAT, math.pi), "E": V(FLOAT, math.e), "TAU": V(FLOAT, 2 * math.pi), "GAMMA": V(FLOAT, gamma)}


def evaluate(src, globals_=None, **inputs):
    """The value of one expression (a V)."""
    return _eval(_Parser(src).parse(), _Scope(globals_ or {}, inputs))


class Evaluator:
    """EvaluatorConservation / EvaluatorPressure (evalexpr.rs:76-290) with the
    pressure form's conversion to conservative variables
    (euler/src/eval.rs: e = p/(γ-1) + rho(u²+v²)/2, rhou = rho·u, rhov = rho·v)."""

    def __init__(self, form, globals_src, exprs, gamma):
        self.form, self.gamma = form, gamma
        self.g = default_globals(gamma)
        if globals_src:
            sc = _Scope(self.g, {})
            _eval(_Parser(globals_src).parse(), sc)
            self.g = {**self.g, **sc.own}
        self.trees = [_Parser(e).parse() for e in exprs]

    def point(self, t, x, y):
        """(rho, rhou, rhov, e) at one point."""
        ins = {"t": t, "x": x, "y": y}
        vals = []
        for i, tree in enumerate(self.trees):
            r = _eval(tree, _Scope(self.g, dict(ins)))
            vals.append(_num(r))
            if i == 0:
                ins["rho"] = vals[0]
            elif self.form == "conservation" and i == 2:
                ins["rhou"], ins["rhov"] = vals[1], vals[2]
            elif self.form == "pressure" and i == 2:
                ins["u"], ins["v"] = vals[1], vals[2]
        if self.form == "conservation":
            return tuple(vals)
        rho, u, v, p = vals
        return rho, u * rho, v * rho, p / (self.gamma - 1.0) + rho * (u * u + v * v) / 2.0

    def field(self, t, X, Y):
        import numpy as np
        out = np.empty((4,) + X.shape)
        for idx in np.ndindex(X.shape):
            out[(slice(None),) + idx] = self.point(t, float(X[idx]), float(Y[idx]))
        return out
