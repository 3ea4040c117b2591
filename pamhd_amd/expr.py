"""Math expressions of the reference's JSON configurations, evaluated over
all cells at once.

The reference hands every string-valued option to muparserx (a third-party
library, cloned from github.com/beltoforion/muparserx at whatever revision
INSTALL.md's `git clone` gives: no pinned version) through
source/math/expression.hpp:58-80 (variables t, x, y, z, radius, lat, lon)
and source/grid/options.hpp:100-290 (grid options, with the variables
`cells` and `volume`), creating its parser as
ParserX(pckCOMMON | pckNON_COMPLEX | pckMATRIX | pckUNIT).  This module
implements the part of that language the configurations use:

* numbers (`1`, `0.5`, `6.371e6`), the constants `pi`, `e`, `true`, `false`;
* `+ - * / ^` (power binds tighter than a sign: `-x^2` is `-(x^2)`; `^` is
  right associative), `%` as C fmod, comparisons `< > <= >= == !=`,
  `and`/`&&`, `or`/`||`, `!`, and `c ? a : b` (lowest precedence, right
  associative);
* vectors `{a, b, c}` and indexing `v[i]`;
* functions sin cos tan asin acos atan sinh cosh tanh asinh acosh atanh exp
  ln log10 log2 sqrt cbrt abs fmod, and min / max of two or more arguments.

All numbers are doubles and `/` divides in floating point.  Transcendental
functions (cbrt aside) and `^` are evaluated element by element with the C
library's (Python's math module calls it), as the reference's std::
functions do; + - * / and sqrt are IEEE-exact in numpy.  Both branches of `?:` are
evaluated for every cell and selected per cell.
"""
import math
import re

import numpy as np

__all__ = ["Expression", "ExpressionError", "evaluate"]


class ExpressionError(ValueError):
    pass


_TOKEN = re.compile(r"""
    (?P<num>(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?)
  | (?P<name>[A-Za-z_][A-Za-z_0-9]*)
  | (?P<op>==|!=|<=|>=|&&|\|\||[-+*/%^<>?:,(){}\[\]!])
  | (?P<space>\s+)
""", re.VERBOSE)

# binding powers: ternary 1, or 2, and 3, equality 5, relational 6,
# additive 9, multiplicative 10, sign 10 (operand parsed above it, so
# that `^` binds inside), power 11 (right associative), index 13
_BINARY = {
    "or": 2, "||": 2, "and": 3, "&&": 3,
    "==": 5, "!=": 5, "<": 6, ">": 6, "<=": 6, ">=": 6,
    "+": 9, "-": 9, "*": 10, "/": 10, "%": 10, "^": 11,
}
_SIGN_BP = 10
_CONSTANTS = {"pi": math.pi, "e": math.e, "true": True, "false": False}


def _libm1(f, fallback):
    """f of the C library element by element; where Python's math module
    reports a domain error or an overflow instead of returning nan or
    +-inf, the value comes from the numpy ufunc (nan / +-inf as C gives)"""
    def one(v):
        try:
            return f(v)
        except (ValueError, OverflowError):
            return float(fallback(v))
    u = np.frompyfunc(one, 1, 1)

    def g(a):
        a = np.asarray(a, dtype=np.float64)
        if a.ndim == 0:
            return np.float64(one(float(a)))
        return u(a).astype(np.float64)
    return g


def _pow_one(x, y):
    try:
        return math.pow(x, y)
    except (ValueError, OverflowError):
        return float(np.power(np.float64(x), np.float64(y)))


_pow_u = np.frompyfunc(_pow_one, 2, 1)


def _pow(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if a.ndim == 0 and b.ndim == 0:
        return np.float64(_pow_one(float(a), float(b)))
    a, b = np.broadcast_arrays(a, b)
    return _pow_u(a, b).astype(np.float64)


_FUNCS1 = {name: _libm1(getattr(math, name), getattr(np, np_name)) for name, np_name in (
    ("sin", "sin"), ("cos", "cos"), ("tan", "tan"), ("asin", "arcsin"), ("acos", "arccos"),
    ("atan", "arctan"), ("sinh", "sinh"), ("cosh", "cosh"), ("tanh", "tanh"),
    ("asinh", "arcsinh"), ("acosh", "arccosh"), ("atanh", "arctanh"), ("exp", "exp"),
    ("log10", "log10"), ("log2", "log2"))}
_FUNCS1["ln"] = _libm1(math.log, np.log)
_FUNCS1["cbrt"] = lambda a: np.cbrt(np.asarray(a, dtype=np.float64))   # numpy's (no math.cbrt here)
_FUNCS1["sqrt"] = lambda a: np.sqrt(np.asarray(a, dtype=np.float64))
_FUNCS1["abs"] = lambda a: np.abs(np.asarray(a, dtype=np.float64))


class Expression:
    """A parsed expression; evaluate(**variables) returns a float or bool
    array broadcast over the variables, or a tuple of them for a vector."""

    def __init__(self, text):
        if not isinstance(text, str):
            raise ExpressionError("expression must be a string: %r" % (text,))
        self.text = text
        self._toks = self._tokenize(text)
        self._pos = 0
        self.tree = self._parse(0)
        if self._peek() is not None:
            raise ExpressionError("unexpected %r in %r" % (self._peek()[1], text))
        del self._toks

    # --- parsing -------------------------------------------------------
    @staticmethod
    def _tokenize(text):
        toks, pos = [], 0
        while pos < len(text):
            m = _TOKEN.match(text, pos)
            if not m:
                raise ExpressionError("invalid character %r in %r" % (text[pos], text))
            pos = m.end()
            kind = m.lastgroup
            if kind == "space":
                continue
            val = m.group(kind)
            if kind == "name" and val in ("and", "or"):
                kind = "op"
            toks.append((kind, val))
        return toks

    def _peek(self):
        return self._toks[self._pos] if self._pos < len(self._toks) else None

    def _next(self):
        t = self._peek()
        if t is None:
            raise ExpressionError("unexpected end of %r" % self.text)
        self._pos += 1
        return t

    def _expect(self, val):
        t = self._next()
        if t[1] != val:
            raise ExpressionError("expected %r, got %r in %r" % (val, t[1], self.text))

    def _parse(self, min_bp):
        left = self._prefix()
        while True:
            t = self._peek()
            if t is None or t[0] != "op":
                break
            op = t[1]
            if op == "[":
                self._next()
                idx = self._parse(0)
                self._expect("]")
                left = ("idx", left, idx)
                continue
            if op == "?":
                if min_bp > 1:
                    break
                self._next()
                a = self._parse(0)
                self._expect(":")
                b = self._parse(1)
                left = ("tern", left, a, b)
                continue
            bp = _BINARY.get(op)
            if bp is None or bp <= min_bp:
                break
            self._next()
            right = self._parse(bp - 1 if op == "^" else bp)
            left = ("bin", op, left, right)
        return left

    def _prefix(self):
        kind, val = self._next()
        if kind == "num":
            return ("num", float(val))
        if kind == "name":
            t = self._peek()
            if t is not None and t[1] == "(":
                self._next()
                args = []
                if self._peek() is not None and self._peek()[1] == ")":
                    self._next()
                else:
                    while True:
                        args.append(self._parse(0))
                        sep = self._next()[1]
                        if sep == ")":
                            break
                        if sep != ",":
                            raise ExpressionError("expected ',' or ')' in %r" % self.text)
                return ("call", val, args)
            if val in _CONSTANTS:
                return ("const", _CONSTANTS[val])
            return ("var", val)
        if val == "(":
            e = self._parse(0)
            self._expect(")")
            return e
        if val == "{":
            items = [self._parse(0)]
            while True:
                t = self._next()
                if t[1] == "}":
                    break
                if t[1] != ",":
                    raise ExpressionError("expected ',' or '}' in %r" % self.text)
                items.append(self._parse(0))
            return ("vec", items)
        if val in ("-", "+", "!"):
            return ({"-": "neg", "+": "pos", "!": "not"}[val], self._parse(_SIGN_BP))
        raise ExpressionError("unexpected %r in %r" % (val, self.text))

    # --- evaluation ----------------------------------------------------
    def evaluate(self, **variables):
        with np.errstate(all="ignore"):
            return self._eval(self.tree, variables)

    def variables(self):
        out = set()

        def walk(n):
            if n[0] == "var":
                out.add(n[1])
            for c in n[1:]:
                if isinstance(c, tuple):
                    walk(c)
                elif isinstance(c, list):
                    for d in c:
                        walk(d)
        walk(self.tree)
        return out

    def _eval(self, n, env):
        k = n[0]
        if k == "num" or k == "const":
            return n[1]
        if k == "var":
            if n[1] not in env:
                raise ExpressionError("unknown variable %r in %r" % (n[1], self.text))
            v = env[n[1]]
            return tuple(v) if isinstance(v, (list, tuple)) else v
        if k == "vec":
            return tuple(self._eval(c, env) for c in n[1])
        if k == "idx":
            v = self._eval(n[1], env)
            i = self._eval(n[2], env)
            if not isinstance(v, tuple):
                raise ExpressionError("indexing a scalar in %r" % self.text)
            i = float(np.asarray(i))
            if i != int(i) or not 0 <= i < len(v):
                raise ExpressionError("index %r out of range in %r" % (i, self.text))
            return v[int(i)]
        if k == "neg":
            return _map1(lambda a: -_num(a), self._eval(n[1], env))
        if k == "pos":
            return self._eval(n[1], env)
        if k == "not":
            return np.logical_not(self._eval(n[1], env))
        if k == "tern":
            c = self._eval(n[1], env)
            a = self._eval(n[2], env)
            b = self._eval(n[3], env)
            if isinstance(a, tuple) or isinstance(b, tuple):
                if not (isinstance(a, tuple) and isinstance(b, tuple) and len(a) == len(b)):
                    raise ExpressionError("?: branches of different shapes in %r" % self.text)
                return tuple(np.where(c, x, y) for x, y in zip(a, b))
            return np.where(c, a, b)
        if k == "bin":
            return self._binary(n[1], self._eval(n[2], env), self._eval(n[3], env))
        if k == "call":
            args = [self._eval(a, env) for a in n[2]]
            return self._call(n[1], args)
        raise ExpressionError("internal: node %r" % (k,))

    def _binary(self, op, a, b):
        if isinstance(a, tuple) or isinstance(b, tuple):
            # matrix package: element-wise sums and differences of vectors,
            # vectors scaled by (or divided by) scalars
            if op in ("+", "-") and isinstance(a, tuple) and isinstance(b, tuple) and len(a) == len(b):
                return tuple(self._binary(op, x, y) for x, y in zip(a, b))
            if op == "*" and isinstance(a, tuple) != isinstance(b, tuple):
                v, s = (a, b) if isinstance(a, tuple) else (b, a)
                return tuple(self._binary("*", x, s) if v is a else self._binary("*", s, x) for x in v)
            if op == "/" and isinstance(a, tuple) and not isinstance(b, tuple):
                return tuple(self._binary("/", x, b) for x in a)
            raise ExpressionError("operator %r on vectors in %r" % (op, self.text))
        if op in ("or", "||"):
            return np.logical_or(a, b)
        if op in ("and", "&&"):
            return np.logical_and(a, b)
        a, b = _num(a), _num(b)
        if op == "+":
            return a + b
        if op == "-":
            return a - b
        if op == "*":
            return a * b
        if op == "/":
            return np.true_divide(a, b)
        if op == "%":
            return np.fmod(a, b)
        if op == "^":
            return _pow(a, b)
        if op == "<":
            return a < b
        if op == ">":
            return a > b
        if op == "<=":
            return a <= b
        if op == ">=":
            return a >= b
        if op == "==":
            return a == b
        if op == "!=":
            return a != b
        raise ExpressionError("operator %r" % op)

    def _call(self, name, args):
        if any(isinstance(a, tuple) for a in args):
            raise ExpressionError("function %s of a vector in %r" % (name, self.text))
        if name in _FUNCS1:
            if len(args) != 1:
                raise ExpressionError("%s takes one argument in %r" % (name, self.text))
            return _FUNCS1[name](_num(args[0]))
        if name in ("min", "max"):
            if len(args) < 1:
                raise ExpressionError("%s needs arguments in %r" % (name, self.text))
            r = _num(args[0])
            for a in args[1:]:
                r = _minmax(name, r, _num(a))
            return r
        if name == "fmod":
            if len(args) != 2:
                raise ExpressionError("fmod takes two arguments in %r" % self.text)
            return np.fmod(_num(args[0]), _num(args[1]))
        raise ExpressionError("unknown function %r in %r" % (name, self.text))


def _minmax(name, a, b):
    # std::min(a, b) = (b < a) ? b : a, std::max(a, b) = (a < b) ? b : a
    if name == "min":
        return np.where(b < a, b, a)
    return np.where(a < b, b, a)


def _num(a):
    if isinstance(a, (bool, np.bool_)) or (isinstance(a, np.ndarray) and a.dtype == bool):
        return np.asarray(a, dtype=np.float64)
    return a


def _map1(f, v):
    return tuple(f(x) for x in v) if isinstance(v, tuple) else f(v)


def evaluate(text, **variables):
    return Expression(text).evaluate(**variables)
