"""Symbolic expressions for ODE right-hand sides.

A small, self-contained expression system with the names heyoka.py exposes for
building dynamical systems (``heyoka/expose_expression.cpp:150-305``,
``heyoka/__init__.py:345-381``): ``expression``, ``make_vars``, the arithmetic
operators, ``pow`` via ``**``, ``sqrt``, ``exp``, ``log``, ``sin``, ``cos``,
``sum``, ``par[i]``, ``time``, ``diff`` and ``subs``.

Expressions are immutable trees with structural equality (``a == b`` is a
bool, as in heyoka.py) and structural hashing, which the decomposer uses for
common-subexpression elimination.
"""

from __future__ import annotations

import math
import numbers

import numpy as np

__all__ = [
    "expression",
    "make_vars",
    "par",
    "time",
    "sin",
    "cos",
    "exp",
    "log",
    "sqrt",
    "square",
    "sum",
    "prod",
    "diff",
    "subs",
    "get_variables",
]

# Kinds of expression nodes.
_VAR, _NUM, _PAR, _TIME, _FUNC = range(5)

# Function names understood by the decomposer.
_FUNCS = {"add", "sub", "mul", "div", "neg", "pow", "sqrt", "exp", "log", "sin", "cos", "square", "sum", "prod"}


def _fmt_num(v: float) -> str:
    # 17 significant digits with trailing zeros kept, e.g. 5.0000000000000000,
    # matching how heyoka.py prints numbers inside expressions.
    if not math.isfinite(v):
        return repr(v)
    return np.format_float_positional(v, precision=17, unique=False, fractional=False, trim="k")


class expression:
    """An immutable symbolic expression.

    ``expression(1.5)`` builds a number, ``expression("x")`` a variable.
    """

    __slots__ = ("_k", "_a", "_h")

    def __init__(self, value=0.0):
        if isinstance(value, expression):
            self._k, self._a, self._h = value._k, value._a, value._h
            return
        if isinstance(value, str):
            self._k, self._a = _VAR, value
        elif isinstance(value, numbers.Real) and not isinstance(value, bool):
            self._k, self._a = _NUM, float(value)
        else:
            raise TypeError(f"Cannot construct an expression from an object of type {type(value)}")
        self._h = None

    # ----- internal constructors ---------------------------------------------------------
    @staticmethod
    def _make(kind, payload):
        e = expression.__new__(expression)
        e._k, e._a, e._h = kind, payload, None
        return e

    @staticmethod
    def _func(name, *args):
        assert name in _FUNCS
        return expression._make(_FUNC, (name, tuple(args)))

    # ----- introspection -----------------------------------------------------------------
    @property
    def kind(self) -> str:
        return ("variable", "number", "param", "time", "func")[self._k]

    def is_number(self) -> bool:
        return self._k == _NUM

    def is_var(self) -> bool:
        return self._k == _VAR

    def is_par(self) -> bool:
        return self._k == _PAR

    def is_time(self) -> bool:
        return self._k == _TIME

    def is_func(self) -> bool:
        return self._k == _FUNC

    @property
    def value(self) -> float:
        if self._k != _NUM:
            raise TypeError("value is defined only for numbers")
        return self._a

    @property
    def name(self) -> str:
        if self._k == _VAR:
            return self._a
        if self._k == _FUNC:
            return self._a[0]
        raise TypeError("name is defined only for variables and functions")

    @property
    def idx(self) -> int:
        if self._k != _PAR:
            raise TypeError("idx is defined only for parameters")
        return self._a

    @property
    def args(self) -> tuple:
        return self._a[1] if self._k == _FUNC else ()

    # ----- structural equality / hashing ---------------------------------------------------
    def _key(self):
        if self._k == _FUNC:
            return (_FUNC, self._a[0], tuple(a._key() for a in self._a[1]))
        if self._k == _NUM:
            # -0.0 and 0.0 compare equal, as numbers do.
            return (_NUM, self._a + 0.0)
        return (self._k, self._a)

    def __hash__(self):
        if self._h is None:
            if self._k == _FUNC:
                self._h = hash((_FUNC, self._a[0], tuple(hash(a) for a in self._a[1])))
            else:
                self._h = hash(self._key())
        return self._h

    def __eq__(self, other):
        if not isinstance(other, expression):
            if isinstance(other, (numbers.Real, str)) and not isinstance(other, bool):
                other = expression(other)
            else:
                return NotImplemented
        if self is other:
            return True
        if hash(self) != hash(other):
            return False
        return self._key() == other._key()

    def __ne__(self, other):
        r = self.__eq__(other)
        return r if r is NotImplemented else not r

    # ----- printing ---------------------------------------------------------------------
    def __repr__(self):
        k = self._k
        if k == _VAR:
            return self._a
        if k == _NUM:
            return _fmt_num(self._a)
        if k == _PAR:
            return f"p{self._a}"
        if k == _TIME:
            return "t"
        name, args = self._a
        if name in ("add", "sub", "mul", "div"):
            op = {"add": "+", "sub": "-", "mul": "*", "div": "/"}[name]
            return f"({args[0]!r} {op} {args[1]!r})"
        if name == "sum":
            return "(" + " + ".join(repr(a) for a in args) + ")"
        if name == "neg":
            return f"-{args[0]!r}"
        if name == "pow":
            return f"{args[0]!r}**{args[1]!r}"
        return f"{name}(" + ", ".join(repr(a) for a in args) + ")"

    __str__ = __repr__

    # ----- arithmetic -------------------------------------------------------------------
    def __add__(self, other):
        return _add(self, _ex(other))

    def __radd__(self, other):
        return _add(_ex(other), self)

    def __sub__(self, other):
        return _sub(self, _ex(other))

    def __rsub__(self, other):
        return _sub(_ex(other), self)

    def __mul__(self, other):
        return _mul(self, _ex(other))

    def __rmul__(self, other):
        return _mul(_ex(other), self)

    def __truediv__(self, other):
        return _div(self, _ex(other))

    def __rtruediv__(self, other):
        return _div(_ex(other), self)

    def __neg__(self):
        return _neg(self)

    def __pos__(self):
        return self

    def __pow__(self, other):
        return _pow(self, _ex(other))

    def __rpow__(self, other):
        return _pow(_ex(other), self)

    def __copy__(self):
        return self

    def __deepcopy__(self, memo):
        return self

    def __reduce__(self):
        return (_rebuild, (self._k, self._a))


def _rebuild(kind, payload):
    return expression._make(kind, payload)


def _ex(x) -> expression:
    if isinstance(x, expression):
        return x
    if isinstance(x, numbers.Real) and not isinstance(x, bool):
        return expression(float(x))
    raise TypeError(f"Cannot convert an object of type {type(x)} to an expression")


def _num(v: float) -> expression:
    return expression._make(_NUM, float(v))


# ----- construction with light simplification -----------------------------------------------
# Folding happens only where it is exact (number op number in double arithmetic, identities
# with 0 and 1), so the numerical meaning of an expression never changes.


def _is_num(e, v=None):
    return e._k == _NUM and (v is None or e._a == v)


def _add(a, b):
    if _is_num(a) and _is_num(b):
        return _num(a._a + b._a)
    if _is_num(a, 0.0):
        return b
    if _is_num(b, 0.0):
        return a
    return expression._func("add", a, b)


def _sub(a, b):
    if _is_num(a) and _is_num(b):
        return _num(a._a - b._a)
    if _is_num(b, 0.0):
        return a
    if _is_num(a, 0.0):
        return _neg(b)
    return expression._func("sub", a, b)


def _mul(a, b):
    if _is_num(a) and _is_num(b):
        return _num(a._a * b._a)
    if _is_num(a, 0.0) or _is_num(b, 0.0):
        return _num(0.0)
    if _is_num(a, 1.0):
        return b
    if _is_num(b, 1.0):
        return a
    return expression._func("mul", a, b)


def _div(a, b):
    if _is_num(a) and _is_num(b):
        return _num(a._a / b._a)
    if _is_num(a, 0.0):
        return _num(0.0)
    if _is_num(b, 1.0):
        return a
    return expression._func("div", a, b)


def _neg(a):
    if _is_num(a):
        return _num(-a._a)
    if a._k == _FUNC and a._a[0] == "neg":
        return a._a[1][0]
    return expression._func("neg", a)


def _pow(a, b):
    if _is_num(a) and _is_num(b):
        return _num(a._a ** b._a)
    if _is_num(b, 1.0):
        return a
    if _is_num(b, 0.0):
        return _num(1.0)
    return expression._func("pow", a, b)


# ----- public factories ---------------------------------------------------------------------


def make_vars(*names):
    """Create one variable per name; a single name returns a single variable."""
    if len(names) == 0:
        raise ValueError("At least one variable name must be provided")
    vs = [expression(str(n)) for n in names]
    return vs[0] if len(vs) == 1 else vs


class _par_generator:
    """``par[i]`` builds the runtime parameter with index ``i`` (``heyoka/__init__.py:345-352``)."""

    def __getitem__(self, idx):
        if not isinstance(idx, numbers.Integral) or idx < 0:
            raise TypeError("A parameter index must be a non-negative integer")
        return expression._make(_PAR, int(idx))

    def __repr__(self):
        return "par"


par = _par_generator()

time = expression._make(_TIME, None)


def _unary(name, x):
    x = _ex(x)
    return expression._func(name, x)


def sin(x):
    x = _ex(x)
    if _is_num(x):
        return _num(math.sin(x._a))
    return expression._func("sin", x)


def cos(x):
    x = _ex(x)
    if _is_num(x):
        return _num(math.cos(x._a))
    return expression._func("cos", x)


def exp(x):
    x = _ex(x)
    if _is_num(x):
        return _num(math.exp(x._a))
    return expression._func("exp", x)


def log(x):
    x = _ex(x)
    if _is_num(x):
        return _num(math.log(x._a))
    return expression._func("log", x)


def sqrt(x):
    x = _ex(x)
    if _is_num(x):
        return _num(math.sqrt(x._a))
    return expression._func("sqrt", x)


def square(x):
    x = _ex(x)
    if _is_num(x):
        return _num(x._a * x._a)
    return expression._func("square", x)


_builtin_sum = __builtins__["sum"] if isinstance(__builtins__, dict) else __builtins__.sum


def sum(terms):
    """n-ary sum. Numbers are folded into one term; an empty sum is 0."""
    terms = [_ex(t) for t in terms]
    nums = [t._a for t in terms if _is_num(t)]
    rest = [t for t in terms if not _is_num(t)]
    if nums:
        c = 0.0
        for v in nums:
            c += v
        if c != 0.0 or not rest:
            rest.append(_num(c))
    if not rest:
        return _num(0.0)
    if len(rest) == 1:
        return rest[0]
    return expression._func("sum", *rest)


def prod(terms):
    terms = [_ex(t) for t in terms]
    if not terms:
        return _num(1.0)
    out = terms[0]
    for t in terms[1:]:
        out = _mul(out, t)
    return out


# ----- analysis / transformation ------------------------------------------------------------


def get_variables(e) -> list:
    """Sorted list of variable names appearing in ``e`` (or in a list of expressions)."""
    seen = set()
    stack = list(e) if isinstance(e, (list, tuple)) else [e]
    while stack:
        x = stack.pop()
        if x._k == _VAR:
            seen.add(x._a)
        elif x._k == _FUNC:
            stack.extend(x._a[1])
    return sorted(seen)


def subs(e, smap):
    """Replace variables (by name or by expression) according to ``smap``."""
    m = {}
    for k, v in smap.items():
        key = k._a if isinstance(k, expression) and k._k == _VAR else k
        m[key] = _ex(v)
    memo = {}

    def rec(x):
        h = id(x)
        if h in memo:
            return memo[h]
        if x._k == _VAR:
            r = m.get(x._a, x)
        elif x._k == _FUNC:
            name, args = x._a
            r = _rebuild_func(name, [rec(a) for a in args])
        else:
            r = x
        memo[h] = r
        return r

    return rec(_ex(e))


def _rebuild_func(name, args):
    if name == "add":
        return _add(*args)
    if name == "sub":
        return _sub(*args)
    if name == "mul":
        return _mul(*args)
    if name == "div":
        return _div(*args)
    if name == "neg":
        return _neg(*args)
    if name == "pow":
        return _pow(*args)
    if name == "sum":
        return sum(args)
    return {"sin": sin, "cos": cos, "exp": exp, "log": log, "sqrt": sqrt, "square": square}[name](*args)


def diff(e, x):
    """Symbolic derivative of ``e`` with respect to the variable (or parameter) ``x``."""
    e = _ex(e)
    if isinstance(x, str):
        x = expression(x)
    if x._k not in (_VAR, _PAR):
        raise TypeError("Derivatives can be taken only with respect to variables or parameters")
    memo = {}
    zero = _num(0.0)

    def d(y):
        key = id(y)
        if key in memo:
            return memo[key][1]
        if y._k in (_VAR, _PAR):
            r = _num(1.0) if (y._k == x._k and y._a == x._a) else zero
        elif y._k == _FUNC:
            name, args = y._a
            if name == "add":
                r = _add(d(args[0]), d(args[1]))
            elif name == "sub":
                r = _sub(d(args[0]), d(args[1]))
            elif name == "sum":
                r = sum([d(a) for a in args])
            elif name == "neg":
                r = _neg(d(args[0]))
            elif name == "mul":
                a, b = args
                r = _add(_mul(d(a), b), _mul(a, d(b)))
            elif name == "div":
                a, b = args
                da, db = d(a), d(b)
                r = _sub(_div(da, b), _div(_mul(a, db), _mul(b, b)))
            elif name == "square":
                a = args[0]
                r = _mul(_mul(_num(2.0), a), d(a))
            elif name == "pow":
                a, c = args
                if c._k != _NUM:
                    raise NotImplementedError("Only constant exponents can be differentiated")
                r = _mul(_mul(c, _pow(a, _num(c._a - 1.0))), d(a))
            elif name == "sqrt":
                a = args[0]
                r = _div(d(a), _mul(_num(2.0), y))
            elif name == "exp":
                r = _mul(y, d(args[0]))
            elif name == "log":
                r = _div(d(args[0]), args[0])
            elif name == "sin":
                r = _mul(cos(args[0]), d(args[0]))
            elif name == "cos":
                r = _neg(_mul(sin(args[0]), d(args[0])))
            else:  # pragma: no cover
                raise NotImplementedError(name)
        else:
            r = zero
        memo[key] = (y, r)
        return r

    return d(e)


def evaluate(e, variables: dict, pars=None, time_value=0.0):
    """Evaluate an expression with numpy: ``variables`` maps names to arrays,
    ``pars`` is indexable by parameter index."""
    import numpy as _np

    memo = {}
    fns = {"sin": _np.sin, "cos": _np.cos, "exp": _np.exp, "log": _np.log, "sqrt": _np.sqrt,
           "square": lambda a: a * a, "neg": lambda a: -a}

    def ev(x):
        k = id(x)
        if k in memo:
            return memo[k][1]
        if x._k == _VAR:
            r = variables[x._a]
        elif x._k == _NUM:
            r = x._a
        elif x._k == _PAR:
            r = pars[x._a]
        elif x._k == _TIME:
            r = time_value
        else:
            name, args = x._a
            a = [ev(y) for y in args]
            if name == "add":
                r = a[0] + a[1]
            elif name == "sub":
                r = a[0] - a[1]
            elif name == "mul":
                r = a[0] * a[1]
            elif name == "div":
                r = a[0] / a[1]
            elif name == "pow":
                r = a[0] ** a[1]
            elif name == "sum":
                r = a[0]
                for y in a[1:]:
                    r = r + y
            else:
                r = fns[name](*a)
        memo[k] = (x, r)
        return r

    return ev(_ex(e))
