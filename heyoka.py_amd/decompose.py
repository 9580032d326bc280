"""Taylor decomposition of an ODE system and its lowering to the op-list format.

The decomposition has the shape heyoka documents for ``taylor_decompose``
(the ``decomposition`` property, ``heyoka/expose_batch_integrators.cpp:231``;
format described in ``doc/notebooks/sympy_interop.ipynb:194-196``): the n state
variables u_0..u_{n-1}, then m elementary operations u_n..u_{n+m-1} over earlier
u's, numbers and runtime parameters, then the n right-hand sides.

Each elementary operation has a Taylor-coefficient recurrence (Jorba & Zou 2005;
``doc/tut_taylor_method.rst:22-81``). The set is the one the GPU kernels
implement: add, sub, neg, n-ary sum, mul, div, square, sqrt, pow with a
constant exponent, exp, log, the sin/cos pair, and time.

The words produced by :meth:`Decomposition.to_words` follow ``include/hyx_oplist.h``.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import expr as _E
from .expr import expression

# Opcodes and operand kinds (include/hyx_oplist.h).
OP_ADD, OP_SUB, OP_NEG, OP_MUL, OP_DIV, OP_SQUARE, OP_SQRT, OP_POW = 1, 2, 3, 4, 5, 6, 7, 8
OP_EXP, OP_LOG, OP_SIN, OP_COS, OP_TIME, OP_SUM = 9, 10, 11, 12, 13, 14
KIND_NODE, KIND_CONST, KIND_PAR = 0, 1, 2
MAGIC = 0x31585948

OP_NAMES = {
    OP_ADD: "add", OP_SUB: "sub", OP_NEG: "neg", OP_MUL: "mul", OP_DIV: "div",
    OP_SQUARE: "square", OP_SQRT: "sqrt", OP_POW: "pow", OP_EXP: "exp", OP_LOG: "log",
    OP_SIN: "sin", OP_COS: "cos", OP_TIME: "time", OP_SUM: "sum",
}

# Degree used for "a full Taylor series" (any order may be non-zero).
INF_DEG = 1 << 20


def operand(kind: int, idx: int) -> int:
    return (kind << 30) | idx


def operand_kind(w: int) -> int:
    return w >> 30


def operand_index(w: int) -> int:
    return w & 0x3FFFFFFF


@dataclass
class Node:
    op: int
    args: tuple  # packed operands
    aux: int = 0


@dataclass
class Decomposition:
    """A lowered ODE system: state variables, elementary nodes and right-hand sides."""

    state_names: list
    n_par: int
    nodes: list  # Node for u_n .. u_{N-1}
    rhs: list  # packed operands, one per state variable
    consts: list  # float values
    sys: list = field(default_factory=list)  # the original (var, rhs) pairs

    @property
    def n_state(self) -> int:
        return len(self.state_names)

    @property
    def n_nodes(self) -> int:
        return self.n_state + len(self.nodes)

    # ------------------------------------------------------------------------------------
    def to_words(self) -> np.ndarray:
        """Encode as the uint32 op-list (include/hyx_oplist.h)."""
        body = []
        for nd in self.nodes:
            body.extend([nd.op, len(nd.args), nd.aux])
            body.extend(nd.args)
        hdr = [MAGIC, self.n_state, self.n_par, self.n_nodes, len(self.consts), 0]
        rhs_off = len(hdr) + len(body)
        hdr[5] = rhs_off
        return np.array(hdr + body + list(self.rhs), dtype=np.uint32)

    def consts_array(self) -> np.ndarray:
        return np.array(self.consts if self.consts else [0.0], dtype=np.float64)

    # ------------------------------------------------------------------------------------
    def degrees(self) -> list:
        """Highest order with a possibly non-zero Taylor coefficient, per node (u index)."""
        deg = [INF_DEG] * self.n_state

        def od(w):
            return deg[operand_index(w)] if operand_kind(w) == KIND_NODE else 0

        for nd in self.nodes:
            ds = [od(a) for a in nd.args]
            if nd.op == OP_TIME:
                d = 1
            elif nd.op in (OP_ADD, OP_SUB, OP_SUM, OP_NEG):
                d = max(ds)
            elif nd.op == OP_MUL:
                d = min(INF_DEG, ds[0] + ds[1])
            elif nd.op == OP_SQUARE:
                d = min(INF_DEG, 2 * ds[0])
            elif nd.op == OP_DIV:
                d = ds[0] if ds[1] == 0 else INF_DEG
            else:  # sqrt, pow, exp, log, sin, cos
                d = 0 if ds[0] == 0 else INF_DEG
            deg.append(d)
        return deg

    def algorithmic_flops(self, order: int, high_accuracy: bool = False) -> dict:
        """Algorithmic FP64 flops of one Taylor step at ``order``.

        Per-order counts (k = order being computed; nodes are evaluated for
        k = 0..p-1, state variables for k = 1..p): var 1; add/sub/neg 1; n-ary sum
        of q terms q-1; mul/div by a scalar 1; mul 2k+1; square k+1; div 2k+1;
        sqrt k+1; pow-const 3k+1; exp 2k+2; log 2k+1; sin/cos pair 4k+3 (counted on
        the sin node); time/par/const 0. Order 0 counts 1 per node, transcendental
        evaluations are tallied separately. Step size 2n, Horner 2n*p (compensated
        summation 4n*p). Nodes whose Taylor series is constant over the step count
        only their order-0 flop.
        """
        p = order
        n = self.n_state
        deg = self.degrees()
        fl = 0
        transc = 0
        for j, nd in enumerate(self.nodes):
            u = n + j
            if nd.op in (OP_SQRT, OP_POW, OP_EXP, OP_LOG, OP_SIN, OP_COS, OP_DIV):
                transc += 1 if nd.op != OP_DIV else 0
            if nd.op == OP_TIME:
                continue
            fl += 1  # order 0
            if deg[u] == 0:
                continue
            scalar_arg = any(operand_kind(a) != KIND_NODE or deg[operand_index(a)] == 0 for a in nd.args)
            for k in range(1, p):
                if nd.op in (OP_ADD, OP_SUB, OP_NEG):
                    fl += 1
                elif nd.op == OP_SUM:
                    fl += len(nd.args) - 1
                elif nd.op == OP_MUL:
                    fl += 1 if scalar_arg else 2 * k + 1
                elif nd.op == OP_DIV:
                    fl += 1 if operand_kind(nd.args[1]) != KIND_NODE or deg[operand_index(nd.args[1])] == 0 else 2 * k + 1
                elif nd.op == OP_SQUARE:
                    fl += k + 1
                elif nd.op == OP_SQRT:
                    fl += k + 1
                elif nd.op == OP_POW:
                    fl += 3 * k + 1
                elif nd.op == OP_EXP:
                    fl += 2 * k + 2
                elif nd.op == OP_LOG:
                    fl += 2 * k + 1
                elif nd.op == OP_SIN:
                    fl += 4 * k + 3
                elif nd.op == OP_COS:
                    pass  # counted with its sin partner
        fl += n * p  # state variables, orders 1..p
        step = 2 * n
        upd = (4 if high_accuracy else 2) * n * p
        return {"recurrences": fl, "step_size": step, "update": upd, "total": fl + step + upd, "transcendentals": transc}

    # ------------------------------------------------------------------------------------
    def as_expressions(self) -> list:
        """heyoka-style decomposition: n var defs, m ops (over u_i variables), n rhs."""
        u = [expression(f"u_{i}") for i in range(self.n_nodes)]

        def ex(w):
            k, i = operand_kind(w), operand_index(w)
            if k == KIND_NODE:
                return u[i]
            if k == KIND_CONST:
                return expression(self.consts[i])
            return _E.par[i]

        out = [expression(nm) for nm in self.state_names]
        for nd in self.nodes:
            a = [ex(w) for w in nd.args]
            if nd.op == OP_ADD:
                e = expression._func("add", *a)
            elif nd.op == OP_SUB:
                e = expression._func("sub", *a)
            elif nd.op == OP_NEG:
                e = expression._func("neg", *a)
            elif nd.op == OP_MUL:
                e = expression._func("mul", *a)
            elif nd.op == OP_DIV:
                e = expression._func("div", *a)
            elif nd.op == OP_SUM:
                e = expression._func("sum", *a)
            elif nd.op == OP_POW:
                e = expression._func("pow", a[0], expression(self.consts[nd.aux]))
            elif nd.op == OP_TIME:
                e = _E.time
            else:
                e = expression._func(OP_NAMES[nd.op], *a)
            out.append(e)
        out.extend(ex(w) for w in self.rhs)
        return out

    def as_heyoka(self) -> list:
        """The decomposition as heyoka's ``taylor_dc_t``: ``(expression, hidden_deps)``
        pairs, the hidden dependency of a sin node being its cos partner and vice versa
        (the shape of the reference's ``decomposition`` property,
        ``heyoka/expose_batch_integrators.cpp:231``)."""
        deps = [[] for _ in range(self.n_nodes + self.n_state)]
        for j, nd in enumerate(self.nodes):
            if nd.op in (OP_SIN, OP_COS):
                deps[self.n_state + j] = [int(nd.aux)]
        return [(e, d) for e, d in zip(self.as_expressions(), deps)]

    @classmethod
    def from_heyoka(cls, dc, n_state=None) -> "Decomposition":
        """Lower a heyoka-shaped Taylor decomposition to an op-list.

        ``dc`` is the list heyoka's ``taylor_decompose`` / ``decomposition`` returns
        (``heyoka/expose_batch_integrators.cpp:231``): the n state variables, then m
        elementary operations over the variables ``u_0 .. u_{n+m-1}`` (each referring
        only to earlier u's), numbers, ``par[i]`` and ``time``, then the n right-hand
        sides (a u variable, a number or a parameter). Entries may be expressions or
        ``(expression, hidden_deps)`` pairs; hidden dependencies are not needed, since a
        sin and a cos of the same argument are paired here. n is the number of leading
        entries that are variables unless ``n_state`` is given.

        Each operation becomes one node (n-ary ``prod`` a chain of binary products), so the
        op-list keeps heyoka's node structure: summation and product orders are those of
        the decomposition. A sin or a cos without its partner gets one added.
        """
        entries = [e[0] if isinstance(e, (tuple, list)) else e for e in dc]
        entries = [_E._ex(e) for e in entries]
        if n_state is None:
            n_state = 0
            while n_state < len(entries) and entries[n_state].is_var():
                n_state += 1
            n_state = min(n_state, len(entries) // 2)
        n = int(n_state)
        m = len(entries) - 2 * n
        if n <= 0 or m < 0:
            raise ValueError(f"A Taylor decomposition of {len(entries)} entries cannot hold {n} state variables "
                             "(n variables, m operations and n right-hand sides are expected)")
        names = []
        for i in range(n):
            if not entries[i].is_var():
                raise ValueError(f"Entry {i} of the decomposition must be a state variable, not '{entries[i]}'")
            names.append(entries[i].name)
        lw = _Lowerer(names)
        w = {i: operand(KIND_NODE, i) for i in range(n)}

        def arg(a, i):
            if a.is_var():
                nm = a.name
                if not (nm.startswith("u_") and nm[2:].isdigit()):
                    raise ValueError(f"Entry {i} of the decomposition refers to '{nm}', which is not a u variable")
                j = int(nm[2:])
                if j >= i or j not in w:
                    raise ValueError(f"Entry {i} of the decomposition refers to u_{j}, which is not defined before it")
                return w[j]
            if a.is_number():
                return lw.const(a.value)
            if a.is_par():
                lw.n_par = max(lw.n_par, a.idx + 1)
                return operand(KIND_PAR, a.idx)
            if a.is_time():
                return lw.node(OP_TIME, ())
            raise ValueError(f"Entry {i} of the decomposition has a nested operand '{a}': every operand of an "
                             "elementary operation must be a u variable, a number, a parameter or time")

        binary = {"add": OP_ADD, "sub": OP_SUB, "div": OP_DIV}
        unary = {"neg": OP_NEG, "square": OP_SQUARE, "sqrt": OP_SQRT, "exp": OP_EXP, "log": OP_LOG}
        for i in range(n, n + m):
            e = entries[i]
            if e.is_time():
                w[i] = lw.node(OP_TIME, ())
                continue
            if not e.is_func():
                raise ValueError(f"Entry {i} of the decomposition must be an elementary operation, not '{e}'")
            name = e.name
            # A power's exponent is a number folded into the node, not an operand.
            a = [arg(x, i) for x in (e.args[:1] if name == "pow" else e.args)]
            if name == "sum" or (name == "add" and len(a) != 2):
                if len(a) < 2:
                    raise ValueError(f"Entry {i} of the decomposition is a sum of {len(a)} term(s)")
                w[i] = lw.node(OP_ADD, a) if len(a) == 2 else lw.node(OP_SUM, a)
            elif name in ("mul", "prod"):
                if len(a) < 2:
                    raise ValueError(f"Entry {i} of the decomposition is a product of {len(a)} factor(s)")
                acc = lw.node(OP_MUL, a[:2])
                for x in a[2:]:
                    acc = lw.node(OP_MUL, [acc, x])
                w[i] = acc
            elif name in binary and len(a) == 2:
                w[i] = lw.node(binary[name], a)
            elif name in unary and len(a) == 1:
                w[i] = lw.node(unary[name], a)
            elif name in ("sin", "cos") and len(a) == 1:
                s, c = lw.sincos(a[0])
                w[i] = s if name == "sin" else c
            elif name == "pow" and len(e.args) == 2:
                ex = e.args[1]
                if not ex.is_number():
                    raise NotImplementedError(f"Entry {i} of the decomposition is a power with a non-constant "
                                              "exponent, which heyoka decomposes further")
                v = ex.value
                if v == 2.0:
                    w[i] = lw.node(OP_SQUARE, a[:1])
                elif v == 0.5:
                    w[i] = lw.node(OP_SQRT, a[:1])
                else:
                    w[i] = lw.node(OP_POW, a[:1], operand_index(lw.const(v)))
            else:
                raise NotImplementedError(f"Unsupported elementary operation '{name}' with {len(a)} argument(s) "
                                          f"in entry {i} of the decomposition")
        rhs = [arg(entries[n + m + j], n + m + j) if not entries[n + m + j].is_time() else lw.node(OP_TIME, ())
               for j in range(n)]
        nodes, rhs, consts = _canonical_consts(lw.nodes, rhs, lw.consts)
        return cls(state_names=names, n_par=lw.n_par, nodes=nodes, rhs=rhs, consts=consts, sys=[])


def _canonical_consts(nodes, rhs, consts):
    """Renumber the constants in order of first use (node operands and pow exponents in
    node order, then the right-hand sides), so that equal node lists give equal op-list
    words whatever order a lowering met the constants in."""
    order, seen = [], {}

    def use(i):
        if i not in seen:
            seen[i] = len(order)
            order.append(i)

    for nd in nodes:
        for a in nd.args:
            if operand_kind(a) == KIND_CONST:
                use(operand_index(a))
        if nd.op == OP_POW:
            use(nd.aux)
    for a in rhs:
        if operand_kind(a) == KIND_CONST:
            use(operand_index(a))
    for i in range(len(consts)):
        use(i)

    def mp(a):
        return operand(KIND_CONST, seen[operand_index(a)]) if operand_kind(a) == KIND_CONST else a

    new_nodes = [Node(nd.op, tuple(mp(a) for a in nd.args), seen[nd.aux] if nd.op == OP_POW else nd.aux)
                 for nd in nodes]
    return new_nodes, [mp(a) for a in rhs], [consts[i] for i in order]


class _Lowerer:
    def __init__(self, state_names):
        self.n = len(state_names)
        self.var_idx = {nm: i for i, nm in enumerate(state_names)}
        self.nodes = []
        self.memo = {}
        self.consts = []
        self.const_idx = {}
        self.n_par = 0
        self.ex_memo = {}

    def const(self, v: float) -> int:
        key = float(v).hex()
        if key not in self.const_idx:
            self.const_idx[key] = len(self.consts)
            self.consts.append(float(v))
        return operand(KIND_CONST, self.const_idx[key])

    def node(self, op, args, aux=0) -> int:
        key = (op, tuple(args), aux)
        if key in self.memo:
            return self.memo[key]
        idx = self.n + len(self.nodes)
        self.nodes.append(Node(op, tuple(args), aux))
        w = operand(KIND_NODE, idx)
        self.memo[key] = w
        return w

    def sincos(self, a):
        ks, kc = ("sin", a), ("cos", a)
        if ks not in self.memo:
            s_idx = self.n + len(self.nodes)
            c_idx = s_idx + 1
            self.nodes.append(Node(OP_SIN, (a,), c_idx))
            self.nodes.append(Node(OP_COS, (a,), s_idx))
            self.memo[ks] = operand(KIND_NODE, s_idx)
            self.memo[kc] = operand(KIND_NODE, c_idx)
        return self.memo[ks], self.memo[kc]

    def flatten_sum(self, e, out):
        if e.is_func() and e.name in ("add", "sum"):
            for a in e.args:
                self.flatten_sum(a, out)
        else:
            out.append(e)

    def _scalar(self, e: expression) -> bool:
        """True when e is a number, a parameter or a function of those only (a degree-0 series)."""
        key = ("scalar", e)
        if key not in self.ex_memo:
            if e.is_var() or e.is_time():
                r = False
            elif e.is_func():
                r = all(self._scalar(a) for a in e.args)
            else:
                r = True
            self.ex_memo[key] = r
        return self.ex_memo[key]

    def lower(self, e: expression) -> int:
        key = e
        if key in self.ex_memo:
            return self.ex_memo[key]
        w = self._lower(e)
        self.ex_memo[key] = w
        return w

    def _lower(self, e: expression) -> int:
        if e.is_var():
            if e.name not in self.var_idx:
                raise ValueError(
                    f"The variable '{e.name}' appears in the right-hand side of the ODE system "
                    "but it is not a state variable"
                )
            return operand(KIND_NODE, self.var_idx[e.name])
        if e.is_number():
            return self.const(e.value)
        if e.is_par():
            self.n_par = max(self.n_par, e.idx + 1)
            return operand(KIND_PAR, e.idx)
        if e.is_time():
            return self.node(OP_TIME, ())
        name, args = e.name, e.args
        if name in ("add", "sum"):
            terms = []
            self.flatten_sum(e, terms)
            ws = [self.lower(t) for t in terms]
            if len(ws) == 2:
                return self.node(OP_ADD, ws)
            return self.node(OP_SUM, ws)
        if name == "prod":
            # heyoka's n-ary product: binary products from the left.
            acc = args[0]
            for a in args[1:]:
                acc = expression._func("mul", acc, a)
            return self.lower(acc)
        if name == "sub":
            return self.node(OP_SUB, [self.lower(args[0]), self.lower(args[1])])
        if name == "neg":
            return self.node(OP_NEG, [self.lower(args[0])])
        if name == "mul":
            a, b = args
            # Scalar factors outermost: (s * c) * b -> s * (c * b) when s holds no variable
            # and no time. The series product then runs on c and b, a multiplication by s
            # costs one flop per order, and no coefficient history of (s * c) is kept.
            if not self._scalar(a) and not self._scalar(b):
                for x, y in ((a, b), (b, a)):
                    if x.is_func() and x.name == "mul":
                        p, q = x.args
                        if self._scalar(p) and not self._scalar(q):
                            return self.lower(expression._func("mul", p, expression._func("mul", q, y)))
                        if self._scalar(q) and not self._scalar(p):
                            return self.lower(expression._func("mul", q, expression._func("mul", p, y)))
            return self.node(OP_MUL, [self.lower(a), self.lower(b)])
        if name == "div":
            return self.node(OP_DIV, [self.lower(args[0]), self.lower(args[1])])
        if name == "square":
            return self.node(OP_SQUARE, [self.lower(args[0])])
        if name == "sqrt":
            return self.node(OP_SQRT, [self.lower(args[0])])
        if name == "exp":
            return self.node(OP_EXP, [self.lower(args[0])])
        if name == "log":
            return self.node(OP_LOG, [self.lower(args[0])])
        if name in ("sin", "cos"):
            s, c = self.sincos(self.lower(args[0]))
            return s if name == "sin" else c
        if name == "pow":
            base, ex = args
            if not ex.is_number():
                # a**b with a non-constant exponent: exp(b * log(a)).
                return self.lower(_E.exp(ex * _E.log(base)))
            a = self.lower(base)
            v = ex.value
            if v == 2.0:
                return self.node(OP_SQUARE, [a])
            if v == 0.5:
                return self.node(OP_SQRT, [a])
            if v == 1.0:
                return a
            if v == 0.0:
                return self.const(1.0)
            return self.node(OP_POW, [a], operand_index(self.const(v)))
        raise NotImplementedError(f"Unsupported function '{name}' in the decomposition")


def taylor_decompose(sys) -> Decomposition:
    """Decompose ``sys`` (a list of ``(variable, rhs)`` pairs) into elementary operations."""
    pairs = []
    for item in sys:
        if not isinstance(item, (tuple, list)) or len(item) != 2:
            raise TypeError("An ODE system must be a list of (variable, expression) pairs")
        lhs, rhs = _E._ex(item[0]), _E._ex(item[1])
        if not lhs.is_var():
            raise ValueError(f"The left-hand side of an ODE must be a variable, not '{lhs}'")
        pairs.append((lhs, rhs))
    names = [l.name for l, _ in pairs]
    if len(set(names)) != len(names):
        raise ValueError("The state variables of an ODE system must be unique")
    if not names:
        raise ValueError("Cannot decompose an empty ODE system")
    lw = _Lowerer(names)
    rhs = [lw.lower(r) for _, r in pairs]
    nodes, rhs, consts = _canonical_consts(lw.nodes, rhs, lw.consts)
    return Decomposition(state_names=names, n_par=lw.n_par, nodes=nodes, rhs=rhs, consts=consts, sys=pairs)


def order_from_tol(tol: float) -> int:
    """Taylor order p = ceil(-ln(tol)/2 + 1).

    Reproduces the orders heyoka prints: 20 at tol = 2.22e-16
    (``doc/notebooks/The adaptive integrator.ipynb:54``) and 12 at tol = 1e-9
    (``doc/notebooks/Customising the adaptive integrator.ipynb:51``).
    """
    if not (tol > 0.0) or not math.isfinite(tol):
        raise ValueError(f"The tolerance in an adaptive Taylor integrator must be finite and positive, but it is {tol} instead")
    p = math.ceil(-math.log(tol) / 2.0 + 1.0)
    if p < 2:
        raise ValueError(f"The Taylor order deduced from the tolerance {tol} is {p}, but it must be at least 2")
    return int(p)
