"""Variational ODE systems: ``var_ode_sys(sys, args, order)``.

Mirrors ``heyoka.var_ode_sys`` (``heyoka/expose_var_ode_sys.cpp:29-59``): the
system is augmented with the equations of the partial derivatives of the state
with respect to the chosen arguments (initial conditions and/or runtime
parameters), up to ``order``.

State layout (``doc/notebooks/var_ode_sys.ipynb`` cells 7 and 31): the original
variables, then for each order 1..order, for each original component i, the
derivatives of x_i for every multi-index of that order in descending
lexicographic order of the exponent vector. The variable of the derivative with
sparse multi-index [(j, e_j), ...] of x_i is named ``∂[(j, e_j), ...]x_i``.

The equations follow from differentiating x_i' = f_i(x, p) along the
arguments: with D_a the total derivative with respect to argument a
(D_a x_k = ∂x_k/∂a, D_a of a derivative variable = the next-order variable,
plus ∂/∂p for a parameter argument), the right-hand side of a derivative
variable of multi-index m is D applied to f_i once per unit of m.
"""

from __future__ import annotations

import enum
import itertools

import numpy as np

from .expr import expression, diff, par as _par, get_variables, _ex, evaluate, time as _time


class var_args(enum.IntFlag):
    vars = 1
    params = 2
    time = 4
    all = 7


def _mname(m, vname):
    parts = [f"({j}, {e})" for j, e in enumerate(m) if e]
    return "∂[" + ", ".join(parts) + "]" + vname


def _mindices(nargs, order):
    """Exponent vectors of total degree `order`, descending lexicographic."""
    out = [m for m in itertools.product(range(order, -1, -1), repeat=nargs) if sum(m) == order]
    return out


class var_ode_sys:
    def __init__(self, sys, args=var_args.vars, order=1):
        if not isinstance(order, int) or order < 1:
            raise ValueError("The 'order' argument to the var_ode_sys constructor must be nonzero")
        pairs = [(_ex(l), _ex(r)) for l, r in sys]
        svars = [l for l, _ in pairs]
        if isinstance(args, var_args):
            vargs = []
            if args & var_args.vars:
                vargs += svars
            if args & var_args.params:
                npar = 0
                stack = [r for _, r in pairs]
                while stack:
                    e = stack.pop()
                    if e.is_par():
                        npar = max(npar, e.idx + 1)
                    elif e.is_func():
                        stack.extend(e.args)
                vargs += [_par[i] for i in range(npar)]
            if args & var_args.time:
                vargs.append(_time)
        else:
            vargs = [_ex(a) for a in args]
            for a in vargs:
                if not (a.is_var() or a.is_par() or a.is_time()):
                    raise TypeError("The arguments of a var_ode_sys must be state variables, parameters or time")
                if a.is_var() and all(a != v for v in svars):
                    raise ValueError(f"The variable '{a}' is not a state variable of the ODE system")
        if not vargs:
            raise ValueError("Cannot formulate the variational equations without any argument")
        self._orig = pairs
        self._vargs = vargs
        self._order = order
        self._n_orig = len(pairs)
        self._build()

    def _build(self):
        n, A = self._n_orig, len(self._vargs)
        names = [l.name for l, _ in self._orig]
        rhs = {(i, (0,) * A): r for i, (_, r) in enumerate(self._orig)}
        var = {(i, (0,) * A): l for i, (l, _) in enumerate(self._orig)}
        self._mindex = [[i] + [0] * A for i in range(n)]
        self._slices = {}
        out = list(self._orig)
        for o in range(1, self._order + 1):
            ms = _mindices(A, o)
            start = len(out)
            for i in range(n):
                for m in ms:
                    var[(i, m)] = expression(_mname(m, names[i]))
            for i in range(n):
                for m in ms:
                    # m = m_prev + e_a for the last argument a present in m.
                    a = max(j for j in range(A) if m[j])
                    mp = tuple(m[j] - (1 if j == a else 0) for j in range(A))
                    r = self._total_derivative(rhs[(i, mp)], a, var, A)
                    rhs[(i, m)] = r
                    out.append((var[(i, m)], r))
                    self._mindex.append([i] + list(m))
            self._slices[o] = (start, len(out))
        self._sys = out
        self._var = var

    def _total_derivative(self, g, a, var, A):
        """D_a g: chain rule through the state and derivative variables, plus ∂/∂p."""
        lookup = {v.name: key for key, v in var.items()}
        terms = []
        for vn in get_variables(g):
            key = lookup.get(vn)
            if key is None:
                raise ValueError(f"Unknown variable '{vn}' in a variational equation")
            i, m = key
            m2 = tuple(m[j] + (1 if j == a else 0) for j in range(A))
            if sum(m2) > self._order:
                continue
            dg = diff(g, expression(vn))
            if dg.is_number() and dg.value == 0.0:
                continue
            terms.append(dg * var[(i, m2)])
        va = self._vargs[a]
        if va.is_par():
            dp = diff(g, va)
            if not (dp.is_number() and dp.value == 0.0):
                terms.append(dp)
        if not terms:
            return expression(0.0)
        out = terms[0]
        for t in terms[1:]:
            out = out + t
        return out

    # ---- properties mirroring heyoka.var_ode_sys -----------------------------------------
    @property
    def sys(self):
        return list(self._sys)

    @property
    def vargs(self):
        return list(self._vargs)

    @property
    def n_orig_sv(self):
        return self._n_orig

    @property
    def order(self):
        return self._order

    def get_vslice(self, order, component=None):
        if order == 0:
            return slice(0, self._n_orig) if component is None else slice(component, component + 1)
        if order not in self._slices:
            raise ValueError(f"Invalid order {order} for a variational system of order {self._order}")
        s, e = self._slices[order]
        if component is None:
            return slice(s, e)
        per = (e - s) // self._n_orig
        return slice(s + component * per, s + (component + 1) * per)

    def get_mindex(self, i):
        return list(self._mindex[i])

    def initial_derivatives(self, x0=None, pars=None, t0=0.0):
        """Initial values of the derivative variables.

        Order 1: d x_i / d x_j(t0) = delta_ij, d x_i / d p = 0, d x_i / d t0 = -f_i(x0, t0)
        (the solution shifted in its initial time); higher orders: 0, except that
        higher derivatives involving t0 are not formulated (NotImplementedError).
        With x0 given ([n_orig] or [n_orig, B]) the values are arrays over the batch."""
        n, A = self._n_orig, len(self._vargs)
        vals = []
        has_t = any(a.is_time() for a in self._vargs)
        fvals = None
        if has_t:
            if x0 is None:
                raise ValueError("The initial state is needed for derivatives with respect to the initial time")
            x0 = [np.asarray(x0)[i] for i in range(n)]
            env = {l.name: x0[i] for i, (l, _) in enumerate(self._orig)}
            fvals = [evaluate(r, env, pars, t0) for _, r in self._orig]
        for idx in range(n, len(self._sys)):
            mi = self._mindex[idx]
            i, m = mi[0], mi[1:]
            v = 0.0
            if sum(m) == 1:
                a = m.index(1)
                va = self._vargs[a]
                if va.is_var() and va.name == self._orig[i][0].name:
                    v = 1.0
                elif va.is_time():
                    v = -fvals[i]
            elif has_t and any(m[j] for j, va in enumerate(self._vargs) if va.is_time()):
                raise NotImplementedError("Initial values of higher-order derivatives with respect to the initial "
                                          "time are not supported")
            vals.append(v)
        return vals

    def __copy__(self):
        return self

    def __deepcopy__(self, memo):
        return self
