"""Compiled functions on the GPU: ``cfunc(fn, vars)``.

Mirrors heyoka.py's ``cfunc`` (``heyoka/cfunc.cpp:90-121`` constructor, ``:121-440``
call operator; tests ``heyoka/_test_cfunc.py``) for FP64: a list of expressions of the
variables ``vars`` (plus ``par[i]`` and ``time``) is lowered to the op-list, compiled
with hipRTC for gfx950 and evaluated with one GPU lane per point.

``cf(inputs, pars=None, time=None, outputs=None)`` takes ``inputs`` of shape
``(nvars,)`` (one point; ``time`` a scalar) or ``(nvars, N)`` (N points, batch index
fastest; ``time`` an array of N), and returns the outputs ``(nouts,)`` or
``(nouts, N)``. ``cf.eval_batch(ta)`` evaluates at every instance of a
``taylor_adaptive_batch`` straight from its device-resident state, parameters and time
(on-device diagnostics such as energies or Jacobi constants); the variables must then
be the integrator's state variables, in order.
"""

from __future__ import annotations

import ctypes as ct

import numpy as np

from . import _lib
from .decompose import taylor_decompose
from .expr import _ex, expression, get_variables

_PAD = "__hyx_cfunc_pad_"


def _walk(es):
    stack = list(es)
    while stack:
        e = stack.pop()
        yield e
        if e.is_func():
            stack.extend(e.args)


class cfunc:
    def __init__(self, fn, vars, **kwargs):
        # LLVM / multiprecision options of the reference are accepted and ignored.
        fp_type = kwargs.pop("fp_type", float)
        if fp_type not in (float, np.float64):
            raise TypeError(f"The floating-point type {fp_type} is not supported by the GPU compiled functions")
        for k in ("high_accuracy", "compact_mode", "parallel_mode", "batch_size", "prec", "opt_level",
                  "force_avx512", "slp_vectorize", "fast_math", "code_model", "parjit"):
            kwargs.pop(k, None)
        if kwargs:
            raise TypeError(f"Unexpected keyword argument(s): {sorted(kwargs)}")
        fn = [_ex(f) for f in fn]
        vars = [_ex(v) for v in vars]
        if not fn:
            raise ValueError("Cannot create a compiled function with no outputs")
        names = []
        for v in vars:
            if not v.is_var():
                raise TypeError(f"The variables of a compiled function must be variables, not '{v}'")
            names.append(v.name)
        if len(set(names)) != len(names):
            raise ValueError("The list of variables of a compiled function contains duplicates")
        missing = sorted(set(get_variables(fn)) - set(names))
        if missing:
            raise ValueError(f"The compiled function's expressions contain variables not in the list of variables: "
                             f"{missing}")
        self._fn, self._vars = fn, vars
        m, q = len(vars), len(fn)
        n = max(m, q, 1)
        rows = vars + [expression(f"{_PAD}{i}") for i in range(n - m)]
        rhs = fn + [expression(0.0)] * (n - q)
        self._dec = taylor_decompose(list(zip(rows, rhs)))
        self._nparams = max([e.idx + 1 for e in _walk(fn) if e.is_par()], default=0)
        self._time_dep = any(e.is_time() for e in _walk(fn))
        self._h = None
        self._L = None

    def _handle(self, source_only=False):
        if self._h is not None and not source_only:
            return self._h
        L = _lib.lib()
        words, consts = _lib.oplist_args(self._dec)
        h = ct.c_void_p()
        flags = _lib.FLAG_SOURCE_ONLY if source_only else 0
        _lib.check(L.hyx_cfunc_create(_lib.dptr(words, ct.c_uint32), len(words), _lib.dptr(consts), len(self._dec.consts),
                                      len(self._vars), len(self._fn), flags, ct.byref(h)))
        if source_only:
            return h
        self._h, self._L = h, L
        return h

    def __del__(self):
        if getattr(self, "_h", None) is not None and self._L is not None:
            self._L.hyx_cfunc_destroy(self._h)
            self._h = None

    def source(self) -> str:
        """The generated HIP source (no GPU needed)."""
        L = _lib.lib()
        h = self._handle(source_only=True)
        try:
            n = L.hyx_cfunc_source(h, None, 0)
            buf = ct.create_string_buffer(n + 1)
            L.hyx_cfunc_source(h, buf, n + 1)
            return buf.value.decode()
        finally:
            L.hyx_cfunc_destroy(h)

    # ---- properties of the reference ---------------------------------------------------
    @property
    def fn(self):
        return list(self._fn)

    @property
    def vars(self):
        return list(self._vars)

    @property
    def nouts(self):
        return len(self._fn)

    @property
    def nparams(self):
        return self._nparams

    @property
    def is_time_dependent(self):
        return self._time_dep

    @property
    def dc(self):
        """Decomposition: the variables, the elementary operations, then the outputs."""
        ex = self._dec.as_expressions()
        n = self._dec.n_state
        m, q = len(self._vars), len(self._fn)
        return ex[:m] + ex[n:n + self._dec.n_nodes - n] + ex[self._dec.n_nodes:self._dec.n_nodes + q]

    # ---- evaluation ---------------------------------------------------------------------
    def __call__(self, inputs, pars=None, time=None, outputs=None):
        x = np.asarray(inputs, dtype=np.float64)
        if x.ndim not in (1, 2):
            raise ValueError(f"The array of inputs provided for the evaluation of a compiled function has {x.ndim} "
                             "dimensions, but it must have either 1 or 2 dimensions instead")
        multi = x.ndim == 2
        if x.shape[0] != len(self._vars):
            raise ValueError(f"The array of inputs provided for the evaluation of a compiled function has "
                             f"{x.shape[0]} rows, but the function has {len(self._vars)} variables")
        N = x.shape[1] if multi else 1
        xi = np.ascontiguousarray(x.reshape(len(self._vars), N))
        pa = None
        if pars is not None:
            pa = np.asarray(pars, dtype=np.float64)
            if pa.ndim != x.ndim:
                raise ValueError(f"The array of parameters provided for the evaluation of a compiled function has "
                                 f"{pa.ndim} dimension(s), but it must have {x.ndim} dimension(s) instead (i.e., the "
                                 "same number of dimensions as the array of inputs)")
            pa = np.ascontiguousarray(pa.reshape(pa.shape[0], N))
        if self._nparams and (pa is None or pa.shape[0] < self._nparams):
            raise ValueError(f"The compiled function needs {self._nparams} parameter rows")
        tm = None
        if time is not None:
            if multi and np.ndim(time) == 0:
                raise TypeError("The time value cannot be a scalar when evaluating a compiled function over batches "
                                "of inputs, it should be an array of values instead")
            if not multi and np.ndim(time) != 0:
                raise TypeError("The time value cannot be an array when evaluating a compiled function over a "
                                "single set of inputs, it should be a scalar instead")
            tm = np.ascontiguousarray(np.broadcast_to(np.asarray(time, dtype=np.float64), (N,)))
        out = np.empty((self.nouts, N))
        h = self._handle()
        _lib.check(self._L.hyx_cfunc_eval(h, _lib.dptr(xi), _lib.dptr(pa), 0 if pa is None else pa.shape[0],
                                          _lib.dptr(tm), N, _lib.dptr(out)))
        res = out if multi else out[:, 0].copy()
        if outputs is not None:
            o = np.asarray(outputs)
            if o.ndim != x.ndim:
                raise ValueError(f"The array of outputs provided for the evaluation of a compiled function has "
                                 f"{o.ndim} dimension(s), but it must have {x.ndim} dimension(s) instead (i.e., the "
                                 "same number of dimensions as the array of inputs)")
            outputs[...] = res
            return outputs
        return res

    def eval_batch(self, ta):
        """Evaluate at every instance of ``ta`` from its device-resident state: returns (nouts, B)."""
        if [v.name for v in self._vars] != list(ta._hyx_dec.state_names):
            raise ValueError("eval_batch() needs the integrator's state variables, in order, as the function's "
                             "variables")
        ta._push()
        out = np.empty((self.nouts, ta.batch_size))
        h = self._handle()
        _lib.check(self._L.hyx_cfunc_eval_batch(h, ta._hyx_batch.h, _lib.dptr(out)))
        return out

    # ---- copies and pickling ------------------------------------------------------------
    def __copy__(self):
        return cfunc(self._fn, self._vars)

    def __deepcopy__(self, memo):
        return cfunc(self._fn, self._vars)

    def __getstate__(self):
        return {"fn": self._fn, "vars": self._vars}

    def __setstate__(self, st):
        self.__init__(st["fn"], st["vars"])
