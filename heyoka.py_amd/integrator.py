"""``taylor_adaptive_batch``: the drop-in batch integrator backed by libhyx.so.

Mirrors the Python surface of ``heyoka.core.taylor_adaptive_batch_dbl``
(``heyoka/expose_batch_integrators.cpp:214-673``): constructor arguments and
their validation texts, ``state`` / ``pars`` writable arrays, read-only
``time`` / ``dtime`` / ``last_h`` / ``tc``, ``step`` / ``step_backward`` /
``step_res``, ``propagate_until`` / ``propagate_for`` / ``propagate_res``,
``update_d_output``, variational accessors, copy, deepcopy and pickling.

The batch lives on the GPU. The host arrays are mirrors: they are uploaded
before every device call and refreshed after it, so writing ``ta.state[:]``
and then propagating behaves as in the reference.

Lanes are independent instances: a lane that finishes or fails stops on its
own without affecting the others' results.
"""

from __future__ import annotations

import copy as _copy
import enum
import threading

import numpy as np

from . import _lib
from .decompose import taylor_decompose, order_from_tol
from .c_output import continuous_output_batch
from .events import nt_event_batch, t_event_batch, _event_ref
from .expr import _ex, expression, get_variables
from .var_ode_sys import var_ode_sys as _var_ode_sys


def _var(name):
    return expression(name)


class taylor_outcome(enum.IntEnum):
    """Outcome codes of heyoka (``heyoka/core.cpp:324-329``)."""

    success = -4294967297
    step_limit = -4294967298
    time_limit = -4294967299
    err_nf_state = -4294967300
    cb_stop = -4294967301


def _outcome(v):
    try:
        return taylor_outcome(int(v))
    except ValueError:
        return int(v)


class _PropRes:
    """A batch's propagate_res kept as its result arrays until it is read: building the
    reference's list of (outcome, min_h, max_h, nsteps) tuples costs about a microsecond
    per lane, which ensembles of many members would otherwise pay for every member."""

    __slots__ = ("oc", "mn", "mx", "ns")

    def __init__(self, oc, mn, mx, ns):
        self.oc, self.mn, self.mx, self.ns = (np.array(oc, dtype=np.int64), np.array(mn, dtype=np.float64),
                                              np.array(mx, dtype=np.float64), np.array(ns, dtype=np.uint64))

    @classmethod
    def of(cls, r, sl=slice(None)):
        return cls(r["outcome"][sl], r["min_h"][sl], r["max_h"][sl], r["nsteps"][sl])

    def copy(self):
        return _PropRes(self.oc, self.mn, self.mx, self.ns)

    def tolist(self):
        ocs = {v: _outcome(v) for v in set(self.oc.tolist())}
        return [(ocs[o], a, b, c) for o, a, b, c in zip(self.oc.tolist(), self.mn.tolist(), self.mx.tolist(),
                                                          self.ns.tolist())]


_MODES = {"persistent": _lib.MODE_PERSISTENT, "plain": _lib.MODE_PLAIN, "graph": _lib.MODE_GRAPH}

_PAD = "__hyx_event_pad_"


def _squares(e, memo=None):
    """e with every product of a subexpression with itself written as its square (the
    square recurrence sums half the terms of the product's)."""
    memo = {} if memo is None else memo
    key = id(e)
    if key in memo:
        return memo[key]
    r = e
    if e.is_func():
        args = [_squares(a, memo) for a in e.args]
        if e.name == "mul" and len(args) == 2 and args[0] == args[1]:
            r = expression._func("square", args[0])
        elif e.name == "pow" and len(args) == 2 and args[1].is_number() and args[1].value == 2.0:
            r = expression._func("square", args[0])
        elif any(a is not b for a, b in zip(args, e.args)):
            r = expression._func(e.name, *args)
    memo[key] = r
    return r


def _event_spec(dec, tes, ntes, tol):
    """The events' op-list: a row per state variable (then padding rows), the event
    functions as right-hand sides, terminal events first."""
    exprs = [_squares(e.expression) for e in tes] + [_squares(e.expression) for e in ntes]
    names = list(dec.state_names)
    missing = sorted(set(get_variables(exprs)) - set(names))
    if missing:
        raise ValueError(f"The event equations contain variables that are not state variables of the system: "
                         f"{missing}")
    n, ne = len(names), len(exprs)
    rows = [_var(v) for v in names]
    rows += [_var(f"{_PAD}{i}") for i in range(max(n, ne) - n)]
    rhs = exprs + [expression(0.0)] * (len(rows) - ne)
    edec = taylor_decompose(list(zip(rows, rhs)))
    return _lib.EventSpec(edec, len(tes), len(ntes), [int(e.direction) for e in tes + ntes],
                          [float(e.cooldown) for e in tes], tol, stopping=[e.callback is None for e in tes])

# Compiled systems keyed by (op-list, constants, order, flags): the analogue of
# heyoka's in-memory JIT cache (heyoka/core.cpp:148-149). Copies of an
# integrator share one compiled system.
_sys_cache: dict = {}
_sys_cache_lock = threading.Lock()


def _call_event(cb, *args):
    """Run an event callback; its failures surface as RuntimeError, as in the reference."""
    try:
        return cb(*args)
    except Exception as e:  # noqa: BLE001
        raise RuntimeError(f"An exception was raised by an event callback: {type(e).__name__}: {e}") from e


def get_system(dec, order, high_accuracy=False, compact=None, events=None):
    words, consts = _lib.oplist_args(dec)
    key = (words.tobytes(), consts.tobytes(), int(order), bool(high_accuracy), compact,
           None if events is None else events.key())
    with _sys_cache_lock:
        s = _sys_cache.get(key)
        if s is None:
            s = _lib.System(dec, order, high_accuracy=high_accuracy, compact=compact, events=events)
            _sys_cache[key] = s
    if compact is None and s.form == "wave" and s.tune_key is not None:
        # The wave form's variants (16, 32, 64 lanes per instance) compute bit-identical
        # results: the automatic choice takes the one the autotuner recorded as fastest.
        rates = _lib.tuned_rates({lab: _lib.tune_key(words, consts, order, lab) for lab in ("wave16", "wave32", "wave")})
        if rates:
            best = max(rates, key=rates.get)
            if best != s.label:
                return get_system(dec, order, high_accuracy, best, events)
    return s


# ---- double-length arithmetic on the host (same algorithm as the kernels) ----------------
def _two_sum(a, b):
    x = a + b
    z = x - a
    return x, (a - (x - z)) + (b - z)


def dfl_add(ah, al, bh, bl):
    s, e = _two_sum(ah, bh)
    t, f = _two_sum(al, bl)
    e = e + t
    x = s + e
    e = e - (x - s)
    s = x
    e = e + f
    x = s + e
    return x, e - (x - s)


def _dfl_norm(hi, lo):
    x = hi + lo
    return x, lo - (x - hi)


class taylor_adaptive_batch:
    """Adaptive Taylor integrator over a batch of independent instances (FP64, GPU)."""

    _internal = frozenset()

    def __init__(self, sys, state, time=None, pars=None, tol=0.0, high_accuracy=False, compact_mode=False,
                 t_events=None, nt_events=None, parallel_mode=False, **kwargs):
        gpu_mode = kwargs.pop("gpu_mode", "persistent")
        code_form = kwargs.pop("code_form", None)  # None (automatic), "compact", "unrolled", "group", "wave[32|16]"
        # LLVM-specific options of the reference are accepted and ignored.
        for k in list(kwargs):
            if k in ("opt_level", "force_avx512", "slp_vectorize", "fast_math", "code_model", "parjit", "fp_type"):
                kwargs.pop(k)
        if kwargs:
            raise TypeError(f"Unexpected keyword argument(s): {sorted(kwargs)}")
        if gpu_mode not in _MODES:
            raise ValueError(f"gpu_mode must be one of {sorted(_MODES)}")
        tes = list(t_events) if t_events is not None else []
        ntes = list(nt_events) if nt_events is not None else []
        for e in tes:
            if not isinstance(e, t_event_batch):
                raise TypeError(f"The list of terminal events can contain only t_event_batch objects, not {type(e)}")
        for e in ntes:
            if not isinstance(e, nt_event_batch):
                raise TypeError(f"The list of non-terminal events can contain only nt_event_batch objects, not "
                                f"{type(e)}")
        # The integrator keeps its own copies of the events (callbacks deep-copied).
        tes = [_copy.copy(e) for e in tes]
        ntes = [_copy.copy(e) for e in ntes]
        self._hyx_init(sys, state, time, pars, tol, high_accuracy, compact_mode, parallel_mode, gpu_mode, code_form,
                       events=(tes, ntes))

    # ------------------------------------------------------------------------------------
    def _hyx_init(self, sys, state, time, pars, tol, high_accuracy, compact_mode, parallel_mode, gpu_mode,
                  code_form, restore=None, events=([], [])):
        d = self.__dict__
        vsys = sys if isinstance(sys, _var_ode_sys) else None
        pairs = vsys.sys if vsys is not None else [(_ex(a), _ex(b)) for a, b in sys]
        st = np.asarray(state)
        if st.ndim != 2:
            raise ValueError(
                "Invalid state vector passed to the constructor of a batch integrator: the expected number of "
                f"dimensions is 2, but the input array has a dimension of {st.ndim}")
        if st.dtype != np.float64:
            if st.size and not np.can_cast(st.dtype, np.float64, casting="safe"):
                raise TypeError(f"Cannot cast the state array from {st.dtype} to float64 according to the rule 'safe'")
            st = st.astype(np.float64)
        B = int(st.shape[1])
        if B == 0:
            raise ValueError("The batch size in an adaptive Taylor integrator cannot be zero")
        dec = taylor_decompose(pairs)
        n = dec.n_state
        if st.shape[0] == 0:
            st = np.zeros((n, B))
        elif vsys is not None and st.shape[0] == vsys.n_orig_sv:
            t_init = np.zeros(B) if time is None else np.asarray(time, dtype=np.float64)
            p_init = None if pars is None else np.asarray(pars, dtype=np.float64)
            ext = vsys.initial_derivatives(st, p_init, t_init)
            ext = np.stack([np.broadcast_to(np.asarray(v, dtype=np.float64), (B,)) for v in ext]) if ext else \
                np.zeros((0, B))
            st = np.concatenate([st, ext], axis=0)
        if st.shape[0] != n:
            raise ValueError(
                f"Inconsistent sizes detected in the initialization of an adaptive Taylor integrator: the state "
                f"vector has a dimension of {st.shape[0]}, while the number of equations is {n}")
        npar = dec.n_par
        if pars is not None:
            pa = np.asarray(pars)
            if pa.ndim != 2 or pa.shape[1] != B:
                raise ValueError(
                    "Invalid parameter vector passed to the constructor of a batch integrator: the expected array "
                    f"shape is (n, {B}), but the input array has either the wrong number of dimensions or the wrong "
                    "shape")
            pa = pa.astype(np.float64)
            if pa.shape[0] < npar:
                raise ValueError(
                    f"Inconsistent sizes detected in the initialization of an adaptive Taylor integrator: the number "
                    f"of parameter rows is {pa.shape[0]}, while the number of parameters in the system is {npar}")
            npar = max(npar, pa.shape[0])
        else:
            pa = np.zeros((npar, B))
        if time is not None:
            tm = np.asarray(time, dtype=np.float64)
            if tm.ndim != 1 or tm.shape[0] != B:
                raise ValueError(
                    "Invalid time vector passed to the constructor of a batch integrator: the expected array shape is "
                    f"({B}), but the input array has either the wrong number of dimensions or the wrong shape")
        else:
            tm = np.zeros(B)
        tol = float(tol)
        if tol == 0.0:
            tol = float(np.finfo(np.float64).eps)
        order = order_from_tol(tol)
        compact = {"compact": True, "unrolled": False, "unrolled_lds": "unrolled_lds",
                   "unrolled_regs": "unrolled_regs", "group": "group", "wave": "wave", "wave32": "wave32",
                   "wave16": "wave16", None: None}[code_form]
        tes, ntes = events
        spec = None
        if tes or ntes:
            spec = _event_spec(dec, tes, ntes, tol)
            if spec.dec.n_par > dec.n_par:
                dec.n_par = spec.dec.n_par
                if pa.shape[0] < dec.n_par:
                    pa = np.concatenate([pa, np.zeros((dec.n_par - pa.shape[0], B))])
                npar = max(npar, dec.n_par)
        system = get_system(dec, order, high_accuracy=bool(high_accuracy), compact=compact, events=spec)
        d["_hyx_tes"] = tes
        d["_hyx_ntes"] = ntes
        d["_hyx_pairs"] = pairs
        d["_hyx_vsys"] = vsys
        d["_hyx_dec"] = dec
        d["_hyx_order"] = order
        d["_hyx_tol"] = tol
        d["_hyx_flags"] = (bool(high_accuracy), bool(compact_mode), bool(parallel_mode), gpu_mode, code_form)
        d["_hyx_system"] = system
        d["_hyx_B"] = B
        d["_hyx_npar"] = npar
        # The device batch is allocated on first use (_hyx_batch): copies made for an
        # ensemble are packed into one device batch and never need their own.
        d["_hyx_batch_obj"] = None
        d["_hyx_cd_pending"] = None
        d["_hyx_state"] = np.ascontiguousarray(st, dtype=np.float64).copy()
        d["_hyx_pars"] = np.ascontiguousarray(pa, dtype=np.float64).copy()
        d["_hyx_thi"] = tm.copy()
        d["_hyx_tlo"] = np.zeros(B)
        d["_hyx_last_h"] = np.zeros(B)
        d["_hyx_prop_res"] = [(taylor_outcome.success, 0.0, 0.0, 0)] * B
        d["_hyx_step_res"] = [(taylor_outcome.success, 0.0)] * B
        d["_hyx_tc"] = np.zeros((n, order + 1, B))
        d["_hyx_dout"] = np.zeros((n, B))
        if restore is not None:
            for k, v in restore.items():
                d[k] = v

    # ---- host <-> device ------------------------------------------------------------------
    @property
    def _hyx_batch(self):
        """The device batch of this integrator, created on first use (with the cooldowns a
        copy inherited)."""
        d = self.__dict__
        b = d["_hyx_batch_obj"]
        if b is None:
            b = _lib.Batch(self._hyx_system, self._hyx_B)
            if d["_hyx_cd_pending"] is not None:
                b.set_cooldowns(d["_hyx_cd_pending"])
                d["_hyx_cd_pending"] = None
            d["_hyx_batch_obj"] = b
        return b

    def _hyx_cooldowns(self):
        """[n_te, 3, B] cooldowns (trigger time hi, lo, length; length < 0: none)."""
        d = self.__dict__
        if d["_hyx_batch_obj"] is None:
            if d["_hyx_cd_pending"] is not None:
                return d["_hyx_cd_pending"].copy()
            return np.full((len(self._hyx_tes), 3, self._hyx_B), -1.0)
        return d["_hyx_batch_obj"].cooldowns()

    def _hyx_set_cooldowns(self, cd):
        d = self.__dict__
        if d["_hyx_batch_obj"] is None:
            d["_hyx_cd_pending"] = np.array(cd, dtype=np.float64)
        else:
            d["_hyx_batch_obj"].set_cooldowns(cd)

    def _push(self):
        b = self._hyx_batch
        b.write(_lib.FIELD_STATE, self._hyx_state)
        if self._hyx_system.npar:
            b.write(_lib.FIELD_PARS, self._hyx_pars[: self._hyx_system.npar])
        b.write(_lib.FIELD_TIME_HI, self._hyx_thi)
        b.write(_lib.FIELD_TIME_LO, self._hyx_tlo)

    def _pull(self):
        b = self._hyx_batch
        self._hyx_state[...] = b.read(_lib.FIELD_STATE, self._hyx_state.shape)
        self._hyx_thi[...] = b.read(_lib.FIELD_TIME_HI, (self._hyx_B,))
        self._hyx_tlo[...] = b.read(_lib.FIELD_TIME_LO, (self._hyx_B,))

    def _lanes(self, v, what):
        a = np.asarray(v, dtype=np.float64)
        if a.ndim == 0:
            return np.full(self._hyx_B, float(a))
        if a.ndim != 1 or a.shape[0] != self._hyx_B:
            raise ValueError(
                f"Invalid {what} passed to a batch integrator: the expected size is {self._hyx_B}, but a vector of "
                f"shape {a.shape} was provided instead")
        return a.copy()

    # ---- stepping -----------------------------------------------------------------------
    def _step_impl(self, mdt, write_tc):
        self._push()
        oc, h = self._hyx_batch.step(mdt, write_tc)
        self._pull()
        self._hyx_last_h[...] = h
        if self.with_events:
            te, cnt, tm, code = self._hyx_batch.events_read()
            busy = (te >= 0) | (cnt > 0)
            if write_tc or busy.any():
                # Callbacks may call update_d_output(): the coefficients of this step.
                self._hyx_tc[...] = self._hyx_batch.read_tc()
            lanes = np.nonzero(busy)[0]
            oc = oc.copy()
            for b, r in zip(lanes, self._dispatch_events(lanes, te[lanes], cnt[lanes], tm[:, lanes], code[:, lanes],
                                                         h[lanes])):
                if r is not None:
                    oc[b] = r[0] if r[1] else -(r[0] + 1)
        elif write_tc:
            self._hyx_tc[...] = self._hyx_batch.read_tc()
        self._hyx_step_res = [(_outcome(o), float(x)) for o, x in zip(oc, h)]
        return oc, h

    def _dispatch_events(self, lanes, te, cnt, tm, code, h):
        """Run the callbacks of one step's events for the given batch elements, as heyoka's
        batch step does: per element, the non-terminal ones in chronological order (the
        state and time are those at the end of the step), then the terminal one. Returns,
        per element, None or (terminal index, whether its callback lets the integration go on)."""
        cap = tm.shape[0]
        out = []
        for j, b in enumerate(lanes):
            k = int(cnt[j])
            if k > cap:
                raise RuntimeError(f"Batch element {int(b)} triggered {k} non-terminal events in one step; at most "
                                   f"{cap} are recorded per step: lower max_delta_t")
            sg = 1.0 if h[j] >= 0 else -1.0
            recs = sorted(range(k), key=lambda r: (sg * tm[r, j], int(code[r, j]) >> 2))
            for r in recs:
                c = int(code[r, j])
                _call_event(self._hyx_ntes[c >> 2].callback, self, float(tm[r, j]), (c & 3) - 1, int(b))
            if te[j] < 0:
                out.append(None)
                continue
            i, d = int(te[j]) >> 2, (int(te[j]) & 3) - 1
            cb = self._hyx_tes[i].callback
            go = False
            if cb is not None:
                r = _call_event(cb, self, d, int(b))
                if not isinstance(r, (bool, np.bool_, int, float, np.integer, np.floating)):
                    raise RuntimeError("The call operator of a terminal event callback returned an object of type "
                                       f"{type(r)} which cannot be converted to bool: the error happened in the "
                                       "construction of the return value of an event callback")
                go = bool(r)
            out.append((i, go))
        return out

    def _prop_parked(self, first, max_steps, grid_out=None):
        """Device-resident propagation with events that need Python: every instance runs in
        the propagate kernels and parks at the end of each step with triggers; only parked
        instances come to the host, where their callbacks run (each sees its own batch
        element's state, time and step coefficients), then they are resumed or stopped.
        first() starts the propagation (with parking on) and returns its records. Returns the
        number of resumes."""
        b = self._hyx_batch
        mode = _MODES[self._hyx_flags[3]]
        n = self._hyx_state.shape[0]
        b.set_event_pause(True)
        rounds = 0
        try:
            r = first()
            while True:
                lanes = b.events_parked()
                if lanes.size == 0:
                    break
                te, cnt, tm, code, h, tc = b.events_gather(lanes)
                self._hyx_state[:, lanes] = b.slots_read(_lib.FIELD_STATE, lanes, n)
                self._hyx_thi[lanes] = b.slots_read(_lib.FIELD_TIME_HI, lanes, 1)[0]
                self._hyx_tlo[lanes] = b.slots_read(_lib.FIELD_TIME_LO, lanes, 1)[0]
                self._hyx_last_h[lanes] = h
                self._hyx_tc[:, :, lanes] = tc
                res = self._dispatch_events(lanes, te, cnt, tm, code, h)
                stop = np.array([0 if (x is None or x[1]) else -(x[0] + 1) for x in res], dtype=np.int64)
                # Callbacks may have changed their element's state, parameters or time.
                b.slots_write(_lib.FIELD_STATE, lanes, self._hyx_state[:, lanes])
                if self._hyx_system.npar:
                    b.slots_write(_lib.FIELD_PARS, lanes, self._hyx_pars[: self._hyx_system.npar][:, lanes])
                b.slots_write(_lib.FIELD_TIME_HI, lanes, self._hyx_thi[lanes][None, :])
                b.slots_write(_lib.FIELD_TIME_LO, lanes, self._hyx_tlo[lanes][None, :])
                b.events_resume(lanes, stop)
                r = b.propagate_resume(max_steps, mode, grid_out)
                rounds += 1
        finally:
            b.set_event_pause(False)
        self._pull()
        self._hyx_prop_res = _PropRes.of(r)
        return rounds

    def step(self, max_delta_t=None, write_tc=False):
        """One adaptive step per lane; ``max_delta_t`` (per lane) clamps |h| and sets its sign."""
        if isinstance(max_delta_t, (bool, np.bool_)):
            write_tc, max_delta_t = bool(max_delta_t), None
        mdt = None if max_delta_t is None else self._lanes(max_delta_t, "max_delta_t vector")
        if mdt is not None and np.isnan(mdt).any():
            raise ValueError("A NaN max_delta_t was passed to the step() function of an adaptive Taylor integrator")
        self._step_impl(mdt, write_tc)

    def step_backward(self, write_tc=False):
        self._step_impl(np.full(self._hyx_B, -np.inf), write_tc)

    # ---- propagation --------------------------------------------------------------------
    def _check_cb(self, callback):
        if callback is None:
            return None
        if not callable(callback):
            raise TypeError(
                f"The object of type '{type(callback)}' cannot be used as a step callback because it is not callable")
        return callback

    def _prop_args(self, max_steps, max_delta_t):
        if not isinstance(max_steps, (int, np.integer)) or max_steps < 0:
            raise ValueError("max_steps must be a non-negative integer")
        B = self._hyx_B
        if max_delta_t is None or (not np.isscalar(max_delta_t) and len(max_delta_t) == 0):
            mdt = np.full(B, np.inf)
        else:
            mdt = self._lanes(max_delta_t, "max_delta_t vector")
        if np.isnan(mdt).any():
            raise ValueError("A nan max_delta_t was passed to the propagate_until() function of an adaptive Taylor "
                             "integrator")
        if (mdt <= 0).any():
            raise ValueError("A non-positive max_delta_t was passed to the propagate_until() function of an adaptive "
                             "Taylor integrator")
        return mdt

    def _prop_impl(self, tf_hi, tf_lo, max_steps, max_delta_t, callback, write_tc, c_output):
        mdt = self._prop_args(max_steps, max_delta_t)
        if not np.all(np.isfinite(tf_hi)) or not np.all(np.isfinite(self._hyx_thi)):
            raise ValueError("A non-finite time was passed to the propagate_until() function of an adaptive Taylor "
                             "integrator")
        cb = self._check_cb(callback)
        if c_output:
            return self._prop_c_output(tf_hi, tf_lo, int(max_steps), mdt, cb)
        if cb is None and not write_tc and not self._device_events_ok():
            self._push()
            self._prop_parked(lambda: self._hyx_batch.propagate_until(tf_hi, tf_lo, mdt, int(max_steps),
                                                                      _MODES[self._hyx_flags[3]]), int(max_steps))
            return None, None
        if cb is None and not write_tc:
            self._push()
            r = self._hyx_batch.propagate_until(tf_hi, tf_lo, mdt, int(max_steps),
                                                _MODES[self._hyx_flags[3]])
            self._pull()
            self._hyx_prop_res = _PropRes.of(r)
            return None, None
        return self._prop_stepwise(tf_hi, tf_lo, int(max_steps), mdt, cb, write_tc)

    def _prop_c_output(self, tf_hi, tf_lo, max_steps, mdt, cb):
        """propagate with c_output=True (continuous_output_batch): every step's Taylor
        coefficients and start time per batch element. On the device (a step log per
        instance, drained when full) unless a step callback, events or the group form
        need host-driven steps."""
        B = self._hyx_B
        t0 = self._hyx_thi + self._hyx_tlo
        n = self._hyx_state.shape[0]
        if cb is None and not self.with_events and self._hyx_system.form not in ("group", "wave"):
            return self._prop_c_output_device(tf_hi, tf_lo, max_steps, mdt, t0), None
        rec = [([], [], [], []) for _ in range(B)]

        def after_step(h, live, fin, tdir):
            th, tl = self._hyx_t_before
            for b in np.nonzero(live)[0]:
                if h[b] != 0.0 or not rec[b][2]:
                    r = rec[b]
                    r[0].append(th[b])
                    r[1].append(tl[b])
                    r[2].append(h[b])
                    r[3].append(self._hyx_tc[:, :, b].copy())

        _, cb = self._prop_stepwise(tf_hi, tf_lo, max_steps, mdt, cb, True, after_step)
        steps = [(np.array(a), np.array(c), np.array(d), np.array(e)) for a, c, d, e in rec]
        return continuous_output_batch._from_steps(steps, n, self._hyx_order, t0, self._hyx_thi + self._hyx_tlo), cb

    def _prop_c_output_device(self, tf_hi, tf_lo, max_steps, mdt, t0):
        b = self._hyx_batch
        B, n, P = self._hyx_B, self._hyx_state.shape[0], self._hyx_order
        mode = _MODES[self._hyx_flags[3]]
        # Log slots per instance: up to 64 steps, within about 1 GiB of device memory.
        cap = int(max(1, min(64, (1 << 30) // (B * n * (P + 1) * 8))))
        parts = [[] for _ in range(B)]

        def drain(lanes):
            nrec, tc, tt = b.clog_drain(lanes)
            for j, lane in enumerate(lanes):
                k = int(nrec[j])
                if k:
                    parts[lane].append((tt[:k, 0, j].copy(), tt[:k, 1, j].copy(), tt[:k, 2, j].copy(),
                                        np.moveaxis(tc[:k, :, :, j], 0, 0).copy()))

        b.set_clog(cap)
        try:
            self._push()
            r = b.propagate_until(tf_hi, tf_lo, mdt, max_steps, mode)
            while True:
                lanes = b.events_parked()
                if lanes.size == 0:
                    break
                drain(lanes)
                b.events_resume(lanes, np.zeros(lanes.size, dtype=np.int64))
                r = b.propagate_resume(max_steps, mode)
            drain(np.arange(B, dtype=np.uint32))
        finally:
            b.set_clog(0)
        self._pull()
        self._hyx_prop_res = _PropRes.of(r)
        steps = []
        for lane in range(B):
            ps = parts[lane]
            if ps:
                steps.append(tuple(np.concatenate([p[q] for p in ps]) for q in range(4)))
            else:
                steps.append((np.zeros(0), np.zeros(0), np.zeros(0), np.zeros((0, n, P + 1))))
        return continuous_output_batch._from_steps(steps, n, P, t0, self._hyx_thi + self._hyx_tlo,
                                                   final=self._hyx_state)

    def _device_events_ok(self):
        """Propagations run entirely on the device unless an event needs the host: every
        non-terminal event and every terminal event with a callback does."""
        return not self._hyx_ntes and all(e.callback is None for e in self._hyx_tes)

    def _prop_stepwise(self, tf_hi, tf_lo, max_steps, mdt, cb, write_tc, after_step=None):
        # Host-driven loop, one device step per iteration, so that a step
        # callback can observe and modify the integrator between steps.
        # after_step(h, live, fin, tdir) runs before the callback (dense output of propagate_grid).
        B = self._hyx_B
        rem0 = dfl_add(tf_hi, tf_lo, -self._hyx_thi, -self._hyx_tlo)[0]
        tdir = rem0 >= 0
        done = np.zeros(B, dtype=bool)
        res_oc = np.full(B, int(taylor_outcome.time_limit), dtype=np.int64)
        min_h = np.full(B, np.inf)
        max_h = np.zeros(B)
        nst = np.zeros(B, dtype=np.int64)
        if cb is not None and hasattr(cb, "pre_hook"):
            cb.pre_hook(self)
        it = 0
        while True:
            rem = dfl_add(tf_hi, tf_lo, -self._hyx_thi, -self._hyx_tlo)[0]
            lim = np.where(tdir, np.minimum(mdt, rem), np.maximum(-mdt, rem))
            lim = np.where(done, 0.0, lim)
            self.__dict__["_hyx_t_before"] = (self._hyx_thi.copy(), self._hyx_tlo.copy())
            oc, h = self._step_impl(lim, write_tc)
            it += 1
            act = ~done
            nst += (act & (h != 0)).astype(np.int64)
            # Only steps that ended normally count towards min_h / max_h: steps clamped by
            # the time limit or shortened by a terminal event (an event index, or
            # -(index + 1) when it stops the lane) do not.
            succ = act & (oc == int(taylor_outcome.success))
            min_h = np.where(succ, np.minimum(min_h, np.abs(h)), min_h)
            max_h = np.where(succ, np.maximum(max_h, np.abs(h)), max_h)
            nf = act & (oc == int(taylor_outcome.err_nf_state))
            res_oc[nf] = int(taylor_outcome.err_nf_state)
            done |= nf
            ste = act & (oc > int(taylor_outcome.success)) & (oc < 0)
            res_oc[ste] = oc[ste]
            done |= ste
            fin = act & ~nf & (oc == int(taylor_outcome.time_limit)) & (h == rem)
            done |= fin
            if after_step is not None:
                after_step(h, act & ~nf, fin, tdir)
            if max_steps and it >= max_steps:
                res_oc[act & ~done] = int(taylor_outcome.step_limit)
                done[:] = True
            stop = False
            if cb is not None:
                r = cb(self)
                if not isinstance(r, (bool, np.bool_)):
                    raise TypeError("The call operator of a step callback is expected to return a boolean, but a "
                                    f"value of type {type(r)} was returned instead")
                if not r:
                    res_oc[~done] = int(taylor_outcome.cb_stop)
                    stop = True
            if stop or done.all():
                break
        self._hyx_prop_res = _PropRes(res_oc, min_h, max_h, nst)
        return None, cb

    def propagate_until(self, t, max_steps=0, max_delta_t=None, callback=None, write_tc=False, c_output=False):
        tf = self._lanes(t, "final time vector")
        ret = self._prop_impl(tf, np.zeros(self._hyx_B), max_steps, max_delta_t, callback, write_tc, c_output)
        return ret

    def propagate_for(self, delta_t, max_steps=0, max_delta_t=None, callback=None, write_tc=False, c_output=False):
        dt = self._lanes(delta_t, "time interval vector")
        if not np.all(np.isfinite(dt)):
            raise ValueError("A non-finite time interval was passed to the propagate_for() function of an adaptive "
                             "Taylor integrator")
        hi, lo = dfl_add(self._hyx_thi, self._hyx_tlo, dt, np.zeros_like(dt))
        return self._prop_impl(hi, lo, max_steps, max_delta_t, callback, write_tc, c_output)

    def propagate_grid(self, grid, max_steps=0, max_delta_t=None, callback=None):
        """Propagate to the last grid time of each instance, returning (callback | None, out[n_pts, n, B]).

        heyoka/expose_batch_integrators.cpp:315-392. Per instance the grid must be finite, strictly
        monotonic and start at the current time (doc/notebooks/The adaptive integrator.ipynb, cell 26).
        The steps are those of propagate_until to the last grid time; the state at each grid time comes
        from the Taylor polynomials of the step containing it (dense output). Points not reached
        (step limit, non-finite state, a callback stop) are NaN (Batch mode overview.ipynb, cell 23)."""
        g = np.asarray(grid)
        if g.ndim != 2:
            raise ValueError(
                "Invalid grid passed to the propagate_grid() method of a batch integrator: the expected number of "
                f"dimensions is 2, but the input array has a dimension of {g.ndim}")
        if g.shape[1] != self._hyx_B:
            raise ValueError(
                "Invalid grid passed to the propagate_grid() method of a batch integrator: the shape must be "
                f"(n, {self._hyx_B}) but the number of columns is {g.shape[1]} instead")
        g = np.ascontiguousarray(g.astype(np.float64, casting="safe"))
        npts = g.shape[0]
        if npts == 0:
            raise ValueError("Cannot invoke propagate_grid() in an adaptive Taylor integrator in batch mode if the "
                             "time grid is empty")
        if not np.all(np.isfinite(g)):
            raise ValueError("A non-finite time value was passed to propagate_grid() in an adaptive Taylor integrator")
        if np.any(g[0] != self._hyx_thi):
            raise ValueError("When invoking propagate_grid(), the first element of the time grid must match the "
                             "current time coordinate of the integrator")
        if npts > 1:
            d = np.diff(g, axis=0)
            fwd = g[-1] > g[0]
            if not np.all(np.where(fwd, d > 0, d < 0)):
                raise ValueError("A non-monotonic time grid was passed to propagate_grid()")
        mdt = self._prop_args(max_steps, max_delta_t)
        cb = self._check_cb(callback)
        if cb is None and not self._device_events_ok():
            self._push()
            box = {}

            def first():
                r, box["out"] = self._hyx_batch.propagate_grid(g, mdt, int(max_steps), _MODES[self._hyx_flags[3]])
                return r

            out = np.empty((npts, self._hyx_state.shape[0], self._hyx_B))
            rounds = self._prop_parked(first, int(max_steps), grid_out=out)
            return None, (box["out"] if rounds == 0 else out)
        if cb is None:
            self._push()
            r, out = self._hyx_batch.propagate_grid(g, mdt, int(max_steps), _MODES[self._hyx_flags[3]])
            self._pull()
            self._hyx_prop_res = _PropRes.of(r)
            return None, out
        # With a step callback: host-driven steps with write_tc, dense output on the host.
        n, B = self._hyx_state.shape[0], self._hyx_B
        out = np.full((npts, n, B), np.nan)
        out[0] = self._hyx_state
        cur = np.ones(B, dtype=np.int64)

        def after_step(h, live, fin, tdir):
            tc = self._hyx_tc
            for b in np.nonzero(live)[0]:
                while cur[b] < npts:
                    d = (g[cur[b], b] - self._hyx_thi[b]) - self._hyx_tlo[b]
                    inside = d <= 0 if tdir[b] else d >= 0
                    if not inside and not fin[b]:
                        break
                    dt = d + h[b]
                    acc = tc[:, -1, b].copy()
                    for k in range(tc.shape[1] - 2, -1, -1):
                        acc = acc * dt + tc[:, k, b]
                    out[cur[b], :, b] = acc
                    cur[b] += 1

        _, cb = self._prop_stepwise(g[-1].copy(), np.zeros(B), int(max_steps), mdt, cb, True, after_step)
        return cb, out

    # ---- dense output from the last step's Taylor coefficients ----------------------------
    def update_d_output(self, t, rel_time=False):
        tt = self._lanes(t, "time vector")
        if rel_time:
            delta = tt + self._hyx_last_h
        else:
            delta = ((tt - self._hyx_thi) - self._hyx_tlo) + self._hyx_last_h
        tc = self._hyx_tc
        acc = tc[:, -1, :].copy()
        for k in range(tc.shape[1] - 2, -1, -1):
            acc = acc * delta + tc[:, k, :]
        self._hyx_dout[...] = acc
        return self.d_output

    # ---- properties -----------------------------------------------------------------------
    @property
    def state(self):
        return self._hyx_state

    @property
    def pars(self):
        return self._hyx_pars

    @staticmethod
    def _ro(a):
        v = a.view()
        v.flags.writeable = False
        return v

    @property
    def time(self):
        return self._ro(self._hyx_thi)

    @property
    def dtime(self):
        return self._ro(self._hyx_thi), self._ro(self._hyx_tlo)

    def set_time(self, time):
        t = self._lanes(time, "time vector")
        self._hyx_thi[...] = t
        self._hyx_tlo[...] = 0.0

    def set_dtime(self, hi_time, lo_time):
        if np.isscalar(hi_time) != np.isscalar(lo_time):
            raise TypeError("The two arguments to the set_dtime() method must be of the same type")
        hi = self._lanes(hi_time, "time vector")
        lo = self._lanes(lo_time, "time vector")
        h, l = _dfl_norm(hi, lo)
        self._hyx_thi[...] = h
        self._hyx_tlo[...] = l

    @property
    def last_h(self):
        return self._ro(self._hyx_last_h)

    @property
    def tc(self):
        return self._ro(self._hyx_tc)

    @property
    def d_output(self):
        return self._ro(self._hyx_dout)

    @property
    def propagate_res(self):
        if isinstance(self._hyx_prop_res, _PropRes):
            self._hyx_prop_res = self._hyx_prop_res.tolist()
        return list(self._hyx_prop_res)

    @property
    def step_res(self):
        return list(self._hyx_step_res)

    @property
    def order(self):
        return self._hyx_order

    @property
    def tol(self):
        return self._hyx_tol

    @property
    def dim(self):
        return self._hyx_dec.n_state

    @property
    def batch_size(self):
        return self._hyx_B

    @property
    def decomposition(self):
        """heyoka's taylor_dc_t: (expression, hidden_deps) pairs (n variables, m operations,
        n right-hand sides)."""
        return self._hyx_dec.as_heyoka()

    @property
    def sys(self):
        return list(self._hyx_pairs)

    @property
    def high_accuracy(self):
        return self._hyx_flags[0]

    @property
    def compact_mode(self):
        return self._hyx_flags[1]

    @property
    def with_events(self):
        return bool(self._hyx_tes or self._hyx_ntes)

    @property
    def t_events(self):
        return [_event_ref(e, self) for e in self._hyx_tes]

    @property
    def nt_events(self):
        return [_event_ref(e, self) for e in self._hyx_ntes]

    @property
    def te_cooldowns(self):
        """Per batch element, per terminal event: None, or (trigger time, cooldown length)."""
        B = self._hyx_B
        if not self._hyx_tes:
            return [[] for _ in range(B)]
        cd = self._hyx_cooldowns()
        return [[(float(cd[e, 0, b] + cd[e, 1, b]), float(cd[e, 2, b])) if cd[e, 2, b] >= 0 else None
                 for e in range(len(self._hyx_tes))] for b in range(B)]

    def reset_cooldowns(self, i=None):
        """Clear the cooldowns of every batch element, or of batch element i."""
        if not self._hyx_tes:
            return
        cd = self._hyx_cooldowns()
        if i is None:
            cd[:, 2, :] = -1.0
        else:
            if not 0 <= int(i) < self._hyx_B:
                raise ValueError(f"Cannot reset the cooldowns at batch index {i}: the batch size for this integrator "
                                 f"is only {self._hyx_B}")
            cd[:, 2, int(i)] = -1.0
        self._hyx_set_cooldowns(cd)

    # ---- variational --------------------------------------------------------------------
    @property
    def is_variational(self):
        return self._hyx_vsys is not None

    @property
    def n_orig_sv(self):
        return self._hyx_vsys.n_orig_sv if self._hyx_vsys is not None else self.dim

    @property
    def vargs(self):
        return self._hyx_vsys.vargs if self._hyx_vsys is not None else []

    @property
    def vorder(self):
        return self._hyx_vsys.order if self._hyx_vsys is not None else 0

    def _need_var(self):
        if self._hyx_vsys is None:
            raise ValueError("The function cannot be invoked on a non-variational integrator")

    def get_vslice(self, order, component=None):
        self._need_var()
        return self._hyx_vsys.get_vslice(order, component)

    def get_mindex(self, i):
        self._need_var()
        return self._hyx_vsys.get_mindex(i)

    @property
    def tstate(self):
        """The state of the original variables given by the last eval_taylor_map (read-only)."""
        self._need_var()
        if "_hyx_tstate" not in self.__dict__:
            self.__dict__["_hyx_tstate"] = np.zeros((self._hyx_vsys.n_orig_sv, self._hyx_B))
        return self._ro(self._hyx_tstate)

    def eval_taylor_map(self, inputs):
        """Evaluate the Taylor map of the variational state at the displacements
        ``inputs`` ([n_vargs, batch_size]): x_i + sum over multi-indices m of
        d^m x_i / m! * prod_j dx_j^m_j (``heyoka/expose_batch_integrators.cpp:573-649``)."""
        self._need_var()
        vs = self._hyx_vsys
        if isinstance(inputs, np.ndarray) and inputs.dtype != np.float64:
            raise TypeError("Invalid dtype detected for the inputs of a Taylor map evaluation: the expected dtype is "
                            f"'float64', but the dtype of the inputs array is '{inputs.dtype}' instead")
        a = np.asarray(inputs, dtype=np.float64)
        nv = len(vs.vargs)
        if a.ndim != 2:
            raise ValueError(f"The array of inputs provided for the evaluation of a Taylor map has {a.ndim} "
                             "dimension(s), but it must have 2 dimensions instead")
        if a.shape[0] != nv:
            raise ValueError(f"The array of inputs provided for the evaluation of a Taylor map has {a.shape[0]} "
                             f"row(s), but it must have {nv} row(s) instead")
        if a.shape[1] != self._hyx_B:
            raise ValueError(f"The array of inputs provided for the evaluation of a Taylor map has {a.shape[1]} "
                             f"column(s), but it must have {self._hyx_B} column(s) instead")
        from math import factorial

        out = self._hyx_state[: vs.n_orig_sv].copy()
        for idx in range(vs.n_orig_sv, self.dim):
            mi = vs.get_mindex(idx)
            i, m = mi[0], mi[1:]
            term = self._hyx_state[idx].copy()
            for j, e in enumerate(m):
                if e:
                    term = term * a[j] ** e / factorial(e)
            out[i] += term
        if "_hyx_tstate" not in self.__dict__:
            self.__dict__["_hyx_tstate"] = np.zeros((vs.n_orig_sv, self._hyx_B))
        self._hyx_tstate[...] = out
        return self.tstate

    # ---- GPU specifics ------------------------------------------------------------------
    @property
    def code_form(self):
        return self._hyx_system.form

    def last_kernel_timing(self):
        """(milliseconds on the device, kernel launches) of the last step/propagate call."""
        return self._hyx_batch.last_timing()

    # ---- copy / pickle --------------------------------------------------------------------
    _STATE_KEYS = ("_hyx_state", "_hyx_pars", "_hyx_thi", "_hyx_tlo", "_hyx_last_h", "_hyx_prop_res",
                   "_hyx_step_res", "_hyx_tc", "_hyx_dout")

    def _user_attrs(self):
        return {k: v for k, v in self.__dict__.items() if not k.startswith("_hyx_")}

    def _events_copy(self):
        return ([_copy.copy(e) for e in self._hyx_tes], [_copy.copy(e) for e in self._hyx_ntes])

    # Shared by copies: the system, its decomposition and the compiled code do not change
    # after construction.
    _SHARED_KEYS = ("_hyx_pairs", "_hyx_vsys", "_hyx_dec", "_hyx_order", "_hyx_tol", "_hyx_flags", "_hyx_system",
                    "_hyx_B", "_hyx_npar")

    def _clone(self, user):
        """A copy that shares the decomposition and the compiled system, with its own host
        state and (on first use) its own device batch."""
        new = taylor_adaptive_batch.__new__(taylor_adaptive_batch)
        d, src = new.__dict__, self.__dict__
        for k in self._SHARED_KEYS:
            d[k] = src[k]
        for k in self._STATE_KEYS:
            v = src[k]
            d[k] = v.copy() if isinstance(v, (np.ndarray, _PropRes)) else list(v)
        d["_hyx_tes"], d["_hyx_ntes"] = self._events_copy()
        d["_hyx_batch_obj"] = None
        d["_hyx_cd_pending"] = self._hyx_cooldowns() if self._hyx_tes else None
        d.update(user)
        return new

    def __copy__(self):
        return self._clone(self._user_attrs())

    def __deepcopy__(self, memo):
        return self._clone(_copy.deepcopy(self._user_attrs(), memo))

    def __getstate__(self):
        st = {k: self.__dict__[k] for k in self._STATE_KEYS}
        if isinstance(st["_hyx_prop_res"], _PropRes):
            st["_hyx_prop_res"] = st["_hyx_prop_res"].tolist()
        st["_hyx_ctor"] = (self._hyx_vsys if self._hyx_vsys is not None else self._hyx_pairs, self._hyx_tol,
                           self._hyx_flags)
        st["_hyx_user"] = self._user_attrs()
        st["_hyx_events"] = (self._hyx_tes, self._hyx_ntes)
        st["_hyx_cooldowns"] = self._hyx_cooldowns() if self._hyx_tes else None
        return st

    def __setstate__(self, st):
        sysarg, tol, f = st["_hyx_ctor"]
        restore = {k: st[k] for k in self._STATE_KEYS}
        self._hyx_init(sysarg, restore["_hyx_state"], restore["_hyx_thi"], restore["_hyx_pars"], tol, f[0], f[1], f[2],
                       f[3], f[4], restore=restore, events=st.get("_hyx_events", ([], [])))
        if st.get("_hyx_cooldowns") is not None:
            self._hyx_set_cooldowns(st["_hyx_cooldowns"])
        self.__dict__.update(st["_hyx_user"])

    def __repr__(self):
        return (f"GPU batch Taylor integrator (FP64)\nTolerance               : {self._hyx_tol!r}\n"
                f"High accuracy           : {str(self.high_accuracy).lower()}\nTaylor order            : "
                f"{self._hyx_order}\nDimension               : {self.dim}\nBatch size              : {self._hyx_B}\n"
                f"Code form               : {self.code_form}\n")
