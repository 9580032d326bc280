"""Ensemble propagation: ``ensemble_propagate_until/for/grid[_batch]``.

Same signature, argument validation and return structure as the reference
(``heyoka/__init__.py:251-336``, ``heyoka/_ensemble_impl.py:13-138``):
``gen(deepcopy(ta), i)`` builds member i and the result is
``[(ta_i, *ret_i) for i in range(n_iter)]``.

On the GPU the members are not run one by one. Their states, parameters and
times are packed side by side into one batch of ``n_iter * batch_size``
instances, propagated in one persistent launch, and unpacked into the member
integrators. Instances are independent and their arithmetic does not depend on
where they sit in the batch, so every member ends bit-for-bit where its own
``propagate_until`` would have taken it (tests/test_gpu_ensemble.py checks this).
Members with a step callback, or whose events call back into Python, are
propagated one at a time instead; terminal events without callbacks stop
instances inside the packed launch (cooldowns are packed with the members).
``ensemble_propagate_grid[_batch]`` splats the 1-D grid over every instance
(``_ensemble_impl.py:13-19``) and returns ``(ta_i, None, out_i)`` with
``out_i`` of shape (n_pts, n, batch_size), cut from one device-side
propagate_grid over the packed batch.
"""

from __future__ import annotations

from collections.abc import Iterable, Sequence
from copy import copy, deepcopy

import numpy as np

from . import _lib
from .integrator import taylor_adaptive_batch, dfl_add, _MODES, _PropRes


def _ensemble_generic(tp, ta, arg, n_iter, gen, **kwargs):
    if not isinstance(n_iter, int):
        raise TypeError(
            f"The n_iter parameter must be an integer, but an object of type {type(n_iter)} was provided instead")
    if n_iter < 0:
        raise ValueError(f"The n_iter parameter must be non-negative, but it is {n_iter} instead")
    if tp in ("until", "for"):
        if isinstance(arg, Iterable):
            raise TypeError("Cannot perform an ensemble propagate_until/for(): the final epoch/time interval must be "
                            "a scalar, not an iterable object")
    else:
        arg = np.array(arg)
        if arg.ndim != 1:
            raise ValueError("Cannot perform an ensemble propagate_grid(): the input time grid must be "
                             f"one-dimensional, but instead it has {arg.ndim} dimensions")
    if "max_delta_t" in kwargs and isinstance(kwargs["max_delta_t"], Iterable):
        raise TypeError('Cannot perform an ensemble propagate_until/for/grid(): the "max_delta_t" argument must be a '
                        "scalar, not an iterable object")
    algo = kwargs.pop("algorithm", "thread")
    allowed = ["thread", "process"]
    if algo not in allowed:
        raise ValueError(f"The parallelisation algorithm must be one of {allowed}, but '{algo}' was provided instead")
    kwargs.pop("max_workers", None)
    if algo == "process":
        kwargs.pop("chunksize", None)
    if "chunksize" in kwargs:
        raise TypeError("propagate_until() got an unexpected keyword argument 'chunksize'")
    if ("callback" in kwargs and kwargs["callback"] is not None) or kwargs.get("c_output"):
        return _ensemble_sequential(tp, ta, arg, n_iter, gen, kwargs)
    if isinstance(ta, taylor_adaptive_batch) and not ta._device_events_ok():
        # Event callbacks receive their own member integrator: one member at a time.
        return _ensemble_sequential(tp, ta, arg, n_iter, gen, kwargs)
    return _ensemble_packed(tp, ta, arg, n_iter, gen, kwargs)


def _splat_grid(grid, ta):
    """The ensemble's 1-D grid as the (n_pts, batch_size) grid of one member (heyoka/_ensemble_impl.py:13-19)."""
    return np.repeat(np.asarray(grid, dtype=np.float64), ta.batch_size).reshape((-1, ta.batch_size))


def _ensemble_sequential(tp, ta, arg, n_iter, gen, kwargs):
    out = []
    for i in range(n_iter):
        kw = copy(kwargs)
        if kwargs.get("callback") is not None:
            kw["callback"] = deepcopy(kwargs["callback"])
        local = gen(deepcopy(ta), i)
        if tp == "until":
            ret = local.propagate_until(arg, **kw)
        elif tp == "for":
            ret = local.propagate_for(arg, **kw)
        else:
            ret = local.propagate_grid(_splat_grid(arg, local), **kw)
        out.append((local,) + tuple(ret))
    return out


# Device time of the last packed ensemble propagation: kernel milliseconds (HIP events on
# its batch's stream), launches, members and instances.
last_run: dict = {}


class ensemble_arrays:
    """Vectorised generator for the ensemble functions: member i takes columns
    ``i * B .. (i + 1) * B - 1`` of ``state`` ([n, n_iter * B]), ``pars``
    ([n_par, n_iter * B]) and ``time`` ([n_iter * B]), where B is the template's batch
    size; arrays left as None keep the template's values.

    It is an ordinary generator (``gen(ta, i)`` writes member i's columns into ``ta``),
    so it works wherever a generator does. The packed GPU path recognises it and writes
    the arrays to the device directly, with no copy of the template per member; the
    members of the result are then built on first access.
    """

    def __init__(self, state=None, pars=None, time=None):
        self.state = None if state is None else np.ascontiguousarray(state, dtype=np.float64)
        self.pars = None if pars is None else np.ascontiguousarray(pars, dtype=np.float64)
        self.time = None if time is None else np.ascontiguousarray(time, dtype=np.float64)
        for a, nd in ((self.state, 2), (self.pars, 2), (self.time, 1)):
            if a is not None and a.ndim != nd:
                raise ValueError(f"ensemble_arrays: expected a {nd}-dimensional array, got shape {a.shape}")

    def _columns(self, ta, n_iter):
        """(state, pars, time hi, time lo) of all members side by side, validated against ta."""
        n, N = ta.dim, n_iter * ta.batch_size

        def cols(a, rows, what):
            if a.shape[-1] != N or (rows is not None and a.shape[0] != rows):
                raise ValueError(f"ensemble_arrays: the {what} array has shape {a.shape}, but the ensemble needs "
                                 f"{'(' + str(rows) + ', ' if rows is not None else '('}{N})")
            return a

        st = np.tile(ta._hyx_state, n_iter) if self.state is None else cols(self.state, n, "state")
        npar = ta._hyx_system.npar
        if self.pars is None:
            pa = np.tile(ta._hyx_pars[:npar], n_iter) if npar else None
        else:
            pa = cols(self.pars, None, "pars")
            if pa.shape[0] < npar:
                raise ValueError(f"ensemble_arrays: the pars array has {pa.shape[0]} rows, but the system has "
                                 f"{npar} parameters")
            pa = pa[:npar] if npar else None
        if self.time is None:
            thi, tlo = np.tile(ta._hyx_thi, n_iter), np.tile(ta._hyx_tlo, n_iter)
        else:
            thi, tlo = cols(self.time, None, "time").copy(), np.zeros(N)
        return st, pa, thi, tlo

    def __call__(self, ta, i):
        B = ta.batch_size
        sl = slice(i * B, (i + 1) * B)
        if self.state is not None:
            ta.state[:] = self.state[:, sl]
        if self.pars is not None:
            ta.pars[: self.pars.shape[0]] = self.pars[:, sl]
        if self.time is not None:
            ta.set_time(self.time[sl])
        return ta


class _packed_members(Sequence):
    """The result list of a vectorised ensemble propagation: entry i is
    ``(ta_i, None, None)`` (``(ta_i, None, out_i)`` for a grid), with ta_i a copy of
    the template holding member i's final state, built on first access."""

    def __init__(self, ta, n_iter, st, pa, thi, tlo, r, cds, gout):
        # A snapshot of the template: later changes to it do not reach the members.
        self._ta, self._n = deepcopy(ta), n_iter
        self._st, self._pa, self._thi, self._tlo, self._r, self._cds, self._gout = st, pa, thi, tlo, r, cds, gout
        self._made = {}

    def __len__(self):
        return self._n

    def __getitem__(self, i):
        if isinstance(i, slice):
            return [self[j] for j in range(*i.indices(self._n))]
        i = int(i)
        if i < 0:
            i += self._n
        if not 0 <= i < self._n:
            raise IndexError("ensemble result index out of range")
        if i not in self._made:
            B = self._ta.batch_size
            m = deepcopy(self._ta)
            sl = slice(i * B, (i + 1) * B)
            if self._pa is not None:
                m._hyx_pars[: self._pa.shape[0]] = self._pa[:, sl]
            _unpack_member(m, sl, self._st, self._thi, self._tlo, self._r, self._cds)
            self._made[i] = (m, None, self._gout[:, :, sl].copy()) if self._gout is not None else (m, None, None)
        return self._made[i]


def _unpack_member(m, sl, st, thi, tlo, r, cds):
    if cds is not None:
        m._hyx_set_cooldowns(cds[:, :, sl])
    m._hyx_state[...] = st[:, sl]
    m._hyx_thi[...] = thi[sl]
    m._hyx_tlo[...] = tlo[sl]
    m._hyx_prop_res = _PropRes.of(r, sl)


def _ensemble_packed(tp, ta, arg, n_iter, gen, kwargs):
    if n_iter == 0:
        return []
    if not isinstance(ta, taylor_adaptive_batch):
        raise TypeError("The template integrator must be a taylor_adaptive_batch")
    max_steps = int(kwargs.pop("max_steps", 0))
    mdt = kwargs.pop("max_delta_t", None)
    kwargs.pop("callback", None)
    kwargs.pop("c_output", None)
    if kwargs.pop("write_tc", False):
        raise NotImplementedError("write_tc is not supported by ensemble propagations on the GPU")
    if kwargs:
        raise TypeError(f"Unexpected keyword argument(s): {sorted(kwargs)}")
    B = ta.batch_size
    n = ta.dim
    N = n_iter * B
    npar = ta._hyx_system.npar
    vectorised = isinstance(gen, ensemble_arrays)
    if vectorised:
        members = None
        st, pa, thi, tlo = gen._columns(ta, n_iter)
        cd_in = np.tile(ta._hyx_cooldowns(), n_iter) if ta._hyx_tes else None
    else:
        members = [gen(deepcopy(ta), i) for i in range(n_iter)]
        for m in members:
            if m.batch_size != B or m.dim != n or m._hyx_system is not ta._hyx_system:
                raise ValueError("The generator must return integrators of the same system and batch size as the "
                                 "template")
        st = np.concatenate([m._hyx_state for m in members], axis=1)
        pa = np.concatenate([m._hyx_pars[:npar] for m in members], axis=1) if npar else None
        thi = np.concatenate([m._hyx_thi for m in members])
        tlo = np.concatenate([m._hyx_tlo for m in members])
        cd_in = np.concatenate([m._hyx_cooldowns() for m in members], axis=2) if ta._hyx_tes else None
    if tp == "until":
        tf_hi, tf_lo = np.full(N, float(arg)), np.zeros(N)
    elif tp == "for":
        tf_hi, tf_lo = dfl_add(thi, tlo, np.full(N, float(arg)), np.zeros(N))
    else:
        grid = np.repeat(np.asarray(arg, dtype=np.float64), N).reshape((-1, N))
        if grid.shape[0] == 0:
            raise ValueError("Cannot invoke propagate_grid() in an adaptive Taylor integrator in batch mode if the "
                             "time grid is empty")
        if not np.all(np.isfinite(grid)):
            raise ValueError("A non-finite time value was passed to propagate_grid() in an adaptive Taylor "
                             "integrator")
        if np.any(grid[0] != thi):
            raise ValueError("When invoking propagate_grid(), the first element of the time grid must match the "
                             "current time coordinate of the integrator")
        if grid.shape[0] > 1:
            d = np.diff(grid[:, 0])
            if not (np.all(d > 0) or np.all(d < 0)):
                raise ValueError("A non-monotonic time grid was passed to propagate_grid()")
        tf_hi = grid[-1].copy()
    if mdt is None:
        md = np.full(N, np.inf)
    else:
        md = np.full(N, float(mdt))
        if np.isnan(md).any() or (md <= 0).any():
            raise ValueError("A non-positive or nan max_delta_t was passed to an ensemble propagation")
    if tp != "grid" and not np.all(np.isfinite(tf_hi)):
        raise ValueError("A non-finite time was passed to an ensemble propagation")
    batch = _lib.Batch(ta._hyx_system, N)
    last_run["members"], last_run["instances"] = n_iter, N
    batch.write(_lib.FIELD_STATE, st)
    if npar:
        batch.write(_lib.FIELD_PARS, pa)
    batch.write(_lib.FIELD_TIME_HI, thi)
    batch.write(_lib.FIELD_TIME_LO, tlo)
    if cd_in is not None:
        # Terminal events without callbacks stop instances on the device; the members'
        # cooldowns travel with them.
        batch.set_cooldowns(cd_in)
    gout = None
    if tp == "grid":
        r, gout = batch.propagate_grid(grid, md, max_steps, _MODES[ta._hyx_flags[3]])
    else:
        r = batch.propagate_until(tf_hi, tf_lo, md, max_steps, _MODES[ta._hyx_flags[3]])
    last_run["kernel_ms"], last_run["launches"] = batch.last_timing()
    last_run["instance_steps"] = int(r["nsteps"].astype(np.int64).sum())
    st = batch.read(_lib.FIELD_STATE, (n, N))
    thi = batch.read(_lib.FIELD_TIME_HI, (N,))
    tlo = batch.read(_lib.FIELD_TIME_LO, (N,))
    cds = batch.cooldowns() if ta._hyx_tes else None
    if vectorised:
        return _packed_members(ta, n_iter, st, pa, thi, tlo, r, cds, gout)
    out = []
    for i, m in enumerate(members):
        sl = slice(i * B, (i + 1) * B)
        _unpack_member(m, sl, st, thi, tlo, r, cds)
        out.append((m, None, gout[:, :, sl].copy()) if tp == "grid" else (m, None, None))
    return out


def ensemble_propagate_until(ta, t, n_iter, gen, **kwargs):
    return _ensemble_generic("until", ta, t, n_iter, gen, **kwargs)


def ensemble_propagate_for(ta, delta_t, n_iter, gen, **kwargs):
    return _ensemble_generic("for", ta, delta_t, n_iter, gen, **kwargs)


def ensemble_propagate_grid(ta, grid, n_iter, gen, **kwargs):
    return _ensemble_generic("grid", ta, grid, n_iter, gen, **kwargs)


def ensemble_propagate_until_batch(ta, t, n_iter, gen, **kwargs):
    return _ensemble_generic("until", ta, t, n_iter, gen, **kwargs)


def ensemble_propagate_for_batch(ta, delta_t, n_iter, gen, **kwargs):
    return _ensemble_generic("for", ta, delta_t, n_iter, gen, **kwargs)


def ensemble_propagate_grid_batch(ta, grid, n_iter, gen, **kwargs):
    return _ensemble_generic("grid", ta, grid, n_iter, gen, **kwargs)
