"""``continuous_output_batch``: dense output over a whole batch propagation.

Mirrors heyoka.py's class (``heyoka/taylor_expose_c_output.cpp:269-520``; tests
``heyoka/test.py:1413-1680``; doc/notebooks/Batch mode overview.ipynb cells 30-36):
``propagate_until/for(..., c_output=True)`` returns it, and calling it with a
time (a scalar for every batch element, a ``[B]`` array, or ``[m, B]`` rows of
times) evaluates the solution there from the Taylor coefficients of the step
that contains each time. Times outside a batch element's integration interval
use its first or last step.

Each batch element keeps its own steps (the GPU does not step the elements in
lockstep); ``times`` ``(n_steps + 1, B)`` and ``tcs`` ``(n_steps, n, p + 1, B)``
pad the elements with fewer steps with zero-length steps at their final time, as
finished elements of the reference's lockstep loop take them.
"""

from __future__ import annotations

import copy as _copy

import numpy as np

_EMPTY = "Cannot use a default-constructed continuous_output_batch object"


class _View(np.ndarray):
    """A read-only array that keeps its continuous_output_batch alive (the reference
    returns views whose base is the object)."""

    def __array_finalize__(self, obj):
        self._owner = getattr(obj, "_owner", None)


def _ro_view(arr, owner):
    v = arr.view(_View)
    v._owner = owner
    v.flags.writeable = False
    return v


class continuous_output_batch:
    def __init__(self):
        self._B = 0
        self._steps = None  # per element: (t_hi [s], t_lo [s], h [s], tc [s, n, p+1])
        self._out = None
        self._n = 0
        self._order = 0

    @classmethod
    def _from_steps(cls, steps, n, order, t0, t1, final=None):
        c = cls()
        c._final = np.zeros((n, len(steps))) if final is None else np.array(final, dtype=np.float64)
        c._B = len(steps)
        c._steps = steps
        c._n, c._order = n, order
        c._out = np.zeros((n, c._B))
        c._bounds = (np.array(t0, dtype=np.float64), np.array(t1, dtype=np.float64))
        return c

    # ---- evaluation ---------------------------------------------------------------------
    def _eval_row(self, tm, out):
        for b, (thi, tlo, h, tc) in enumerate(self._steps):
            s = len(h)
            if s == 0:
                # No step taken (already at the final time): the solution is the state there.
                out[:, b] = self._final[:, b]
                continue
            fwd = not (h[0] < 0)
            t = float(tm[b])
            # The step whose interval holds t; before the start the first, after the end the last.
            ends = thi + h  # step end times (hi parts), monotonic in the direction of integration
            j = int(np.searchsorted(ends, t, side="right")) if fwd else int(np.searchsorted(-ends, -t, side="right"))
            j = min(max(j, 0), s - 1)
            dt = (t - thi[j]) - tlo[j]
            acc = tc[j, :, -1].copy()
            for k in range(tc.shape[2] - 2, -1, -1):
                acc = acc * dt + tc[j, :, k]
            out[:, b] = acc

    def __call__(self, time):
        if self._steps is None:
            raise ValueError(_EMPTY)
        if np.ndim(time) == 0:
            self._eval_row(np.full(self._B, float(time)), self._out)
            return _ro_view(self._out, self)
        tm = np.asarray(time)
        if tm.dtype != np.float64:
            tm = tm.astype(np.float64, casting="safe")
        if tm.ndim not in (1, 2):
            raise ValueError("Invalid time array passed to a continuous_output_batch object: the number of dimensions "
                             f"must be 1 or 2, but it is {tm.ndim} instead")
        if tm.ndim == 1:
            if tm.shape[0] != self._B:
                raise ValueError("Invalid time array passed to a continuous_output_batch object: the length must be "
                                 f"{self._B} but it is {tm.shape[0]} instead")
            self._eval_row(tm, self._out)
            return _ro_view(self._out, self)
        if tm.shape[1] != self._B:
            raise ValueError("Invalid time array passed to a continuous_output_batch object: the number of columns "
                             f"must be {self._B} but it is {tm.shape[1]} instead")
        ret = np.empty((tm.shape[0], self._n, self._B))
        for i in range(tm.shape[0]):
            self._eval_row(tm[i], ret[i])
        return ret

    # ---- properties -----------------------------------------------------------------------
    @property
    def output(self):
        return None if self._steps is None else _ro_view(self._out, self)

    @property
    def n_steps(self):
        if self._steps is None:
            raise ValueError(_EMPTY)
        return max(len(h) for _, _, h, _ in self._steps)

    def _padded(self):
        ns = self.n_steps
        times = np.empty((ns + 1, self._B))
        tcs = np.zeros((ns, self._n, self._order + 1, self._B))
        for b, (thi, tlo, h, tc) in enumerate(self._steps):
            s = len(h)
            times[0, b] = thi[0] + tlo[0] if s else self._bounds[0][b]
            times[1:s + 1, b] = thi + tlo + h
            tcs[:s, :, :, b] = tc
            if s == 0:
                times[1:, b] = times[0, b]
                tcs[:, :, 0, b] = self._final[:, b]
            elif s < ns:
                end = times[s, b]
                times[s + 1:, b] = end
                # Zero-length steps at the final state.
                last = np.zeros((self._n, self._order + 1))
                acc = tc[-1, :, -1].copy()
                dt = h[-1]
                for k in range(self._order - 1, -1, -1):
                    acc = acc * dt + tc[-1, :, k]
                last[:, 0] = acc
                tcs[s:, :, :, b] = last
        return times, tcs

    @property
    def times(self):
        if self._steps is None:
            return None
        return _ro_view(self._padded()[0], self)

    @property
    def tcs(self):
        if self._steps is None:
            return None
        return _ro_view(self._padded()[1], self)

    @property
    def bounds(self):
        if self._steps is None:
            raise ValueError(_EMPTY)
        return self._bounds[0].copy(), self._bounds[1].copy()

    @property
    def batch_size(self):
        return self._B

    def __repr__(self):
        if self._steps is None:
            return "Default-constructed continuous_output_batch"
        fwd = all(not (len(h) and h[0] < 0) for _, _, h, _ in self._steps)
        return (f"Directions : [{', '.join('forward' if fwd else 'backward' for _ in range(self._B))}]\n"
                f"Time ranges: {list(zip(self._bounds[0].tolist(), self._bounds[1].tolist()))}\n"
                f"N of steps : {self.n_steps}\n")

    # ---- copies / pickling ----------------------------------------------------------------
    def __copy__(self):
        c = continuous_output_batch.__new__(continuous_output_batch)
        c.__dict__.update(self.__dict__)
        c._out = None if self._out is None else self._out.copy()
        return c

    def __deepcopy__(self, memo):
        c = continuous_output_batch.__new__(continuous_output_batch)
        for k, v in self.__dict__.items():
            c.__dict__[k] = _copy.deepcopy(v, memo)
        return c

    def __getstate__(self):
        return dict(self.__dict__)

    def __setstate__(self, st):
        self.__dict__.update(st)


continuous_output_batch_dbl = continuous_output_batch
