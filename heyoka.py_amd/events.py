"""Batch event classes: ``t_event_batch``, ``nt_event_batch`` and ``event_direction``.

Mirror the classes heyoka.py binds in ``heyoka/taylor_expose_events.cpp:195-300``
(construction arguments, the callable check and its TypeError text, the
``expression`` / ``callback`` / ``direction`` / ``cooldown`` properties, copies and
pickling). As in the reference, the callback is deep-copied when an event is
created or copied (doc/notebooks/Event detection.ipynb, cell 6).

Callback signatures (batch mode, doc/notebooks/Batch mode overview.ipynb cell 37):
non-terminal ``cb(ta, time, d_sgn, batch_idx)``; terminal
``cb(ta, d_sgn, batch_idx) -> bool`` (False, or no callback, stops the
integration of that batch element).
"""

from __future__ import annotations

import copy as _copy
import enum

import numpy as np

from .expr import _ex


class event_direction(enum.IntEnum):
    """heyoka's event_direction: the sign of the event function's time derivative at a trigger."""

    negative = -1
    any = 0
    positive = 1


def _direction(d):
    try:
        return event_direction(int(d))
    except (ValueError, TypeError):
        raise ValueError(f"Invalid value for the event direction: {d!r}") from None


def _check_callable(cb):
    if not callable(cb):
        raise TypeError(f"An object of type '{type(cb)}' cannot be used as an event callback because it is not "
                        "callable")


class nt_event_batch:
    """Non-terminal event: ``callback(ta, time, d_sgn, batch_idx)`` runs at every trigger."""

    def __init__(self, expression, callback, direction=event_direction.any, fp_type=float):
        if fp_type not in (float, np.float64):
            raise TypeError(f'The floating-point type "{fp_type}" is not recognized/supported')
        _check_callable(callback)
        self._ex = _ex(expression)
        self._cb = _copy.deepcopy(callback)
        self._dir = _direction(direction)

    @property
    def expression(self):
        return self._ex

    @property
    def callback(self):
        return self._cb

    @property
    def direction(self):
        return self._dir

    def __repr__(self):
        return (f"C++ datatype: double\nEvent type     : non-terminal\nEvent equation : {self._ex}\n"
                f"Event direction: {self._dir.name}\n")

    def __copy__(self):
        return nt_event_batch(self._ex, self._cb, self._dir)

    def __deepcopy__(self, memo):
        return nt_event_batch(self._ex, self._cb, self._dir)

    def __getstate__(self):
        return {"ex": self._ex, "cb": self._cb, "dir": int(self._dir)}

    def __setstate__(self, st):
        self._ex, self._cb, self._dir = st["ex"], st["cb"], event_direction(st["dir"])


class t_event_batch:
    """Terminal event: the step ends at the trigger; ``callback(ta, d_sgn, batch_idx)``
    (optional) returns whether the integration of that batch element continues."""

    def __init__(self, expression, callback=None, direction=event_direction.any, cooldown=-1.0, fp_type=float):
        if fp_type not in (float, np.float64):
            raise TypeError(f'The floating-point type "{fp_type}" is not recognized/supported')
        # The reference binds cooldown without implicit conversion: a float is required.
        if not isinstance(cooldown, (float, np.floating)):
            raise TypeError(f"The cooldown of a terminal event must be a float, not {type(cooldown)}")
        if callback is not None:
            _check_callable(callback)
        self._ex = _ex(expression)
        self._cb = None if callback is None else _copy.deepcopy(callback)
        self._dir = _direction(direction)
        self._cd = float(cooldown)

    @property
    def expression(self):
        return self._ex

    @property
    def callback(self):
        return self._cb

    @property
    def direction(self):
        return self._dir

    @property
    def cooldown(self):
        return self._cd

    def __repr__(self):
        cd = "auto" if self._cd < 0 else repr(self._cd)
        return (f"C++ datatype: double\nEvent type     : terminal\nEvent equation : {self._ex}\n"
                f"Event direction: {self._dir.name}\nWith callback  : {'yes' if self._cb is not None else 'no'}\n"
                f"Cooldown       : {cd}\n")

    def __copy__(self):
        return t_event_batch(self._ex, self._cb, self._dir, self._cd)

    def __deepcopy__(self, memo):
        return t_event_batch(self._ex, self._cb, self._dir, self._cd)

    def __getstate__(self):
        return {"ex": self._ex, "cb": self._cb, "dir": int(self._dir), "cd": self._cd}

    def __setstate__(self, st):
        self._ex, self._cb, self._dir, self._cd = st["ex"], st["cb"], event_direction(st["dir"]), st["cd"]


class _event_ref:
    """What ``ta.t_events`` / ``ta.nt_events`` hand out: a view of the integrator's own
    event that keeps the integrator alive, as the reference's properties do."""

    __slots__ = ("_ev", "_owner")

    def __init__(self, ev, owner):
        self._ev = ev
        self._owner = owner

    def __getattr__(self, name):
        return getattr(self._ev, name)

    def __repr__(self):
        return repr(self._ev)

    def __copy__(self):
        return _copy.copy(self._ev)

    def __deepcopy__(self, memo):
        return _copy.deepcopy(self._ev, memo)
