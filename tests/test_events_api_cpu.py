"""Event classes and the event op-list on the CPU (no GPU): construction checks and
their error texts (heyoka/taylor_expose_events.cpp:195-300), properties, copies,
pickling, and the generated event code compiling for gfx950."""

import copy
import pickle

import numpy as np
import pytest

import heyoka_amd as hy
from heyoka_amd import _lib
from heyoka_amd.integrator import _event_spec

x, v = hy.make_vars("x", "v")
PEND = [(x, v), (v, -9.8 * hy.sin(x))]


class _cb:
    def __init__(self):
        self.lst = []

    def __call__(self, *args):
        return True


def test_event_classes():
    c = _cb()
    ev = hy.nt_event_batch(v * v - 1e-10, c, direction=hy.event_direction.negative)
    assert ev.direction == hy.event_direction.negative and str(ev.expression) == str(v * v - 1e-10)
    assert ev.callback is not c and ev.callback.lst is not c.lst  # deep-copied on construction
    assert id(ev.callback) == id(ev.callback)
    with pytest.raises(TypeError) as e:
        hy.nt_event_batch(v, 3)
    assert "cannot be used as an event callback because it is not callable" in str(e.value)
    te = hy.t_event_batch(v)
    assert te.callback is None and te.cooldown == -1.0 and te.direction == hy.event_direction.any
    with pytest.raises(TypeError):
        hy.t_event_batch(v, cooldown=1)  # cooldown is not converted implicitly
    with pytest.raises(TypeError):
        hy.t_event_batch(v, callback="nope")
    te = hy.t_event_batch(x, callback=c, cooldown=0.5, direction=hy.event_direction.positive)
    for e2 in (copy.copy(te), copy.deepcopy(te), pickle.loads(pickle.dumps(te))):
        assert e2.cooldown == 0.5 and e2.direction == hy.event_direction.positive
        assert e2.callback is not te.callback
    assert "terminal" in repr(te) and "non-terminal" in repr(ev)
    assert [int(d) for d in hy.event_direction] == [-1, 0, 1]


def test_event_oplist_and_compile():
    dec = hy.taylor_decompose(PEND)
    tes = [hy.t_event_batch(v), hy.t_event_batch(x - 0.1, cooldown=0.25)]
    ntes = [hy.nt_event_batch(v * v - 1e-10, _cb()), hy.nt_event_batch(hy.sin(x) * hy.par[1], _cb())]
    spec = _event_spec(dec, tes, ntes, 1e-15)
    assert spec.n_te == 2 and spec.n_nt == 2 and list(spec.cooldowns) == [-1.0, 0.25]
    # Four events over two state variables: two padding rows.
    assert spec.dec.n_state == 4 and spec.dec.n_par == 2
    with pytest.raises(ValueError):
        _event_spec(dec, [hy.t_event_batch(hy.make_vars("z"))], [], 1e-15)
    dec.n_par = 2
    for compact in (False, True):
        s = _lib.System(dec, 19, compact=compact, events=spec, source_only=True)
        src = s.source()
        assert "hyx_ev_norms" in src and "#define HYX_NEV 4" in src and "#define HYX_NTE 2" in src
        _lib.System(dec, 19, compact=compact, events=spec, compile_only=True)
    with pytest.raises(_lib.HyxError):
        _lib.System(dec, 19, compact="group", events=spec, source_only=True)
