"""The wave form on the GPU: the decomposition's nodes spread over the lanes of a
group (64, 32 or 16 lanes per instance), every coefficient row in LDS, run by the
group kernels.

Against the f64 / real128 oracle (final states within 1e-10 relative on the
well-conditioned instances, step counts within one), against the compact form
(the same recurrences in the same order: states within 1e-12), and through the
paths the group kernels carry: propagate_grid, step(write_tc), events, all three
propagation modes.
"""

import numpy as np
import pytest

import heyoka_amd as hy
import pyoracle as O
import events as OE
from heyoka_amd import workloads as W
from test_gpu_parity import _gpu_vs_oracle, _nbody3

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("form", ["wave", "wave32"])
def test_j2var_wave_vs_oracle(form):
    print(f"[gpu] j2var 64 instances t=5 code_form={form}", flush=True)
    ta, _, _ = _gpu_vs_oracle(W.j2var(64), 64, 5.0, 1e-10, code_form=form)
    assert ta.code_form == "wave"


@pytest.mark.parametrize("mode", ["persistent", "plain", "graph"])
def test_cr3bp_wave_modes(mode):
    print(f"[gpu] cr3bp wave form mode={mode}", flush=True)
    ta, _, _ = _gpu_vs_oracle(W.cr3bp(100), 100, 10.0, 1e-10, code_form="wave", mode=mode)
    assert ta.code_form == "wave"


@pytest.mark.parametrize("form", ["wave", "wave16"])
def test_nbody3_wave(form):
    print(f"[gpu] nbody3 code_form={form}", flush=True)
    _gpu_vs_oracle(_nbody3(), 37, 3.0, 1e-10, code_form=form)


def test_all_recurrences_wave_vs_compact():
    # Every recurrence kind (sqrt, div, exp, log, pow, sin / cos, time, parameters,
    # constants, linear nodes): the wave form against the compact form and the oracle.
    print("[gpu] all recurrence kinds, wave vs compact", flush=True)
    x, v, y = hy.make_vars("x", "v", "y")
    sys_ = [(x, v),
            (v, -hy.sqrt(x * x + 1.5) / (1.0 + y * y) + hy.exp(-x) * hy.par[0] + hy.log(2.0 + y * y) * hy.time),
            (y, hy.cos(x) * hy.par[1] - (v * v + 0.5) ** 1.5 + hy.time)]
    rng = np.random.default_rng(5)
    B = 70
    st = np.stack([rng.uniform(-0.5, 0.5, B), rng.uniform(-0.2, 0.2, B), rng.uniform(-0.5, 0.5, B)])
    pars = np.stack([rng.uniform(0.1, 0.5, B), rng.uniform(0.5, 1.5, B)])
    runs = {}
    for form in ("wave", "compact"):
        ta = hy.taylor_adaptive_batch(sys_, st, pars=pars, tol=1e-15, code_form=form)
        assert ta.code_form == form
        ta.propagate_until(2.0)
        runs[form] = ta
    w, c = runs["wave"], runs["compact"]
    np.testing.assert_allclose(w.state, c.state, rtol=1e-12, atol=1e-14)
    assert all(a[0] == hy.taylor_outcome.time_limit for a in w.propagate_res)
    ns_w = np.array([r[3] for r in w.propagate_res])
    ns_c = np.array([r[3] for r in c.propagate_res])
    assert np.max(np.abs(ns_w - ns_c)) <= 1
    dec = hy.taylor_decompose(sys_)
    r = O.propagate_until(dec, w.order, st, np.full(B, 2.0), pars=pars, nthreads=8, lane_batch=1)
    np.testing.assert_allclose(w.state, r["state"], rtol=1e-10, atol=1e-12)


def test_wave_step_tc_and_grid():
    print("[gpu] wave form step(write_tc) and propagate_grid", flush=True)
    wl = W.j2var(40)
    ta = hy.taylor_adaptive_batch(wl.sys, wl.state, tol=1e-15, code_form="wave")
    ta.step(write_tc=True)
    dec = hy.taylor_decompose(wl.sys.sys)
    r = O.step_batch(dec, ta.order, wl.state, want_tc=True)
    scale = np.max(np.abs(r["tc"]), axis=1, keepdims=True) + 1e-300
    assert np.max(np.abs(ta.tc - r["tc"]) / scale) < 1e-12
    # h = rho * (|x| / |x^[p]|)^(1/p): the top orders' few-ulp differences over ~400 nodes.
    np.testing.assert_allclose([h for _, h in ta.step_res], r["h"], rtol=1e-12)
    grid = np.repeat(np.linspace(0.0, 20.0, 11), 40).reshape(11, 40)
    tg = hy.taylor_adaptive_batch(wl.sys, wl.state, tol=1e-15, code_form="wave")
    tc_ = hy.taylor_adaptive_batch(wl.sys, wl.state, tol=1e-15, code_form="compact")
    _, og = tg.propagate_grid(grid)
    _, oc = tc_.propagate_grid(grid)
    np.testing.assert_allclose(og, oc, rtol=1e-11, atol=1e-12)


def test_wave_events_vs_oracle():
    # A terminal event on the J2 orbit's z coordinate (equator crossings upwards) and a
    # non-terminal one on x: the device detection inside the wave form's step.
    print("[gpu] wave form events", flush=True)
    wl = W.j2var(24)
    xs = wl.sys.sys[0][0]
    zs = wl.sys.sys[2][0]
    hits = []
    ta = hy.taylor_adaptive_batch(wl.sys, wl.state, tol=1e-15, code_form="wave",
                                  t_events=[hy.t_event_batch(zs, direction=hy.event_direction.positive)],
                                  nt_events=[hy.nt_event_batch(xs, lambda ta, t, d, b: hits.append((b, t, d)))])
    assert ta.code_form == "wave"
    ta.propagate_until(np.full(24, 30.0))
    tb = hy.taylor_adaptive_batch(wl.sys, wl.state, tol=1e-15, code_form="compact",
                                  t_events=[hy.t_event_batch(zs, direction=hy.event_direction.positive)],
                                  nt_events=[hy.nt_event_batch(xs, lambda ta, t, d, b: None)])
    tb.propagate_until(np.full(24, 30.0))
    oc_w = [int(r[0]) for r in ta.propagate_res]
    oc_c = [int(r[0]) for r in tb.propagate_res]
    assert oc_w == oc_c and -1 in oc_w
    np.testing.assert_allclose(ta.time, tb.time, rtol=1e-12)
    stopped = np.array(oc_w) == -1
    assert np.all(np.abs(ta.state[2, stopped]) < 1e-12)
    dec = hy.taylor_decompose(wl.sys.sys)
    ref = OE.propagate(hy, wl.sys.sys, wl.state[:, :6], 30.0, t_events=[(zs, 1, -1.0, True)], nt_events=[(xs, 0)])
    assert list(ref["outcome"]) == oc_w[:6]
    np.testing.assert_allclose(ta.time[:6], ref["t_hi"], rtol=1e-12)
    for b in range(6):
        got = sorted(t for bb, t, _ in hits if bb == b)
        want = [t for _, t, _ in ref["nt"][b]]
        np.testing.assert_allclose(got, want, rtol=1e-12)


def test_wave_backward_and_continuous_output():
    # Backward propagation and c_output=True (host-driven steps for the group kernels) in the
    # wave form, against the compact form.
    print("[gpu] wave form backward propagation and continuous output", flush=True)
    wl = W.j2var(20)
    res = {}
    for form in ("wave", "compact"):
        ta = hy.taylor_adaptive_batch(wl.sys, wl.state, tol=1e-15, code_form=form)
        ta.propagate_until(np.full(20, -6.0))
        assert all(p[0] == hy.taylor_outcome.time_limit for p in ta.propagate_res)
        back = ta.state.copy()
        tb = hy.taylor_adaptive_batch(wl.sys, wl.state, tol=1e-15, code_form=form)
        co, _ = tb.propagate_for(np.full(20, 4.0), c_output=True)
        res[form] = (back, tb.state.copy(), co(np.full(20, 2.5)).copy())
    for a, b in zip(res["wave"], res["compact"]):
        np.testing.assert_allclose(a, b, rtol=1e-11, atol=1e-12)
