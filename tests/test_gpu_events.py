"""Event detection on the GPU: the reference's batch event tests, the notebook
goldens, and parity with the CPU oracle (oracle/events.py).

Tolerances: event times of well-conditioned events (|dg/dt| of order one) agree
with the oracle and the notebooks to a few ulps of the time; the kernels refine
roots with a different bracketing method and evaluate with fused multiply-adds,
so 1e-14 relative is allowed. Roots of v^2 - eps (|dg/dt| ~ 1e-6 at the root)
are ill-conditioned and agree to ~1e-12.
"""

import copy
import sys

import numpy as np
import pytest

import heyoka_amd as hy
import events as OE
from heyoka_amd import taylor_outcome as to

pytestmark = pytest.mark.gpu

x, v = hy.make_vars("x", "v")
PEND = [(x, v), (v, -9.8 * hy.sin(x))]


def test_reference_events_batch():
    # heyoka/_test_batch_integrator.py:555-600
    print("[gpu] reference test_events", flush=True)

    def cb0(ta, t, d_sgn, bidx):
        pass

    ta = hy.taylor_adaptive_batch(sys=PEND, state=np.array([[0.0, 0.001], [0.25, 0.2501]]),
                                  nt_events=[hy.nt_event_batch(v * v - 1e-6, cb0)], t_events=[hy.t_event_batch(v)])
    assert ta.with_events and len(ta.t_events) == 1 and len(ta.nt_events) == 1
    ta.propagate_until([1e9, 1e9])
    assert all(int(r[0]) == -1 for r in ta.propagate_res)
    assert ta.te_cooldowns[0][0] is not None and ta.te_cooldowns[1][0] is not None
    ta.reset_cooldowns(0)
    assert ta.te_cooldowns[0][0] is None and ta.te_cooldowns[1][0] is not None
    ta.reset_cooldowns()
    assert ta.te_cooldowns[0][0] is None and ta.te_cooldowns[1][0] is None
    # Stopped at the first zero of v, as the oracle.
    ref = OE.propagate(hy, PEND, [[0.0, 0.001], [0.25, 0.2501]], 1e9, t_events=[(v, 0, -1.0, True)],
                       nt_events=[(v * v - 1e-6, 0)])
    np.testing.assert_allclose(ta.time, ref["t_hi"], rtol=1e-14)
    np.testing.assert_allclose(ta.state, ref["state"], rtol=0, atol=1e-15)
    assert [r[3] for r in ta.propagate_res] == list(ref["nsteps"])


def test_reference_nt_counts_and_identity():
    # heyoka/test.py:693-800
    print("[gpu] reference nt event counts", flush=True)
    counter = [0, 0]
    cur = [0.0, 0.0]
    ta_id = [None]
    seen = []

    def cb0(ta, t, d_sgn, bidx):
        assert t > cur[bidx]
        assert counter[bidx] % 3 in (0, 2)
        assert ta_id[0] == id(ta)
        counter[bidx] += 1
        cur[bidx] = t
        seen.append((bidx, 0, t))

    def cb1(ta, t, d_sgn, bidx):
        assert t > cur[bidx]
        assert counter[bidx] % 3 == 1
        assert ta_id[0] == id(ta)
        counter[bidx] += 1
        cur[bidx] = t
        seen.append((bidx, 1, t))

    ta = hy.taylor_adaptive_batch(sys=PEND, state=[[0.0, 0.001], [0.25, 0.2501]],
                                  nt_events=[hy.nt_event_batch(v * v - 1e-10, cb0), hy.nt_event_batch(v, cb1)])
    ta_id[0] = id(ta)
    ta.propagate_until([4.0, 4.0])
    assert all(r[0] == to.time_limit for r in ta.propagate_res)
    assert counter == [12, 12]
    ref = OE.propagate(hy, PEND, [[0.0, 0.001], [0.25, 0.2501]], 4.0, nt_events=[(v * v - 1e-10, 0), (v, 0)])
    for b in range(2):
        mine = [(i, t) for bb, i, t in seen if bb == b]
        assert [i for i, _ in mine] == [i for i, _, _ in ref["nt"][b]]
        for (i, t), (_, rt, _) in zip(mine, ref["nt"][b]):
            assert abs(t - rt) <= (1e-14 * abs(rt) if i == 1 else 1e-11), (b, i, t, rt)
    assert [r[3] for r in ta.propagate_res] == list(ref["nsteps"])


def test_reference_event_objects():
    # heyoka/test.py:770-860 and :905-1000: property identity, refcounts, deep copies,
    # callback errors.
    print("[gpu] reference event objects", flush=True)

    class cbn:
        def __init__(self):
            self.lst = []

        def __call__(self, ta, t, d_sgn, bidx):
            pass

    class cbt:
        def __init__(self):
            self.lst = []

        def __call__(self, ta, d_sgn, bidx):
            pass

    for kind, mk in (("nt_events", lambda e, c: hy.nt_event_batch(e, c)),
                     ("t_events", lambda e, c: hy.t_event_batch(e, callback=c))):
        c = cbn if kind == "nt_events" else cbt
        ta = hy.taylor_adaptive_batch(sys=PEND, state=[[0.0, 0.001], [0.25, 0.2501]],
                                      **{kind: [mk(v * v - 1e-10, c()), mk(v, c()), mk(v, c())]})
        rc = sys.getrefcount(ta)
        lst = getattr(ta, kind)
        assert sys.getrefcount(ta) == rc + 3
        del lst
        for i in range(3):
            evs = getattr(ta, kind)
            assert id(evs[i].callback) == id(getattr(ta, kind)[i].callback)
            assert id(evs[i].callback.lst) == id(getattr(ta, kind)[i].callback.lst)
        tc = copy.deepcopy(ta)
        for i in range(3):
            assert id(getattr(tc, kind)[i].callback) != id(getattr(ta, kind)[i].callback)
            assert id(getattr(tc, kind)[i].callback.lst) != id(getattr(ta, kind)[i].callback.lst)

    def bad(ta, t):
        pass

    ta = hy.taylor_adaptive_batch(sys=PEND, state=[[0.0, 0.001], [0.25, 0.2501]],
                                  nt_events=[hy.nt_event_batch(v * v - 1e-10, bad)])
    with pytest.raises(RuntimeError):
        ta.propagate_until([4.0, 4.0])
    ta = hy.taylor_adaptive_batch(sys=PEND, state=[[0.0, 0.001], [0.25, 0.2501]],
                                  t_events=[hy.t_event_batch(v * v - 1e-10, callback=bad)])
    with pytest.raises(RuntimeError):
        ta.propagate_until([4.0, 4.0])

    def cb3(ta, d_sgn, bidx):
        return "hello"

    ta = hy.taylor_adaptive_batch(sys=PEND, state=[[0.0, 0.001], [0.25, 0.2501]],
                                  t_events=[hy.t_event_batch(v * v - 1e-10, callback=cb3)])
    with pytest.raises(RuntimeError) as e:
        ta.propagate_until([4.0, 4.0])
    assert "in the construction of the return value of an event callback" in str(e.value)


def test_reference_terminal_steps():
    # heyoka/test.py:860-905: terminal event with a continuing callback plus a
    # non-terminal event, driven by step().
    print("[gpu] reference terminal steps", flush=True)
    counter_t = [0, 0]
    counter_nt = [0, 0]
    cur = [0.0, 0.0]
    ta_id = [None]

    def cb0(ta, t, d_sgn, bidx):
        assert t > cur[bidx] and ta_id[0] == id(ta)
        counter_nt[bidx] += 1
        cur[bidx] = t

    def cb1(ta, d_sgn, bidx):
        assert ta.time[bidx] > cur[bidx] and ta_id[0] == id(ta)
        counter_t[bidx] += 1
        cur[bidx] = ta.time[bidx]
        return True

    ta = hy.taylor_adaptive_batch(sys=PEND, state=[[0.0, 0.001], [0.25, 0.2501]],
                                  nt_events=[hy.nt_event_batch(v * v - 1e-10, cb0)],
                                  t_events=[hy.t_event_batch(v, callback=cb1)])
    ta_id[0] = id(ta)
    while True:
        ta.step()
        if all(r[0] > to.success for r in ta.step_res):
            break
    assert all(int(r[0]) == 0 for r in ta.step_res)
    assert all(t < 1 for t in ta.time)
    assert counter_nt == [1, 1] and counter_t == [1, 1]
    while True:
        ta.step()
        if all(r[0] > to.success for r in ta.step_res):
            break
    assert all(int(r[0]) == 0 for r in ta.step_res)
    assert all(t > 1 for t in ta.time)
    assert counter_nt == [3, 3] and counter_t == [2, 2]


@pytest.mark.parametrize("code_form", ["unrolled", "compact"])
def test_notebook_goldens(goldens, code_form):
    # doc/notebooks/Event detection.ipynb cells 5-21 (scalar pendulum, as batch lanes).
    print(f"[gpu] event notebook goldens code_form={code_form}", flush=True)
    g = goldens["ev_zero_vel_times"]
    times, xs = [], []

    def cb(ta, t, d_sgn, bidx):
        ta.update_d_output([t, t])
        xs.append(ta.d_output[0, bidx])
        times.append(t)

    ic = np.repeat(np.array(g["ic"])[:, None], 2, axis=1)
    ta = hy.taylor_adaptive_batch(PEND, ic, nt_events=[hy.nt_event_batch(v, cb)], code_form=code_form)
    _, out = ta.propagate_grid(np.repeat(np.linspace(0, 5, 1000), 2).reshape(1000, 2))
    assert len(times) == 10
    np.testing.assert_allclose(times[0::2], g["times"], rtol=1e-15, atol=0)
    np.testing.assert_allclose(times[1::2], g["times"], rtol=1e-15, atol=0)
    np.testing.assert_allclose(xs[0::2], goldens["ev_x_at_zero_vel"]["x"], rtol=0, atol=5e-17)
    assert np.isfinite(out).all()
    gp = goldens["ev_positive_times"]
    pt = []
    ta = hy.taylor_adaptive_batch(PEND, ic[:, :1], code_form=code_form, nt_events=[
        hy.nt_event_batch(v, lambda ta, t, d, b: pt.append(t), direction=hy.event_direction.positive)])
    ta.propagate_until(gp["t_end"])
    np.testing.assert_allclose(pt, gp["times"], rtol=1e-15, atol=0)
    g2 = goldens["ev_two_events"]
    log = []
    ta = hy.taylor_adaptive_batch(PEND, ic[:, :1], code_form=code_form, nt_events=[
        hy.nt_event_batch(v, lambda ta, t, d, b: log.append((0, t))),
        hy.nt_event_batch(v * v - 1e-12, lambda ta, t, d, b: log.append((1, t)))])
    ta.propagate_until(g2["t_end"])
    assert [i for i, _ in log] == [i for i, _ in g2["events"]]
    for (i, t), (_, gt) in zip(log, g2["events"]):
        assert abs(t - gt) <= (4e-16 * abs(gt) if i == 0 else 2e-12), (i, t, gt)
    oc, mn, mx, ns = ta.propagate_res[0]
    assert oc == to.time_limit and ns == g2["nsteps"]
    np.testing.assert_allclose([mn, mx], [g2["min_h"], g2["max_h"]], rtol=1e-15)


@pytest.mark.parametrize("code_form", ["unrolled", "compact"])
@pytest.mark.parametrize("mode", ["persistent", "plain"])
def test_device_terminal_events_vs_oracle(code_form, mode):
    # Terminal events without callbacks run entirely on the device: every lane stops at
    # its first accepted trigger with outcome -(index + 1). Directions and two
    # competing events; times, states and step counts against the oracle.
    print(f"[gpu] device terminal events code_form={code_form} mode={mode}", flush=True)
    rng = np.random.default_rng(5)
    B = 96
    ic = np.stack([rng.uniform(-0.8, 0.8, B), rng.uniform(-1.0, 1.0, B)])
    tes = [hy.t_event_batch(x - 0.3, direction=hy.event_direction.positive),
           hy.t_event_batch(v + 1.2, direction=hy.event_direction.negative)]
    ta = hy.taylor_adaptive_batch(PEND, ic, t_events=tes, code_form=code_form, gpu_mode=mode)
    assert ta.code_form == code_form
    ta.propagate_until(np.full(B, 6.0))
    ref = OE.propagate(hy, PEND, ic, 6.0, t_events=[(x - 0.3, 1, -1.0, True), (v + 1.2, -1, -1.0, True)])
    oc = np.array([int(r[0]) for r in ta.propagate_res])
    np.testing.assert_array_equal(oc, ref["outcome"])
    assert set(oc) <= {-1, -2, int(to.time_limit)} and (oc == -1).any()
    np.testing.assert_allclose(ta.time, ref["t_hi"], rtol=1e-13, atol=1e-14)
    np.testing.assert_allclose(ta.state, ref["state"], rtol=0, atol=1e-13)
    assert [r[3] for r in ta.propagate_res] == list(ref["nsteps"])
    stopped = oc == -1
    np.testing.assert_allclose(ta.state[0, stopped], 0.3, atol=1e-14)


def test_device_grid_with_stopping_event():
    print("[gpu] grid with stopping event", flush=True)
    ic = np.array([[0.05, 0.2, -0.3], [0.0, 0.5, 0.1]])
    ta = hy.taylor_adaptive_batch(PEND, ic, t_events=[hy.t_event_batch(x - 0.1)])
    grid = np.repeat(np.linspace(0.0, 3.0, 31), 3).reshape(31, 3)
    _, out = ta.propagate_grid(grid)
    oc = [int(r[0]) for r in ta.propagate_res]
    ref = OE.propagate(hy, PEND, ic, 3.0, t_events=[(x - 0.1, 0, -1.0, True)])
    assert oc == list(ref["outcome"])
    for b in range(3):
        reach = np.isfinite(out[:, 0, b])
        if oc[b] == -1:
            # Points after the trigger are NaN; those before it are not.
            last = np.nonzero(reach)[0].max()
            assert grid[last, b] <= ta.time[b] < grid[min(last + 1, 30), b] or last == 30
            assert not reach[last + 1:].any()
        else:
            assert reach.all()
    # The same through the host path (a step callback forces it).
    tb = hy.taylor_adaptive_batch(PEND, ic, t_events=[hy.t_event_batch(x - 0.1)])
    _, o2 = tb.propagate_grid(grid, callback=lambda t: True)
    np.testing.assert_allclose(o2, out, rtol=0, atol=1e-14)


def test_backward_and_cooldowns():
    print("[gpu] backward events and cooldowns", flush=True)
    log = []
    ic = np.array([[0.1, -0.2], [0.0, 0.3]])
    ta = hy.taylor_adaptive_batch(PEND, ic, nt_events=[hy.nt_event_batch(x, lambda ta, t, d, b: log.append((b, t, d)))])
    ta.propagate_until([-3.0, -3.0])
    ref = OE.propagate(hy, PEND, ic, -3.0, nt_events=[(x, 0)])
    for b in range(2):
        mine = [(t, d) for bb, t, d in log if bb == b]
        assert len(mine) == len(ref["nt"][b]) > 0
        for (t, d), (_, rt, rd) in zip(mine, ref["nt"][b]):
            assert d == rd and abs(t - rt) <= 1e-14 * max(1.0, abs(rt))
        assert all(np.diff([t for t, _ in mine]) < 0)
    # A continuing terminal event: the cooldown keeps it from re-firing at the restart.
    hits = []

    def cbt(ta, d, b):
        hits.append((b, ta.time[b]))
        return True

    ta = hy.taylor_adaptive_batch(PEND, ic, t_events=[hy.t_event_batch(x, callback=cbt)])
    ta.propagate_until([4.0, 4.0])
    assert all(r[0] == to.time_limit for r in ta.propagate_res)
    for b in range(2):
        ts = [t for bb, t in hits if bb == b]
        assert len(ts) >= 3 and np.min(np.diff(ts)) > 0.5
    cd = ta.te_cooldowns
    assert len(cd) == 2 and len(cd[0]) == 1


@pytest.mark.parametrize("mode", ["persistent", "plain"])
def test_parked_path_matches_host_stepped(mode):
    # Events whose callbacks need Python run in the propagate kernels, parking an
    # instance at the end of each step with triggers; the host-stepped path (forced by a
    # step callback) gives the same triggers, states and step counts bit for bit.
    print(f"[gpu] parked vs host-stepped mode={mode}", flush=True)
    rng = np.random.default_rng(11)
    B = 40
    ic = np.stack([rng.uniform(-1.0, 1.0, B), rng.uniform(-1.5, 1.5, B)])
    logs = ([], [])

    def mk(log):
        def nt(ta, t, d, b):
            log.append(("nt", b, t, d, ta.time[b]))

        def tcb(ta, d, b):
            log.append(("te", b, ta.time[b], d))
            ta.pars[0, b] += 1.0
            return ta.pars[0, b] < 3.0

        return ([hy.t_event_batch(x - 0.2, callback=tcb, direction=hy.event_direction.positive)],
                [hy.nt_event_batch(v, nt), hy.nt_event_batch(x * v - 0.05, nt)])

    sysp = [(x, v), (v, -9.8 * hy.sin(x) - 0.01 * hy.par[0] * v)]
    runs = []
    for k, log in enumerate(logs):
        te, nt = mk(log)
        ta = hy.taylor_adaptive_batch(sysp, ic, pars=np.zeros((1, B)), t_events=te, nt_events=nt, gpu_mode=mode)
        if k == 0:
            ta.propagate_until(np.full(B, 8.0))
        else:
            ta.propagate_until(np.full(B, 8.0), callback=lambda t: True)
        runs.append(ta)
    a, b = runs
    np.testing.assert_array_equal(a.state, b.state)
    np.testing.assert_array_equal(a.pars, b.pars)
    assert a.propagate_res == b.propagate_res
    for lane in range(B):
        la = [e for e in logs[0] if e[1] == lane]
        lb = [e for e in logs[1] if e[1] == lane]
        assert la == lb, lane
    ocs = [int(r[0]) for r in a.propagate_res]
    assert -1 in ocs and int(to.time_limit) in ocs


def test_parked_grid_and_large_batch():
    print("[gpu] parked grid / large batch", flush=True)
    B = 4096
    rng = np.random.default_rng(3)
    ic = np.stack([rng.uniform(-1.0, 1.0, B), rng.uniform(-1.0, 1.0, B)])
    counts = np.zeros(B, dtype=int)
    last = np.full(B, -np.inf)

    def cb(ta, t, d, b):
        assert t > last[b] and d == 1
        last[b] = t
        counts[b] += 1

    ta = hy.taylor_adaptive_batch(PEND, ic, nt_events=[hy.nt_event_batch(v, cb, direction=hy.event_direction.positive)])
    grid = np.repeat(np.linspace(0.0, 10.0, 21), B).reshape(21, B)
    _, out = ta.propagate_grid(grid)
    assert np.isfinite(out).all() and all(r[0] == to.time_limit for r in ta.propagate_res)
    # Every lane crosses v = 0 upwards once per period (about 2 s at these amplitudes).
    assert counts.min() >= 4 and counts.max() <= 6
    plain = hy.taylor_adaptive_batch(PEND, ic)
    _, o2 = plain.propagate_grid(grid)
    sub = slice(0, 64)
    ref = OE.propagate(hy, PEND, ic[:, sub], 10.0, nt_events=[(v, 1)])
    assert [len(e) for e in ref["nt"]] == list(counts[sub])
    np.testing.assert_allclose(ta.state[:, sub], ref["state"], rtol=0, atol=1e-13)
    # The events change the step sizes (their norms join the step's), not the solution.
    np.testing.assert_allclose(out, o2, rtol=0, atol=1e-12)


def test_ensembles_with_events():
    print("[gpu] ensembles with events", flush=True)
    ic = np.array([[0.05, 0.2], [0.0, 0.4]])

    def gen(t, i):
        t.state[:] = ic + 0.01 * i
        return t

    # Stopping terminal events: one packed launch, members equal to their own runs.
    ta = hy.taylor_adaptive_batch(PEND, ic, t_events=[hy.t_event_batch(x - 0.1)])
    ret = hy.ensemble_propagate_until_batch(ta, 5.0, 4, gen)
    for i, (m, _, _) in enumerate(ret):
        one = hy.taylor_adaptive_batch(PEND, ic + 0.01 * i, t_events=[hy.t_event_batch(x - 0.1)])
        one.propagate_until(5.0)
        np.testing.assert_array_equal(m.state, one.state)
        assert m.propagate_res == one.propagate_res
        assert m.te_cooldowns == one.te_cooldowns
    # Non-terminal events with callbacks: each member's callback sees its own integrator.
    seen = {}

    class cb:
        def __call__(self, t_, t, d, b):
            seen.setdefault(id(t_), []).append(t)

    ta = hy.taylor_adaptive_batch(PEND, ic, nt_events=[hy.nt_event_batch(v, cb())])
    ret = hy.ensemble_propagate_until_batch(ta, 5.0, 3, gen)
    assert len(seen) == 3 and all(id(m) in seen for m, _, _ in ret)


def test_group_form_events_vs_compact_and_oracle():
    # The 6-body ensemble keeps the group form (16 lanes per instance) with events: the
    # event jet comes from the state rows in LDS, the group detects together and
    # re-evaluates its owned variables at a terminal trigger.
    print("[gpu] group form events", flush=True)
    from heyoka_amd import workloads as W
    rng = np.random.default_rng(21)
    B = 64
    wl = W.nbody6(B)
    nm = hy.taylor_decompose(wl.sys).state_names
    x0, y0, x1, y1 = (hy.expression(nm[i]) for i in (0, 1, 6, 7))
    pars = rng.uniform(-0.9, 0.9, (1, B))
    ev = x1 - x0 - hy.par[0]
    runs = {}
    for form in ("group", "compact"):
        ta = hy.taylor_adaptive_batch(wl.sys, wl.state, pars=pars, tol=1e-15, code_form=form,
                                      t_events=[hy.t_event_batch(ev, direction=hy.event_direction.positive)])
        assert ta.code_form == form
        ta.propagate_until(np.full(B, 5.0))
        runs[form] = ta
    g, c = runs["group"], runs["compact"]
    og = [int(r[0]) for r in g.propagate_res]
    assert og == [int(r[0]) for r in c.propagate_res]
    assert og.count(-1) > B // 4 and og.count(int(to.time_limit)) > 0
    np.testing.assert_allclose(g.time, c.time, rtol=1e-12)
    scale = np.max(np.abs(c.state), axis=1, keepdims=True)
    assert np.max(np.abs(g.state - c.state) / scale) < 1e-11
    stopped = np.array(og) == -1
    np.testing.assert_allclose((g.state[6] - g.state[0])[stopped], pars[0, stopped], atol=1e-13)
    ref = OE.propagate(hy, wl.sys, wl.state[:, :4], 5.0, pars=pars[:, :4], tol=1e-15,
                       t_events=[(ev, 1, -1.0, True)])
    assert og[:4] == list(ref["outcome"])
    np.testing.assert_allclose(g.time[:4], ref["t_hi"], rtol=1e-12)
    assert [r[3] for r in g.propagate_res[:4]] == list(ref["nsteps"])
    # Non-terminal events with a Python callback: the parked path in the group form.
    logs = {}
    for form in ("group", "compact"):
        log = []
        ta = hy.taylor_adaptive_batch(wl.sys, wl.state, pars=pars, tol=1e-15, code_form=form,
                                      nt_events=[hy.nt_event_batch(y1 - y0, lambda t_, t, d, b: log.append((b, t, d)))])
        ta.propagate_until(np.full(B, 5.0))
        logs[form] = sorted(log)
    assert len(logs["group"]) == len(logs["compact"]) > 0
    for (bg, tg_, dg), (bc, tc_, dc) in zip(logs["group"], logs["compact"]):
        assert bg == bc and dg == dc and abs(tg_ - tc_) <= 1e-12 * max(1.0, abs(tc_))


def _wavy_ramp(log):
    # doc/notebooks/Batch mode overview.ipynb cell 46: a ball bouncing on a wavy ramp and on
    # the ground; terminal events whose callbacks reflect the velocity and continue.
    px, py, vx, vy = hy.make_vars("x", "y", "vx", "vy")
    eqns = [(px, vx), (py, vy), (vx, hy.expression(0.0)), (vy, hy.expression(-9.8))]
    eq_curve = py - (1.0 - px + 0.05 * hy.cos(11 * np.pi * px))
    CR = 0.8
    last = {}

    def cb_curve(ta, d_sgn, b):
        if ta.time[b] - last.get(b, -1.0) < 1e-6:
            return False
        last[b] = ta.time[b]
        x_, y_ = ta.state[0:2, b]
        log.append((b, "curve", float(ta.time[b]), float(x_), float(y_)))
        grad = np.array([1 + 0.05 * 11 * np.pi * np.sin(11 * np.pi * x_), 1])
        u = grad / np.linalg.norm(grad)
        vel = ta.state[2:4, b]
        vel += (1.0 + CR) * (-np.dot(vel, u) * u)
        return True

    def cb_bottom(ta, d_sgn, b):
        if ta.time[b] - last.get(b, -1.0) < 1e-6:
            return False
        last[b] = ta.time[b]
        log.append((b, "ground", float(ta.time[b]), float(ta.state[0, b]), float(ta.state[1, b])))
        ta.state[3, b] = -CR * ta.state[3, b]
        return True

    B = 4
    ic = [[0.0] * B, [1.2 - i / 50 for i in range(B)], [0.0] * B, [0.0] * B]
    tes = [hy.t_event_batch(eq_curve, callback=cb_curve, direction=hy.event_direction.negative),
           hy.t_event_batch(py, callback=cb_bottom, direction=hy.event_direction.negative)]
    return eqns, ic, tes


def test_wavy_ramp_batch():
    print("[gpu] wavy ramp batch", flush=True)
    logs = ([], [])
    outs = []
    grid = np.repeat(np.linspace(0, 10, 2001), 4).reshape(2001, -1)
    for k, log in enumerate(logs):
        eqns, ic, tes = _wavy_ramp(log)
        ta = hy.taylor_adaptive_batch(eqns, ic, t_events=tes)
        if k == 0:
            _, res = ta.propagate_grid(grid)
        else:
            _, res = ta.propagate_grid(grid, callback=lambda t: True)  # host-driven steps
        outs.append((res, ta.state.copy(), list(ta.propagate_res)))
    (r0, s0, p0), (r1, s1, p1) = outs
    # Parked device path and host-driven steps take the same steps and bounces bit for bit;
    # the grid points are evaluated on the device from the step start and on the host from
    # the step end, so they differ by rounding.
    np.testing.assert_array_equal(s0, s1)
    np.testing.assert_allclose(r0, r1, rtol=0, atol=1e-13)
    assert p0 == p1 and sorted(logs[0]) == sorted(logs[1])
    curve = lambda x: 1.0 - x + 0.05 * np.cos(11 * np.pi * x)
    bounces = logs[0]
    assert sum(1 for e in bounces if e[1] == "curve") >= 8
    for b, kind, t, x_, y_ in bounces:
        # Each bounce sits on its surface.
        if kind == "curve":
            assert abs(y_ - curve(x_)) < 1e-12
        else:
            assert abs(y_) < 1e-12
    # The ball never goes through the ramp or the ground.
    fin = np.isfinite(r0[:, 0, :])
    assert np.all(r0[:, 1, :][fin] >= -1e-9)
    above = r0[:, 0, :][fin] < 0.9
    assert np.all((r0[:, 1, :][fin] - curve(r0[:, 0, :][fin]))[above] >= -1e-9)
