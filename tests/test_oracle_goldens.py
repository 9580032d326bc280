"""The CPU oracle against the values printed in the reference's executed notebooks.

These pin the oracle's order rule, step-size rule (absolute and relative
branches), Horner update, double-length time and propagate_until bookkeeping
(nsteps, min_h / max_h excluding clamped steps) to heyoka's own outputs
(tests/golden/notebook_goldens.json; each entry names its notebook file:line).
"""

import numpy as np
import pytest

import heyoka_amd as hy
import pyoracle as O


def ulp_diff(a, b):
    a, b = np.float64(a), np.float64(b)
    ia = np.frombuffer(a.tobytes(), dtype=np.int64)[0]
    ib = np.frombuffer(b.tobytes(), dtype=np.int64)[0]
    return abs(int(ia) - int(ib))


@pytest.fixture(scope="module")
def pend():
    x, v = hy.make_vars("x", "v")
    return hy.taylor_decompose([(x, v), (v, -9.8 * hy.sin(x))])


@pytest.fixture(scope="module")
def fdp():
    x, v = hy.make_vars("x", "v")
    return hy.taylor_decompose([(x, v), (v, hy.cos(hy.time) - hy.par[0] * v - hy.sin(x))])


def test_order_rule(goldens):
    assert hy.order_from_tol(np.finfo(float).eps) == 20
    assert hy.order_from_tol(1e-9) == goldens["pend_tol1e9"]["order"] == 12
    assert hy.order_from_tol(1e-15) == 19


def test_pendulum_single_steps(pend, goldens):
    g = goldens["pend_first_step"]
    r = O.step_batch(pend, 20, [[0.05], [0.025]])
    assert r["outcome"][0] == O.SUCCESS
    assert ulp_diff(r["h"][0], g["h"]) <= 2
    np.testing.assert_allclose(r["state"][:, 0], g["state"], rtol=0, atol=5e-17)
    # Backward step from there.
    r2 = O.step_batch(pend, 20, r["state"], t_hi=r["t_hi"], t_lo=r["t_lo"], max_dt=-np.inf)
    assert ulp_diff(r2["h"][0], goldens["pend_step_backward"]["h"]) <= 2
    # Clamped steps: +0.01 then -0.02, both time_limit.
    r3 = O.step_batch(pend, 20, r2["state"], t_hi=r2["t_hi"], t_lo=r2["t_lo"], max_dt=0.01)
    assert r3["outcome"][0] == O.TIME_LIMIT and r3["h"][0] == 0.01
    r4 = O.step_batch(pend, 20, r3["state"], t_hi=r3["t_hi"], t_lo=r3["t_lo"], max_dt=-0.02)
    assert r4["outcome"][0] == O.TIME_LIMIT and r4["h"][0] == -0.02
    g = goldens["pend_after_clamped"]
    # The first step differs from the golden by at most 2 ulp (of 0.216), which the
    # summed time carries along.
    assert abs(r4["t_hi"][0] - g["time"]) <= 6e-17
    np.testing.assert_allclose(r4["state"][:, 0], g["state"], rtol=0, atol=5e-9)


def test_pendulum_propagate(pend, goldens):
    st = np.array([[0.05], [0.025]])
    r = O.propagate_until(pend, 20, st, 5.0)
    for key, (rr) in (("pend_propagate_for_5", r),):
        g = goldens[key]
        assert rr["outcome"][0] == O.TIME_LIMIT
        assert int(rr["nsteps"][0]) == g["nsteps"]
        assert ulp_diff(rr["min_h"][0], g["min_h"]) <= 4
        assert ulp_diff(rr["max_h"][0], g["max_h"]) <= 4
        assert rr["t_hi"][0] == g["time"]
    r2 = O.propagate_until(pend, 20, r["state"], 20.0, t_hi=r["t_hi"], t_lo=r["t_lo"])
    g = goldens["pend_propagate_until_20"]
    assert int(r2["nsteps"][0]) == g["nsteps"]
    assert ulp_diff(r2["min_h"][0], g["min_h"]) <= 4 and ulp_diff(r2["max_h"][0], g["max_h"]) <= 4
    assert r2["t_hi"][0] == 20.0
    r3 = O.propagate_until(pend, 20, r2["state"], 0.0, t_hi=r2["t_hi"], t_lo=r2["t_lo"])
    g = goldens["pend_propagate_until_0"]
    assert int(r3["nsteps"][0]) == g["nsteps"]
    assert ulp_diff(r3["min_h"][0], g["min_h"]) <= 4 and ulp_diff(r3["max_h"][0], g["max_h"]) <= 4
    assert r3["t_hi"][0] == 0.0
    # Round trip back to t=0 reproduces the printed 17-digit state to a few ulp.
    np.testing.assert_allclose(r3["state"][:, 0], g["state"], rtol=0, atol=2e-16)


def test_pendulum_clamped_propagation_times(pend, goldens):
    # propagate_until(0.5, max_delta_t=0.1): every step is clamped; the times the
    # step callback printed follow from double-length time accumulation.
    st = np.array([[0.05], [0.025]])
    t_hi, t_lo = np.zeros(1), np.zeros(1)
    times = []
    for _ in range(5):
        rem = (0.5 - t_hi[0]) - t_lo[0]
        r = O.step_batch(pend, 20, st, t_hi=t_hi, t_lo=t_lo, max_dt=min(0.1, rem))
        st, t_hi, t_lo = r["state"], r["t_hi"], r["t_lo"]
        times.append(t_hi[0])
    assert times == goldens["pend_callback_times"]["times"]


def test_pendulum_tol_1e9_round_trip(goldens):
    x, v = hy.make_vars("x", "v")
    dec = hy.taylor_decompose([(x, v), (v, -9.8 * hy.sin(x))])
    p = hy.order_from_tol(1e-9)
    r = O.propagate_until(dec, p, [[0.05], [0.025]], 10.0)
    r = O.propagate_until(dec, p, r["state"], 0.0, t_hi=r["t_hi"], t_lo=r["t_lo"])
    # The round-trip error (~1e-12) is a property of the order-12 truncation; it
    # must be reproduced, not just bounded.
    np.testing.assert_allclose(r["state"][:, 0], goldens["pend_tol1e9"]["state"], rtol=0, atol=3e-15)


def test_fdp_batch_step(fdp, goldens):
    s = goldens["fdp_setup"]
    r = O.step_batch(fdp, 20, s["state"], pars=s["pars"])
    assert np.all(r["outcome"] == goldens["fdp_step"]["outcome"])
    assert r["h"].tolist() == goldens["fdp_step"]["h"]  # bit-exact in the relative-error branch
    g = goldens["fdp_step_state"]
    np.testing.assert_allclose(r["state"], g["state"], rtol=0, atol=5e-9)
    np.testing.assert_allclose(r["t_hi"], g["time"], rtol=0, atol=5e-9)
    r2 = O.step_batch(fdp, 20, r["state"], pars=s["pars"], t_hi=r["t_hi"], t_lo=r["t_lo"],
                      max_dt=[0.010, 0.011, 0.012, 0.013])
    g = goldens["fdp_clamped"]
    assert np.all(r2["outcome"] == g["outcome"]) and r2["h"].tolist() == g["h"]
    np.testing.assert_allclose(r2["state"], g["state"], rtol=0, atol=5e-9)
    np.testing.assert_allclose(r2["t_hi"], g["time"], rtol=0, atol=5e-9)


def test_fdp_batch_propagate(fdp, goldens):
    s = goldens["fdp_setup"]
    r = O.step_batch(fdp, 20, s["state"], pars=s["pars"])
    r = O.step_batch(fdp, 20, r["state"], pars=s["pars"], t_hi=r["t_hi"], t_lo=r["t_lo"],
                     max_dt=[0.010, 0.011, 0.012, 0.013])
    # propagate_for(delta_t): final time = current time + delta_t in double length.
    dt = np.array(goldens["fdp_propagate_for"]["delta_t"])
    tf_hi = r["t_hi"] + dt
    tf_lo = (dt - (tf_hi - r["t_hi"])) + (r["t_hi"] - (tf_hi - (tf_hi - r["t_hi"]))) + r["t_lo"]
    p = O.propagate_until(fdp, 20, r["state"], tf_hi, tf_lo=tf_lo, pars=s["pars"], t_hi=r["t_hi"], t_lo=r["t_lo"],
                          lane_batch=4)
    g = goldens["fdp_propagate_for"]
    for b in range(4):
        oc, mn, mx, ns = g["res"][b]
        assert p["outcome"][b] == oc and int(p["nsteps"][b]) == ns
        # ~40 steps of ulp-level differences in the state move h by a few ulp.
        np.testing.assert_allclose([p["min_h"][b], p["max_h"][b]], [mn, mx], rtol=1e-14, atol=0)
    np.testing.assert_allclose(p["state"], g["state"], rtol=0, atol=5e-8)
    np.testing.assert_allclose(p["t_hi"], g["time"], rtol=0, atol=5e-8)
    q = O.propagate_until(fdp, 20, p["state"], goldens["fdp_propagate_until"]["t"], pars=s["pars"], t_hi=p["t_hi"],
                          t_lo=p["t_lo"], lane_batch=4)
    g = goldens["fdp_propagate_until"]
    for b in range(4):
        oc, mn, mx, ns = g["res"][b]
        assert q["outcome"][b] == oc and int(q["nsteps"][b]) == ns
        np.testing.assert_allclose([q["min_h"][b], q["max_h"][b]], [mn, mx], rtol=1e-13, atol=0)
    np.testing.assert_allclose(q["state"], g["state"], rtol=0, atol=5e-8)
    assert q["t_hi"].tolist() == g["t"]


def test_fdp_taylor_coefficients(fdp, goldens):
    s = goldens["fdp_setup"]
    r = O.step_batch(fdp, 20, s["state"], pars=s["pars"], want_tc=True)
    tc = np.array(goldens["fdp_tc"]["tc"])  # [2][21][4], printed with 9 significant digits
    np.testing.assert_allclose(r["tc"], tc, rtol=6e-9, atol=1e-15)


def test_fdp_dense_output_from_tc(fdp, goldens):
    # update_d_output(t) after one step is the Taylor polynomial of the step
    # evaluated at t - t0 (here t0 = 0).
    s = goldens["fdp_setup"]
    r = O.step_batch(fdp, 20, s["state"], pars=s["pars"], want_tc=True)
    g = goldens["fdp_d_output"]
    t = np.array(g["t"])
    out = np.zeros((2, 4))
    for i in range(2):
        acc = r["tc"][i, 20, :].copy()
        for k in range(19, -1, -1):
            acc = acc * t + r["tc"][i, k, :]
        out[i] = acc
    np.testing.assert_allclose(out, g["out"], rtol=0, atol=5e-9)


def test_variational_order2_goldens(goldens):
    # var_ode_sys(FDP, var_args.vars, order=2) from [0.2, 0.3], par 0.4, to t = 3
    # (doc/notebooks/var_ode_sys.ipynb:299,361,529): pins the variational
    # formulation (layout + equations) together with the propagate loop.
    x, v = hy.make_vars("x", "v")
    vs = hy.var_ode_sys([(x, v), (v, hy.cos(hy.time) - hy.par[0] * v - hy.sin(x))], hy.var_args.vars, order=2)
    dec = hy.taylor_decompose(vs.sys)
    g = goldens["var2_state_t3"]
    st = np.array(g["ic"] + vs.initial_derivatives())[:, None]
    r = O.propagate_until(dec, 20, st, 3.0, pars=[[0.4]])
    np.testing.assert_allclose(r["state"][:, 0], g["state"], rtol=0, atol=5e-9)
    sl = vs.get_vslice(2)
    assert (sl.start, sl.stop) == (6, 12)
    assert [vs.get_mindex(i) for i in range(sl.start, sl.stop)] == goldens["var2_order2"]["mindex"]
    np.testing.assert_allclose(r["state"][sl, 0], goldens["var2_order2"]["values"], rtol=1e-13, atol=0)
    # The order-1 + order-2 Taylor map reproduces the displaced non-variational run
    # (var_ode_sys.ipynb:650-651, 707: error ~6.6e-13).
    dx = np.array([goldens["var2_displaced"]["dx"], goldens["var2_displaced"]["dv"]])
    sv = r["state"][:, 0]
    approx = sv[:2] + sv[2:6].reshape(2, 2) @ dx
    o2 = sv[6:12].reshape(2, 3)
    approx = approx + o2 @ np.array([dx[0] ** 2 / 2, dx[0] * dx[1], dx[1] ** 2 / 2])
    np.testing.assert_allclose(approx, goldens["var2_displaced"]["nonvar_state"], rtol=0, atol=5e-9)


def test_pendulum_propagate_grid(pend, goldens):
    # ta.propagate_grid([0.0, 0.1, ..., 1.0]) from the notebook's initial state:
    # 5 steps, the last clamped to the final grid time, and the dense output.
    g = goldens["pend_grid"]
    grid = np.array(g["grid"])[:, None]
    r = O.propagate_grid(pend, 20, np.array([[0.05], [0.025]]), grid, lane_batch=1)
    assert r["outcome"][0] == O.TIME_LIMIT and int(r["nsteps"][0]) == g["nsteps"]
    assert ulp_diff(r["min_h"][0], g["min_h"]) <= 4 and ulp_diff(r["max_h"][0], g["max_h"]) <= 4
    np.testing.assert_allclose(r["out"][:, :, 0], g["out"], rtol=0, atol=5e-9)
    assert r["t_hi"][0] == 1.0


def test_fdp_batch_propagate_grid(fdp, goldens):
    # Batch mode overview: grid linspace(0, 100, 1000) on the 4 lanes from the
    # initial conditions; the reference prints propagate_res afterwards.
    s = goldens["fdp_setup"]
    grid = np.repeat(np.linspace(0, 100, 1000), 4).reshape(1000, -1)
    r = O.propagate_grid(fdp, 20, s["state"], grid, pars=s["pars"], lane_batch=4)
    for b, (oc, mn, mx, ns) in enumerate(goldens["fdp_grid_res"]["res"]):
        assert r["outcome"][b] == oc and int(r["nsteps"][b]) == ns
        # ~400 steps of a forced, weakly chaotic system: round-off differences in the
        # state grow to ~1e-13 in h (the step counts agree exactly).
        np.testing.assert_allclose([r["min_h"][b], r["max_h"][b]], [mn, mx], rtol=5e-13)
    assert np.all(np.isfinite(r["out"])) and np.all(r["t_hi"] == 100.0)


def test_grid_dense_output_matches_propagate_until(pend):
    # The dense output at an interior grid time agrees with propagating to that
    # time directly (different steps, same trajectory) to near round-off, and a
    # one-point grid returns the current state with no step.
    x0 = np.array([[0.05, -0.3, 0.7], [0.025, 0.4, -0.2]])
    grid = np.repeat(np.linspace(0.0, 3.0, 31), 3).reshape(31, 3)
    r = O.propagate_grid(pend, 20, x0, grid, lane_batch=3)
    for j in (1, 7, 18, 30):
        p = O.propagate_until(pend, 20, x0, grid[j], lane_batch=3)
        np.testing.assert_allclose(r["out"][j], p["state"], rtol=0, atol=2e-14)
    np.testing.assert_array_equal(r["out"][-1], r["state"])
    one = O.propagate_grid(pend, 20, x0, grid[:1], lane_batch=3)
    np.testing.assert_array_equal(one["out"][0], x0)
    assert np.all(one["nsteps"] == 0) and np.all(one["outcome"] == O.TIME_LIMIT)


def test_grid_backward_and_step_limit(pend):
    # Backward grids walk the same rule in the other direction; a step limit
    # leaves the unreached points NaN (batch propagate_grid fills them so).
    x0 = np.array([[0.05, 0.3], [0.025, -0.1]])
    grid = np.repeat(np.linspace(0.0, -2.0, 9), 2).reshape(9, 2)
    r = O.propagate_grid(pend, 20, x0, grid, lane_batch=2)
    p = O.propagate_until(pend, 20, x0, grid[4], lane_batch=2)
    np.testing.assert_allclose(r["out"][4], p["state"], rtol=0, atol=2e-14)
    grid = np.repeat(np.linspace(0.0, 20.0, 41), 2).reshape(41, 2)
    r = O.propagate_grid(pend, 20, x0, grid, max_steps=3, lane_batch=2)
    assert np.all(r["outcome"] == O.STEP_LIMIT) and np.all(r["nsteps"] == 3)
    reached = np.isfinite(r["out"][:, 0, 0])
    k = int(reached.sum())
    assert 1 < k < 41 and reached[:k].all() and not reached[k:].any()


def test_high_accuracy_update_is_compensated(fdp, goldens):
    # high_accuracy replaces Horner by the compensated sum of x^[k] h^k (Customising the
    # adaptive integrator.ipynb, cell 12). With the step's own coefficients and h, the
    # result must be within 2 ulp of the sum taken in extended precision (h^k formed by
    # repeated multiplication, as the update does).
    s = goldens["fdp_setup"]
    rng = np.random.default_rng(11)
    st = np.vstack([rng.uniform(-1, 1, 64), rng.uniform(-2, 2, 64)])
    pa = rng.uniform(0.05, 0.2, (1, 64))
    r = O.step_batch(fdp, 20, st, pars=pa, high_accuracy=True, want_tc=True)
    tc = r["tc"].astype(np.longdouble)
    h = r["h"].astype(np.longdouble)
    exact = np.zeros((2, 64), dtype=np.longdouble)
    hk = np.ones(64, dtype=np.longdouble)
    for k in range(21):
        exact += tc[:, k, :] * hk
        hk = hk * h
    ref = exact.astype(np.float64)
    u = [[ulp_diff(a, b) for a, b in zip(ra, rb)] for ra, rb in zip(r["state"], ref)]
    assert max(max(x) for x in u) <= 2
    p = O.step_batch(fdp, 20, st, pars=pa, want_tc=True)
    assert np.array_equal(p["h"], r["h"])  # same coefficients and step, different summation
    assert np.max(np.abs(p["state"] - r["state"])) < 1e-14
