"""Continuous output (propagate_until/for with c_output=True) on the GPU: the
reference's batch test (heyoka/test.py:1413-1680), the notebook values
(doc/notebooks/Batch mode overview.ipynb cells 31-36) and agreement with
propagate_grid's dense output."""

import copy
import pickle
import sys

import numpy as np
import pytest

import heyoka_amd as hy

pytestmark = pytest.mark.gpu

x, v = hy.make_vars("x", "v")
PEND = [(x, v), (v, -9.8 * hy.sin(x))]


def _fdp():
    return [(x, v), (v, hy.cos(hy.time) - hy.par[0] * v - hy.sin(x))]


def test_notebook_c_output(goldens):
    print("[gpu] c_output notebook", flush=True)
    s = goldens["fdp_setup"]
    g = goldens["fdp_c_output"]
    ta = hy.taylor_adaptive_batch(_fdp(), s["state"], pars=s["pars"])
    c_out, cb = ta.propagate_until(g["t_final"], c_output=True)
    assert cb is None
    d = goldens["fdp_d_output"]
    np.testing.assert_allclose(c_out(d["t"]), d["out"], rtol=0, atol=0.5 * 10.0 ** -d["digits"] * 10)
    c_arr = c_out(g["times"])
    assert c_arr.shape == (3, 2, 4)
    assert abs(c_arr[0, 1, 2] - g["c_arr_0_1_2"]) <= 4e-16 * abs(g["c_arr_0_1_2"])
    np.testing.assert_array_equal(ta.time, g["t_final"])


def test_reference_batch_c_output():
    print("[gpu] c_output reference batch test", flush=True)
    ic = [[0.0, 0.01, 0.02, 0.03], [0.25, 0.26, 0.27, 0.28]]
    arr_ic = np.array(ic)
    ta = hy.taylor_adaptive_batch(sys=PEND, state=ic)
    final_tm = [10.0, 10.4, 10.5, 11.0]
    check_tm = [0.1, 1.3, 5.6, 9.1]

    def reset():
        ta.state[:] = ic
        ta.set_time([0.0] * 4)

    c_out, cb = ta.propagate_until(final_tm)
    assert cb is None and c_out is None
    reset()
    c_out, cb = ta.propagate_until(final_tm, c_output=False)
    assert cb is None and c_out is None
    reset()
    c_out, cb = ta.propagate_until(final_tm, c_output=True)
    assert cb is None and c_out is not None
    assert c_out(check_tm).shape == (2, 4)
    with pytest.raises(ValueError):
        c_out(check_tm)[0] = 0.5
    rc = sys.getrefcount(c_out)
    tmp_out = c_out(check_tm)
    assert sys.getrefcount(c_out) == rc + 1
    del tmp_out
    with pytest.raises(ValueError) as e:
        c_out(np.zeros((1, 1, 1)))
    assert ("Invalid time array passed to a continuous_output_batch object: the number of dimensions must be 1 or "
            "2, but it is 3 instead") in str(e.value)
    for k in (1, 0, 5):
        with pytest.raises(ValueError) as e:
            c_out(np.zeros((k,)))
        assert (f"Invalid time array passed to a continuous_output_batch object: the length must be 4 but it is {k} "
                "instead") in str(e.value)
    # Against one-element integrators (the reference compares with scalar ones).
    singles = []
    for idx in range(4):
        t1 = hy.taylor_adaptive_batch(sys=PEND, state=arr_ic[:, idx:idx + 1])
        singles.append(t1.propagate_until([final_tm[idx]], c_output=True)[0])
    c_out(check_tm)
    for idx in range(4):
        singles[idx]([check_tm[idx]])
        np.testing.assert_allclose(singles[idx].output[:, 0], c_out.output[:, idx], rtol=2.2e-15, atol=2.2e-15)
    scal = copy.deepcopy(np.array(c_out(0.42)))
    assert np.all(scal == c_out([0.42] * 4))
    nc = np.vstack([check_tm, np.zeros(4)]).T.flatten()[::2]
    c_out(nc)
    for idx in range(4):
        singles[idx]([check_tm[idx]])
        np.testing.assert_allclose(singles[idx].output[:, 0], c_out.output[:, idx], rtol=2.2e-15, atol=2.2e-15)
    with pytest.raises(ValueError) as e:
        c_out(np.zeros((5, 3)))
    assert ("Invalid time array passed to a continuous_output_batch object: the number of columns must be 4 but it "
            "is 3 instead") in str(e.value)
    b_check = np.repeat(check_tm, 5, axis=0).reshape((4, 5)).T
    out_b = c_out(b_check)
    assert out_b.shape == (5, 2, 4)
    for j in range(5):
        np.testing.assert_array_equal(out_b[j], c_out(check_tm))
    assert c_out(np.zeros((0, 4))).shape == (0, 2, 4)
    assert c_out.times.shape == (c_out.n_steps + 1, 4) and np.all(np.isfinite(c_out.times))
    with pytest.raises(ValueError):
        c_out.times[0] = 0.5
    assert c_out.tcs.shape == (c_out.n_steps, 2, ta.order + 1, 4)
    assert np.all(c_out.bounds[0] == [0.0] * 4)
    np.testing.assert_allclose(c_out.bounds[1], final_tm, rtol=2.2e-15)
    assert c_out.batch_size == 4 and "forward" in repr(c_out)
    for c2 in (copy.copy(c_out), copy.deepcopy(c_out), pickle.loads(pickle.dumps(c_out))):
        assert c2.tcs.shape == (c_out.n_steps, 2, ta.order + 1, 4)
        np.testing.assert_array_equal(c2(check_tm), c_out(check_tm))

    class foo:
        pass

    cc = copy.deepcopy(c_out)
    cc.bar = foo()
    assert id(cc.bar) == id(copy.copy(cc).bar)
    assert id(cc.bar) != id(copy.deepcopy(cc).bar)
    c_out.foo = []
    assert pickle.loads(pickle.dumps(c_out)).foo == []


def test_c_output_matches_grid_and_backward():
    print("[gpu] c_output vs grid", flush=True)
    rng = np.random.default_rng(2)
    ic = np.stack([rng.uniform(-0.5, 0.5, 6), rng.uniform(-0.5, 0.5, 6)])
    ta = hy.taylor_adaptive_batch(PEND, ic)
    c_out, _ = ta.propagate_until(np.full(6, 7.0), c_output=True)
    tb = hy.taylor_adaptive_batch(PEND, ic)
    grid = np.repeat(np.linspace(0.0, 7.0, 29), 6).reshape(29, 6)
    _, out = tb.propagate_grid(grid)
    np.testing.assert_allclose(c_out(grid), out, rtol=0, atol=1e-14)
    # Backward: bounds reversed, evaluation from the steps taken backwards.
    back, _ = ta.propagate_until(np.full(6, 0.0), c_output=True)
    assert "backward" in repr(back)
    np.testing.assert_allclose(back(np.full(6, 0.0)), ic, rtol=0, atol=1e-13)
    np.testing.assert_allclose(back(grid[::-1]), out[::-1], rtol=0, atol=1e-13)


@pytest.mark.parametrize("mode", ["persistent", "plain"])
def test_device_log_matches_host_steps(mode):
    # The device path (per-instance step logs, instances park when their 64-slot log is
    # full) against host-driven steps (forced by a step callback): same steps, same
    # coefficients, same evaluation, bit for bit.
    print(f"[gpu] c_output device log vs host steps mode={mode}", flush=True)
    rng = np.random.default_rng(9)
    B = 24
    ic = np.stack([rng.uniform(-1.0, 1.0, B), rng.uniform(-1.0, 1.0, B)])
    tf = rng.uniform(40.0, 80.0, B)
    ta = hy.taylor_adaptive_batch(PEND, ic, gpu_mode=mode)
    c1, _ = ta.propagate_until(tf, c_output=True)
    tb = hy.taylor_adaptive_batch(PEND, ic, gpu_mode=mode)
    c2, _ = tb.propagate_until(tf, c_output=True, callback=lambda t: True)
    assert c1.n_steps > 64  # several drains per instance
    np.testing.assert_array_equal(ta.state, tb.state)
    assert ta.propagate_res == tb.propagate_res
    np.testing.assert_array_equal(c1.tcs, c2.tcs)
    np.testing.assert_array_equal(c1.times, c2.times)
    tq = np.stack([rng.uniform(0.0, tf) for _ in range(7)])
    np.testing.assert_array_equal(c1(tq), c2(tq))
    # Step counts match propagate_res per element.
    ns = [r[3] for r in ta.propagate_res]
    assert max(ns) == c1.n_steps


def test_large_batch_c_output_vs_grid():
    print("[gpu] c_output large batch", flush=True)
    B = 8192
    rng = np.random.default_rng(4)
    ic = np.stack([rng.uniform(-1.0, 1.0, B), rng.uniform(-1.0, 1.0, B)])
    ta = hy.taylor_adaptive_batch(PEND, ic)
    c_out, _ = ta.propagate_until(np.full(B, 20.0), c_output=True)
    tb = hy.taylor_adaptive_batch(PEND, ic)
    grid = np.repeat(np.linspace(0.0, 20.0, 9), B).reshape(9, B)
    _, out = tb.propagate_grid(grid)
    np.testing.assert_allclose(c_out(grid), out, rtol=0, atol=1e-13)
