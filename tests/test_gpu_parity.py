"""GPU kernels (through libhyx.so) against the notebook goldens and the CPU oracle.

Tolerances: the kernels use fused multiply-adds and reciprocal multiplies where
the oracle rounds every operation, so results agree to a few ulp per step;
final states of non-chaotic runs must agree within 1e-10 relative (1e-12 where
stated), step counts exactly or within one step.
"""

import numpy as np
import pytest

import heyoka_amd as hy
import pyoracle as O
from heyoka_amd import workloads as W

pytestmark = pytest.mark.gpu


def ulp_diff(a, b):
    ia = np.frombuffer(np.float64(a).tobytes(), dtype=np.int64)[0]
    ib = np.frombuffer(np.float64(b).tobytes(), dtype=np.int64)[0]
    return abs(int(ia) - int(ib))


def _pend_sys():
    x, v = hy.make_vars("x", "v")
    return [(x, v), (v, -9.8 * hy.sin(x))]


def _fdp_sys():
    x, v = hy.make_vars("x", "v")
    return [(x, v), (v, hy.cos(hy.time) - hy.par[0] * v - hy.sin(x))]


@pytest.mark.parametrize("code_form", ["unrolled", "compact"])
def test_pendulum_goldens(goldens, code_form):
    print(f"[gpu] pendulum goldens code_form={code_form}", flush=True)
    ta = hy.taylor_adaptive_batch(_pend_sys(), [[0.05], [0.025]], code_form=code_form)
    assert ta.order == 20
    ta.step()
    assert ta.step_res[0][0] == hy.taylor_outcome.success
    assert ulp_diff(ta.step_res[0][1], goldens["pend_first_step"]["h"]) <= 3
    np.testing.assert_allclose(ta.state[:, 0], goldens["pend_first_step"]["state"], rtol=0, atol=1e-16)
    ta.step_backward()
    assert ulp_diff(ta.step_res[0][1], goldens["pend_step_backward"]["h"]) <= 3
    ta.step(max_delta_t=[0.01])
    assert ta.step_res[0] == (hy.taylor_outcome.time_limit, 0.01)
    ta.step(max_delta_t=[-0.02])
    assert ta.step_res[0] == (hy.taylor_outcome.time_limit, -0.02)
    ta.set_time(0.0)
    ta.state[:] = [[0.05], [0.025]]
    for key, call in (("pend_propagate_for_5", lambda: ta.propagate_for(5.0)),
                      ("pend_propagate_until_20", lambda: ta.propagate_until(20.0)),
                      ("pend_propagate_until_0", lambda: ta.propagate_until(0.0))):
        call()
        g = goldens[key]
        oc, mn, mx, ns = ta.propagate_res[0]
        assert oc == hy.taylor_outcome.time_limit
        assert ns == g["nsteps"], key
        np.testing.assert_allclose([mn, mx], [g["min_h"], g["max_h"]], rtol=1e-14)
        assert ta.time[0] == g["time"]
    np.testing.assert_allclose(ta.state[:, 0], goldens["pend_propagate_until_0"]["state"], rtol=0, atol=5e-16)


def test_fdp_batch_goldens(goldens):
    print("[gpu] fdp batch goldens", flush=True)
    s = goldens["fdp_setup"]
    ta = hy.taylor_adaptive_batch(_fdp_sys(), s["state"], pars=s["pars"])
    ta.step(write_tc=True)
    hs = [h for _, h in ta.step_res]
    for a, b in zip(hs, goldens["fdp_step"]["h"]):
        assert ulp_diff(a, b) <= 3
    np.testing.assert_allclose(ta.tc, np.array(goldens["fdp_tc"]["tc"]), rtol=6e-9, atol=1e-15)
    np.testing.assert_allclose(ta.update_d_output([0.01, 0.02, 0.03, 0.04]), goldens["fdp_d_output"]["out"], rtol=0,
                               atol=5e-9)
    np.testing.assert_allclose(ta.update_d_output(ta.time), ta.state, rtol=1e-15, atol=1e-15)
    ta.step([0.010, 0.011, 0.012, 0.013])
    assert [r for r in ta.step_res] == [(hy.taylor_outcome.time_limit, h) for h in goldens["fdp_clamped"]["h"]]
    np.testing.assert_allclose(ta.state, goldens["fdp_clamped"]["state"], rtol=0, atol=5e-9)
    ta.propagate_for([10.0, 11.0, 12.0, 13.0])
    g = goldens["fdp_propagate_for"]
    for b in range(4):
        oc, mn, mx, ns = ta.propagate_res[b]
        assert int(oc) == g["res"][b][0] and ns == g["res"][b][3]
        np.testing.assert_allclose([mn, mx], g["res"][b][1:3], rtol=1e-13)
    np.testing.assert_allclose(ta.state, g["state"], rtol=0, atol=5e-8)
    ta.propagate_until([20.0, 21.0, 22.0, 23.0])
    g = goldens["fdp_propagate_until"]
    for b in range(4):
        oc, mn, mx, ns = ta.propagate_res[b]
        assert int(oc) == g["res"][b][0] and ns == g["res"][b][3]
    np.testing.assert_allclose(ta.state, g["state"], rtol=0, atol=5e-8)
    assert ta.time.tolist() == g["t"]


def _gpu_vs_oracle(wl, n, tf, rtol, compare_steps=True, code_form=None, mode="persistent", min_frac=0.8):
    """Final states of the GPU vs the f64 oracle on the instances whose f64 and
    real128 oracle runs agree to 1e-12 (the well-conditioned ones), plus the
    GPU's error against the real128 oracle over all instances."""
    st = wl.state[:, :n]
    pars = None if wl.pars is None else wl.pars[:, :n]
    ta = hy.taylor_adaptive_batch(wl.sys, st, pars=pars, tol=wl.tol, code_form=code_form, gpu_mode=mode)
    tfv = np.broadcast_to(np.asarray(tf, dtype=float), (n,)).copy()
    ta.propagate_until(tfv)
    dec = hy.taylor_decompose(wl.sys.sys if hasattr(wl.sys, "n_orig_sv") else wl.sys)
    r = O.propagate_until(dec, ta.order, st, tfv, pars=pars, nthreads=8, lane_batch=1)
    q = O.propagate_until(dec, 2 * ta.order, st, tfv, pars=pars, nthreads=8, lane_batch=1, f128=True)
    qs = q["state"] + q["state_lo"]
    scale = np.maximum(np.abs(qs), 1.0)
    e64 = np.max(np.abs(r["state"] - qs) / scale, axis=0)
    egpu = np.max(np.abs(ta.state - qs) / scale, axis=0)
    ok = e64 < 1e-12
    assert ok.sum() >= min_frac * n, ok.sum()
    err = np.max(np.abs(ta.state - r["state"])[:, ok] / np.maximum(np.abs(r["state"][:, ok]), 1.0))
    assert err <= rtol, err
    # Over every instance the GPU is as accurate as the f64 oracle against real128.
    assert np.median(egpu) <= 10 * np.median(e64) + 1e-15, (np.median(egpu), np.median(e64))
    ns = np.array([p[3] for p in ta.propagate_res]).astype(np.int64)
    if compare_steps:
        assert np.max(np.abs(ns[ok] - r["nsteps"].astype(np.int64)[ok])) <= 1
    assert all(p[0] == hy.taylor_outcome.time_limit for p in ta.propagate_res)
    print(f"[gpu] {wl.name}: n={n} well-conditioned={ok.sum()} max|gpu-f64|={err:.2e} "
          f"median err vs real128: gpu {np.median(egpu):.2e} f64 {np.median(e64):.2e}", flush=True)
    return ta, r, err


def test_pendulum_config_vs_oracle_and_energy():
    print("[gpu] pendulum config B=4 t=100", flush=True)
    wl = W.pendulum()
    ta, r, err = _gpu_vs_oracle(wl, 4, 100.0, 1e-12)
    en = lambda s: 0.5 * s[1] ** 2 + 1.0 - np.cos(s[0])
    e0 = en(wl.state)
    drift = np.max(np.abs(en(ta.state) - e0) / np.abs(e0))
    assert drift < 1e-13, drift


@pytest.mark.parametrize("code_form", ["unrolled", "compact"])
def test_cr3bp_vs_oracle(code_form):
    print(f"[gpu] cr3bp 256 instances t=20 code_form={code_form}", flush=True)
    _gpu_vs_oracle(W.cr3bp(256), 256, 20.0, 1e-10, code_form=code_form)


@pytest.mark.parametrize("code_form", ["group", "compact"])
def test_nbody6_short_vs_oracle(code_form):
    print(f"[gpu] nbody6 128 instances t=2 code_form={code_form}", flush=True)
    ta, _, _ = _gpu_vs_oracle(W.nbody6(128), 128, 2.0, 1e-10, code_form=code_form)
    assert ta.code_form == code_form


def _nbody3():
    sys_ = hy.model.nbody(3, masses=[1.0, 0.3, 0.01])
    base = np.array([0, 0, 0, 0, 0, 0, 1, 0, 0.1, 0, 1.1, 0, -2, 0.5, 0, 0.2, -0.6, 0.05], dtype=float)
    rng = np.random.default_rng(7)
    st = base[:, None] + rng.normal(0, 1e-3, (18, 37))
    return W.Workload("nbody3", sys_, st, None, np.full(37, 3.0), 1e-15)


@pytest.mark.parametrize("mode", ["persistent", "plain", "graph"])
def test_group_form_nbody3_modes(mode):
    # G = 4 lanes per instance, 16 instances per wave, a batch that is not a multiple of 16.
    print(f"[gpu] nbody3 group form mode={mode}", flush=True)
    ta, _, _ = _gpu_vs_oracle(_nbody3(), 37, 3.0, 1e-10, code_form="group", mode=mode)
    assert ta.code_form == "group"


def test_group_form_step_tc_and_lane_failure():
    print("[gpu] group form step(write_tc) and a non-finite instance", flush=True)
    wl = _nbody3()
    st = wl.state[:, :5].copy()
    ta = hy.taylor_adaptive_batch(wl.sys, st, tol=1e-15, code_form="group")
    ta.step(write_tc=True)
    dec = hy.taylor_decompose(wl.sys)
    r = O.step_batch(dec, ta.order, st, want_tc=True)
    np.testing.assert_allclose(ta.tc, r["tc"], rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose([h for _, h in ta.step_res], r["h"], rtol=1e-13)
    st[3, 2] = np.inf
    tb = hy.taylor_adaptive_batch(wl.sys, st, tol=1e-15, code_form="group")
    tb.propagate_until(1.0)
    oc = [p[0] for p in tb.propagate_res]
    assert oc[2] == hy.taylor_outcome.err_nf_state
    assert all(o == hy.taylor_outcome.time_limit for i, o in enumerate(oc) if i != 2)


@pytest.mark.parametrize("code_form", [None, "compact"])
def test_j2var_short_vs_oracle(code_form):
    # Automatic form: the wave form (32 lanes per instance); the compact form's tape in device memory.
    print(f"[gpu] j2var 64 instances t=5 code_form={code_form}", flush=True)
    ta, _, _ = _gpu_vs_oracle(W.j2var(64), 64, 5.0, 1e-10, code_form=code_form)
    assert ta.code_form == (code_form or "wave")


def test_modes_bitwise_identical():
    print("[gpu] persistent / plain / graph bitwise", flush=True)
    wl = W.cr3bp(300, log_t=(0.5, 1.5))
    outs = []
    for mode in ("persistent", "plain", "graph"):
        ta = hy.taylor_adaptive_batch(wl.sys, wl.state, pars=wl.pars, tol=wl.tol, gpu_mode=mode)
        ta.propagate_until(wl.t_final)
        outs.append((ta.state.copy(), ta.time.copy(), ta.propagate_res))
    for o in outs[1:]:
        assert np.array_equal(o[0], outs[0][0]) and np.array_equal(o[1], outs[0][1]) and o[2] == outs[0][2]


def test_vector_vs_scalar_arguments_bitwise():
    # _test_batch_integrator.py:144-168, 250-274: scalar and per-lane arguments agree bit for bit.
    print("[gpu] vector vs scalar args", flush=True)
    ic = np.array([[0.0, 0.1, 0.2, 0.3], [0.25, 0.26, 0.27, 0.28]])
    ta = hy.taylor_adaptive_batch(_pend_sys(), ic)
    ta.propagate_until([10.0] * 4, max_delta_t=[1e-1] * 4)
    st, res = ta.state.copy(), ta.propagate_res
    ta.set_time(0.0)
    ta.state[:] = ic
    ta.propagate_until(10.0, max_delta_t=1e-1)
    assert np.array_equal(ta.state, st) and ta.propagate_res == res


def test_lane_independence_and_outcomes():
    print("[gpu] outcomes: err_nf_state lane, step_limit", flush=True)
    ic = np.array([[0.1, np.nan, 0.3, 0.4], [0.0, 0.0, 0.0, 0.0]])
    ta = hy.taylor_adaptive_batch(_pend_sys(), ic)
    ta.propagate_until(10.0)
    oc = [p[0] for p in ta.propagate_res]
    assert oc[1] == hy.taylor_outcome.err_nf_state
    assert all(o == hy.taylor_outcome.time_limit for i, o in enumerate(oc) if i != 1)
    ref = hy.taylor_adaptive_batch(_pend_sys(), ic[:, [0, 2, 3]])
    ref.propagate_until(10.0)
    assert np.array_equal(ta.state[:, [0, 2, 3]], ref.state)
    ta = hy.taylor_adaptive_batch(_pend_sys(), ic[:, [0, 2]])
    ta.propagate_until(100.0, max_steps=5)
    assert all(p[0] == hy.taylor_outcome.step_limit and p[3] == 5 for p in ta.propagate_res)


def test_resume_after_budget_is_bitwise():
    # A propagation split by the persistent kernel's iteration budget, and one
    # split into two propagate_until calls, match the straight run bit for bit.
    print("[gpu] split propagation bitwise", flush=True)
    ic = np.array([[0.0, 0.1, 0.2, 0.3], [0.25, 0.26, 0.27, 0.28]])
    a = hy.taylor_adaptive_batch(_pend_sys(), ic)
    a.propagate_until(50.0)
    b = hy.taylor_adaptive_batch(_pend_sys(), ic)
    b.propagate_until(20.0)
    b.propagate_until(50.0)
    assert np.allclose(a.state, b.state, rtol=1e-12)
    c = hy.taylor_adaptive_batch(_pend_sys(), ic)
    c._hyx_batch.set_iter_budget(7)
    c.propagate_until(50.0)
    assert c.last_kernel_timing()[1] > 10  # relaunched many times
    assert np.array_equal(a.state, c.state) and c.propagate_res == a.propagate_res


@pytest.mark.parametrize("code_form", ["unrolled", "wave"])
def test_budget_relaunch_reaches_instances_past_the_first_claims(code_form):
    # After the first launch of a budget-limited persistent propagation, only the
    # instances at the end of the queue are unfinished, beyond every position the
    # resident waves claim first: the relaunches must still carry them to their final
    # time (every outcome time_limit at t = tf, states equal to the plain form's).
    print(f"[gpu] budget relaunch past the first claim round, code_form={code_form}", flush=True)
    N = 200_000 if code_form == "unrolled" else 40_000
    rng = np.random.default_rng(2)
    ic = np.stack([rng.uniform(-1.0, 1.0, N), rng.uniform(-0.5, 0.5, N)])
    tf = np.full(N, 0.5)
    tf[-300:] = 40.0
    runs = []
    for mode in ("persistent", "plain"):
        ta = hy.taylor_adaptive_batch(_pend_sys(), ic, code_form=code_form, gpu_mode=mode)
        if mode == "persistent":
            ta._hyx_batch.set_iter_budget(64)
        ta.propagate_until(tf)
        runs.append(ta)
    a, b = runs
    assert np.array_equal(a.time, tf)
    assert all(p[0] == hy.taylor_outcome.time_limit for p in a.propagate_res)
    assert np.array_equal(a.state, b.state) and a.propagate_res == b.propagate_res


def test_copy_pickle_continuation_bitwise():
    # _test_batch_integrator.py:602-648: a pickled copy continues bit for bit.
    import pickle
    from copy import deepcopy

    print("[gpu] pickle continuation", flush=True)
    ta = hy.taylor_adaptive_batch(_pend_sys(), [[0, 0.01], [0.25, 0.26]])
    for _ in range(4):
        ta.step()
    ta2 = pickle.loads(pickle.dumps(ta))
    ta3 = deepcopy(ta)
    assert np.array_equal(ta.state, ta2.state) and np.array_equal(ta.time, ta2.time)
    ta.step()
    ta2.step()
    ta3.step()
    assert np.array_equal(ta.state, ta2.state) and np.array_equal(ta.state, ta3.state)
    assert np.array_equal(ta.time, ta2.time)


def test_callback_path_matches(goldens):
    print("[gpu] step callback path", flush=True)
    ta = hy.taylor_adaptive_batch(_pend_sys(), [[0.05], [0.025]])
    times = []

    def cb(t):
        times.append(t.time[0])
        return True

    ta.propagate_until(0.5, max_delta_t=0.1, callback=cb)
    assert times == goldens["pend_callback_times"]["times"]
