"""The benchmark configurations at (reduced) scale against the oracle on the GPU.

- Serving: a seeded Poisson stream over the pendulum, the CR3BP, the 6-body and the
  J2 + variational systems, sizes floor(10**U(0, log10 4096)), horizons 10**U(0, 1):
  every result alone == in the mix == propagate_until, bit for bit, and a sample of
  each request against the f64 oracle (1e-10 relative).
- Ensemble (heyoka/_ensemble_impl.py:50-66): the 6-body workload as 16 members of a
  4,096-instance template through ensemble_propagate_until_batch, the members' first
  instances against the oracle, two members against their own propagate_until
  bitwise, and the vectorised generator (ensemble_arrays) bitwise equal to the
  Python generator.
- J2 + variational over 200 orbital periods on 64 instances: the state transition
  matrix against the Richardson extrapolation of central finite differences
  (step 1e-6 max(1, |x_j|) and half of it), symplecticity, energy and h_z.
- The ragged CR3BP with the generator's own per-instance horizons clipped to 50,
  persistent and plain forms, against the oracle on the well-conditioned instances (where
  the f64 and real128 oracles agree) and the Jacobi constant on all.
- Snapshot / restore: slots propagated part way, snapshotted, restored into other
  slots of a new batch and finished equal the oracle's straight run (1e-10, on the
  well-conditioned instances) and the GPU's straight run (bitwise), as pickling does in the reference
  (heyoka/_test_batch_integrator.py:631-648).
- Terminal events whose callback lets the integration go on do not count towards
  min_h / max_h (the steps they shorten are not the step rule's), on the device
  parked path, the host-stepped path and the events oracle alike.
- A decomposition in heyoka's shape (Decomposition.from_heyoka) runs on the GPU.
- A tuned launch shape recorded for a system is applied to every new batch of it and
  does not change results.
"""

import json

import numpy as np
import pytest

import heyoka_amd as hy
from heyoka_amd import _lib
from heyoka_amd import workloads as W
from heyoka_amd.decompose import Decomposition
from heyoka_amd.integrator import get_system
from heyoka_amd.serving import Server

import events as OE
import pyoracle as O

pytestmark = pytest.mark.gpu


def _pairs(sys_):
    return sys_.sys if hasattr(sys_, "n_orig_sv") else sys_


def _oracle_check(wl_sys, state, pars, t_final, got, rtol=1e-10, order=19):
    dec = hy.taylor_decompose(_pairs(wl_sys))
    r = O.propagate_until(dec, order, state, t_final, pars=pars, nthreads=8, lane_batch=1)
    err = np.max(np.abs(got - r["state"]) / np.maximum(np.abs(r["state"]), 1.0))
    assert err <= rtol, err
    return err


def _conditioned_check(wl_sys, state, pars, t_final, got, min_frac, rtol=1e-10, order=19):
    """GPU against the f64 oracle on the instances whose f64 and real128 oracle runs agree
    to 1e-12 (chaotic ones amplify every rounding difference), and the GPU as accurate as
    the f64 oracle against real128 over all instances. Returns the well-conditioned mask."""
    dec = hy.taylor_decompose(_pairs(wl_sys))
    n = state.shape[1]
    r = O.propagate_until(dec, order, state, t_final, pars=pars, nthreads=16, lane_batch=1)
    q = O.propagate_until(dec, 2 * order, state, t_final, pars=pars, nthreads=16, lane_batch=1, f128=True)
    qs = q["state"] + q["state_lo"]
    scale = np.maximum(np.abs(qs), 1.0)
    e64 = np.max(np.abs(r["state"] - qs) / scale, axis=0)
    egpu = np.max(np.abs(got - qs) / scale, axis=0)
    ok = e64 < 1e-12
    assert ok.sum() >= min_frac * n, ok.sum()
    err = np.max(np.abs(got - r["state"])[:, ok] / np.maximum(np.abs(r["state"][:, ok]), 1.0))
    print(f"[gpu] {ok.sum()}/{n} well-conditioned, max|gpu - f64 oracle| {err:.2e}, median err vs real128 "
          f"gpu {np.median(egpu):.2e} f64 {np.median(e64):.2e}", flush=True)
    assert err <= rtol, err
    assert np.median(egpu) <= 10 * np.median(e64) + 1e-15
    return ok, r


def _jacobi(st, mu):
    x, y, vx, vy = st
    r1 = np.sqrt((x + mu) ** 2 + y ** 2)
    r2 = np.sqrt((x - 1 + mu) ** 2 + y ** 2)
    return x ** 2 + y ** 2 + 2 * (1 - mu) / r1 + 2 * mu / r2 - (vx ** 2 + vy ** 2)


# ---- serving config -----------------------------------------------------------------------
def _serving_requests(n_req=12, seed=21):
    rng = np.random.default_rng(seed)
    kinds = ["pendulum", "cr3bp", "nbody6", "j2var"]
    gens = {"pendulum": W.pendulum, "cr3bp": W.cr3bp, "nbody6": W.nbody6, "j2var": W.j2var}
    shared = {k: gens[k](1, seed=0).sys for k in kinds}
    out = []
    t = 0.0
    for k in range(n_req):
        t += rng.exponential(1.0 / 50.0)
        kind = kinds[k % 4]
        m = int(10 ** rng.uniform(0, np.log10(4096)))
        wl = gens[kind](m, seed=500 + k)
        out.append(dict(kind=kind, sys=shared[kind], state=wl.state, pars=wl.pars, t_final=10 ** rng.uniform(0, 1, m),
                        arrival=t))
    return out


def test_serving_config_poisson_stream_vs_alone_propagate_and_oracle():
    print("[gpu] serving config: Poisson stream over P / CR / S6 / J2V", flush=True)
    import time

    reqs = _serving_requests()
    srv = Server(capacity=16384, epoch_iters=64)
    t0 = time.perf_counter()
    hs = [srv.submit(r["sys"], r["state"], r["t_final"], pars=r["pars"], tol=1e-15, arrival=t0 + r["arrival"])
          for r in reqs]
    srv.run_until_idle()
    assert all(h.finished is not None for h in hs)
    print(f"[gpu] serving config: {len(hs)} requests, {sum(h.n for h in hs)} instances, {srv.epochs} epochs",
          flush=True)
    for r, h in zip(reqs, hs):
        print(f"[gpu] serving config: {r['kind']} x{h.n}", flush=True)
        alone = Server(capacity=16384, epoch_iters=1 << 20)
        a = alone.submit(r["sys"], r["state"], r["t_final"], pars=r["pars"], tol=1e-15)
        alone.run_until_idle()
        assert np.array_equal(h.result_state, a.result_state)
        assert np.array_equal(h.result, a.result)
        ta = hy.taylor_adaptive_batch(r["sys"], r["state"], pars=r["pars"], tol=1e-15)
        ta.propagate_until(r["t_final"])
        assert np.array_equal(ta.state, h.result_state)
        assert [int(p[3]) for p in ta.propagate_res] == h.result["nsteps"].tolist()
        assert np.all(h.result["outcome"] == int(hy.taylor_outcome.time_limit))
        k = min(h.n, 4)
        pa = None if r["pars"] is None else r["pars"][:, :k]
        _oracle_check(r["sys"], r["state"][:, :k], pa, r["t_final"][:k], h.result_state[:, :k])


# ---- ensemble at the 6-body scale ---------------------------------------------------------
def test_nbody6_ensemble_16x4096_vs_oracle_and_members():
    print("[gpu] nbody6 ensemble: 16 members x 4096 instances", flush=True)
    wl = W.nbody6(65536)
    B, T = 4096, 5.0
    tmpl = hy.taylor_adaptive_batch(wl.sys, wl.state[:, :B].copy(), tol=1e-15)

    def gen(ta, i):
        ta.state[:] = wl.state[:, i * B:(i + 1) * B]
        return ta

    res = hy.ensemble_propagate_until_batch(tmpl, T, 16, gen)
    assert len(res) == 16 and all(r[1] is None and r[2] is None for r in res)
    firsts = np.stack([res[i][0].state[:, 0] for i in range(16)], axis=1)
    _oracle_check(wl.sys, np.ascontiguousarray(wl.state[:, ::B][:, :16]), None, T, firsts)
    for i in (0, 11):
        own = gen(hy.taylor_adaptive_batch(wl.sys, wl.state[:, :B].copy(), tol=1e-15), i)
        own.propagate_until(T)
        assert np.array_equal(own.state, res[i][0].state)
        assert own.propagate_res == res[i][0].propagate_res
    vec = hy.ensemble_propagate_until_batch(tmpl, T, 16, hy.ensemble_arrays(state=wl.state))
    assert len(vec) == 16
    for i in (0, 7, 15):
        assert np.array_equal(vec[i][0].state, res[i][0].state)
        assert vec[i][0].propagate_res == res[i][0].propagate_res
        assert np.array_equal(vec[i][0].time, res[i][0].time)
    # The template is unchanged by the ensemble.
    assert np.array_equal(tmpl.state, wl.state[:, :B]) and np.all(tmpl.time == 0.0)


# ---- J2 + variational over 200 periods ----------------------------------------------------
def test_j2var_stm_200_periods_richardson_fd():
    print("[gpu] j2var: STM over 200 periods on 64 instances vs Richardson FD", flush=True)
    nb = 64
    wl = W.j2var(nb)
    tf = wl.t_final  # 200 periods of each orbit
    ta = hy.taylor_adaptive_batch(wl.sys, wl.state, tol=1e-15)
    ta.propagate_until(tf)
    assert all(p[0] == hy.taylor_outcome.time_limit for p in ta.propagate_res)
    phi = ta.state[6:42].reshape(6, 6, nb)
    x0 = wl.state[:6]
    cols = []
    deltas = []
    for j in range(6):
        d = 1e-6 * np.maximum(1.0, np.abs(x0[j]))
        deltas.append(d)
        for scale in (1.0, 0.5):
            for sgn in (1.0, -1.0):
                xp = x0.copy()
                xp[j] += sgn * scale * d
                cols.append(xp)
    big = np.concatenate(cols, axis=1)  # [6, 24 * nb]
    tb = hy.taylor_adaptive_batch(hy.model.j2(), big, tol=1e-15)
    tb.propagate_until(np.tile(tf, 24))
    out = tb.state.reshape(6, 24, nb)
    pmax = np.max(np.abs(phi), axis=(0, 1))
    worst_fd, worst_rich = 0.0, 0.0
    for j in range(6):
        fd1 = (out[:, 4 * j, :] - out[:, 4 * j + 1, :]) / (2 * deltas[j])
        fd2 = (out[:, 4 * j + 2, :] - out[:, 4 * j + 3, :]) / deltas[j]
        rich = (4.0 * fd2 - fd1) / 3.0
        worst_fd = max(worst_fd, float(np.max(np.abs(fd1 - phi[:, j, :]) / pmax)))
        worst_rich = max(worst_rich, float(np.max(np.abs(rich - phi[:, j, :]) / pmax)))
    print(f"[gpu] j2var 200 periods: max|Phi| {pmax.max():.3e}, FD rel err {worst_fd:.2e}, "
          f"Richardson rel err {worst_rich:.2e}", flush=True)
    assert worst_rich < 1e-6, worst_rich
    assert worst_rich < worst_fd or worst_fd < 1e-9
    Jm = np.block([[np.zeros((3, 3)), np.eye(3)], [-np.eye(3), np.zeros((3, 3))]])
    for b in range(nb):
        P = phi[:, :, b]
        assert np.max(np.abs(P.T @ Jm @ P - Jm)) < 1e-10 * max(1.0, float(np.max(np.abs(P))) ** 2)
    mu, J2, R = 1.0, 1.0826e-3, 1.0

    def energy(s):
        r = np.sqrt(s[0] ** 2 + s[1] ** 2 + s[2] ** 2)
        return 0.5 * (s[3] ** 2 + s[4] ** 2 + s[5] ** 2) - mu / r + 0.5 * mu * J2 * R * R * (
            3 * s[2] ** 2 / r ** 2 - 1) / r ** 3

    de = np.max(np.abs(energy(ta.state[:6]) - energy(x0)) / np.abs(energy(x0)))
    dh = np.max(np.abs((ta.state[0] * ta.state[4] - ta.state[1] * ta.state[3]) - (x0[0] * x0[4] - x0[1] * x0[3])))
    print(f"[gpu] j2var 200 periods: energy drift {de:.2e}, h_z drift {dh:.2e}", flush=True)
    assert de < 1e-10 and dh < 1e-10


# ---- ragged CR3BP, own horizons -------------------------------------------------------------
@pytest.mark.parametrize("mode", ["persistent", "plain"])
def test_cr3bp_ragged_own_horizons_vs_oracle(mode):
    # Near L4 / L5 with offsets of 0.05 many orbits are chaotic: within the horizon the
    # f64 and real128 oracles themselves part for about half of the instances, so the
    # comparison is on the well-conditioned ones, and the Jacobi constant on all.
    print(f"[gpu] cr3bp ragged horizons (clipped to 50), mode={mode}", flush=True)
    n = 192
    wl = W.cr3bp(n)
    tf = np.minimum(wl.t_final, 50.0)
    ta = hy.taylor_adaptive_batch(wl.sys, wl.state, pars=wl.pars, tol=1e-15, gpu_mode=mode)
    ta.propagate_until(tf)
    assert np.array_equal(ta.time, tf)
    assert all(p[0] == hy.taylor_outcome.time_limit for p in ta.propagate_res)
    # Up to t = 50 the well-conditioned instances still carry the chaotic neighbours'
    # sensitivity: 1e-9 here (the f64 oracle itself agrees with real128 to 1e-12 on them).
    ok, r = _conditioned_check(wl.sys, wl.state, wl.pars, tf, ta.state, min_frac=0.3, rtol=1e-9)
    ns = np.array([p[3] for p in ta.propagate_res], dtype=np.int64)
    assert np.max(np.abs(ns[ok] - r["nsteps"].astype(np.int64)[ok])) <= 1
    c0, c1 = _jacobi(wl.state, wl.pars[0]), _jacobi(ta.state, wl.pars[0])
    drift = np.abs((c1 - c0) / c0)
    print(f"[gpu] cr3bp ragged: Jacobi drift max {drift.max():.2e} median {np.median(drift):.2e} "
          f"p90 {np.percentile(drift, 90):.2e}", flush=True)
    # Close approaches to a primary degrade the invariant's conservation for a few orbits.
    assert np.median(drift) < 1e-13 and np.percentile(drift, 90) < 1e-11 and np.max(drift[ok]) < 1e-10


# ---- snapshot / restore --------------------------------------------------------------------
def test_snapshot_restore_finish_vs_oracle_straight_run():
    print("[gpu] snapshot -> restore -> finish vs the oracle's straight run", flush=True)
    n = 48
    wl = W.cr3bp(n, log_t=(1.0, 1.5))
    dec = hy.taylor_decompose(wl.sys)
    system = get_system(dec, 19)
    cap = 128
    a = _lib.Batch(system, cap)
    a.slots_reset()
    slots = np.arange(3, 3 + n, dtype=np.uint32)
    a.slots_write(_lib.FIELD_STATE, slots, wl.state)
    a.slots_write(_lib.FIELD_PARS, slots, wl.pars)
    a.slots_write(_lib.FIELD_TIME_HI, slots, np.zeros((1, n)))
    a.slots_write(_lib.FIELD_TIME_LO, slots, np.zeros((1, n)))
    a.slots_arm(slots, wl.t_final)
    a.slots_launch(37)  # part way: at most 37 steps per instance
    assert a.slots_wait() > 0
    blob = a.slots_snapshot(slots)
    a.close()
    b = _lib.Batch(system, cap)
    b.slots_reset()
    new = np.arange(cap - n, cap, dtype=np.uint32)[::-1].copy()
    b.slots_restore(new, blob)
    b.slots_launch(1 << 30)
    assert b.slots_wait() == 0
    done, res = b.slots_results(new)
    assert done.all() and np.all(res["outcome"] == int(hy.taylor_outcome.time_limit))
    got = b.slots_read(_lib.FIELD_STATE, new, 4)
    _conditioned_check(wl.sys, wl.state, wl.pars, wl.t_final, got, min_frac=0.3, rtol=1e-9)
    ta = hy.taylor_adaptive_batch(wl.sys, wl.state, pars=wl.pars, tol=1e-15)
    ta.propagate_until(wl.t_final)
    assert np.array_equal(ta.state, got)
    assert [int(p[3]) for p in ta.propagate_res] == res["nsteps"].tolist()
    assert [p[1] for p in ta.propagate_res] == res["min_h"].tolist()


# ---- min_h / max_h with terminal events that continue ---------------------------------------
@pytest.mark.parametrize("path", ["device", "stepwise"])
def test_terminal_event_steps_not_in_min_max_h(path):
    print(f"[gpu] terminal events that continue: min_h / max_h, path={path}", flush=True)
    x, v = hy.make_vars("x", "v")
    sys_ = [(x, v), (v, -9.8 * hy.sin(x))]
    ic = np.array([[0.05, 0.1, -0.2, 0.3], [0.025, 0.0, 0.1, -0.05]])
    calls = []

    def cb(ta, d, b):
        calls.append(b)
        return True

    ta = hy.taylor_adaptive_batch(sys_, ic, t_events=[hy.t_event_batch(v, callback=cb)])
    if path == "device":
        ta.propagate_until(5.0)
    else:
        ta.propagate_until(5.0, callback=lambda t: True)
    assert calls
    ref = OE.propagate(hy, sys_, ic, 5.0, t_events=[(v, 0, -1.0, False)])
    for b in range(4):
        oc, mn, mx, ns = ta.propagate_res[b]
        assert int(oc) == int(hy.taylor_outcome.time_limit) and ns == ref["nsteps"][b]
        np.testing.assert_allclose([mn, mx], [ref["min_h"][b], ref["max_h"][b]], rtol=1e-13)
    # Without the events the step rule alone picks (nearly) the same sizes: min_h from the
    # shortened steps would be far smaller.
    plain = hy.taylor_adaptive_batch(sys_, ic)
    plain.propagate_until(5.0)
    for b in range(4):
        assert ta.propagate_res[b][1] > 0.5 * plain.propagate_res[b][1]


# ---- heyoka-shaped decomposition on the GPU -------------------------------------------------
def test_from_heyoka_decomposition_runs_on_gpu():
    print("[gpu] from_heyoka decomposition", flush=True)

    def u(i):
        return hy.expression(f"u_{i}")

    x, v, y = hy.make_vars("x", "v", "y")
    dc = [(x, []), (v, []), (y, []), (hy.sin(u(0)), [4]), (hy.cos(u(0)), [3]),
          (hy.expression._func("pow", u(0), hy.expression(2.0)), []),
          (hy.expression._func("prod", u(5), u(2), u(3)), []),
          (hy.expression._func("mul", hy.expression(-0.5), u(6)), []),
          (hy.expression._func("sub", u(7), u(1)), []), (u(1), []), (u(8), []), (hy.expression(0.0), [])]
    dec = Decomposition.from_heyoka(dc)
    rng = np.random.default_rng(3)
    st = np.stack([rng.uniform(-1, 1, 100), rng.uniform(-0.5, 0.5, 100), rng.uniform(0.5, 1.5, 100)])
    b = _lib.Batch(get_system(dec, 19), 100)
    b.write(_lib.FIELD_STATE, st)
    b.write(_lib.FIELD_TIME_HI, np.zeros(100))
    b.write(_lib.FIELD_TIME_LO, np.zeros(100))
    r = b.propagate_until(np.full(100, 5.0))
    assert np.all(r["outcome"] == int(hy.taylor_outcome.time_limit))
    ref = O.propagate_until(dec, 19, st, 5.0, nthreads=8)
    got = b.read(_lib.FIELD_STATE, st.shape)
    err = np.max(np.abs(got - ref["state"]) / np.maximum(np.abs(ref["state"]), 1.0))
    assert err < 1e-13, err


# ---- tuned launch shapes in the product path -------------------------------------------------
def test_tuned_launch_shape_applied_and_bitwise(tmp_path, monkeypatch):
    print("[gpu] tuned launch shape applied to new batches", flush=True)
    wl = W.cr3bp(300, log_t=(0.5, 1.5))
    dec = hy.taylor_decompose(wl.sys)
    system = get_system(dec, 19)
    ref = hy.taylor_adaptive_batch(wl.sys, wl.state, pars=wl.pars, tol=1e-15)
    ref.propagate_until(wl.t_final)
    monkeypatch.setenv("HYX_CACHE_DIR", str(tmp_path))
    with open(tmp_path / "autotune.json", "w") as f:
        json.dump({system.tune_key: {"waves_per_cu": 1, "refill_min": 32, "rate": 1.0}}, f)
    ta = hy.taylor_adaptive_batch(wl.sys, wl.state, pars=wl.pars, tol=1e-15)
    assert ta._hyx_batch.tuned[:2] == (1, 32)
    ta.propagate_until(wl.t_final)
    assert np.array_equal(ta.state, ref.state) and ta.propagate_res == ref.propagate_res
