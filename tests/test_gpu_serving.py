"""Continuous-batching server: results alone == in the mix == plain propagate_until.

Three systems (pendulum lane form, ragged CR3BP lane form, 3-body group form)
share the server; small epochs and a small capacity force requests to be
admitted while others run and to wait for free slots. Every result must be bit
for bit the one the request gets alone and the one propagate_until gives.
"""

import numpy as np
import pytest

import heyoka_amd as hy
from heyoka_amd import workloads as W
from heyoka_amd.serving import Server

pytestmark = pytest.mark.gpu


def _requests():
    rng = np.random.default_rng(11)
    out = []
    pend = hy.model.pendulum()
    nb3 = hy.model.nbody(3, masses=[1.0, 0.3, 0.01])
    base3 = np.array([0, 0, 0, 0, 0, 0, 1, 0, 0.1, 0, 1.1, 0, -2, 0.5, 0, 0.2, -0.6, 0.05], dtype=float)
    cr = W.cr3bp(64, log_t=(0.5, 1.5))
    for i in range(9):
        kind = i % 3
        if kind == 0:
            m = int(rng.integers(1, 6))
            out.append(dict(sys=pend, state=np.stack([rng.uniform(-1, 1, m), rng.uniform(-0.5, 0.5, m)]),
                            t_final=rng.uniform(5, 20, m), pars=None))
        elif kind == 1:
            sl = slice(7 * i, 7 * i + 20)
            out.append(dict(sys=cr.sys, state=cr.state[:, sl], t_final=cr.t_final[sl], pars=cr.pars[:, sl]))
        else:
            m = int(rng.integers(3, 12))
            out.append(dict(sys=nb3, state=base3[:, None] + rng.normal(0, 1e-3, (18, m)),
                            t_final=np.full(m, rng.uniform(1, 3)), pars=None))
    return out


def _run(reqs, capacity=48, epoch_iters=40, stagger=True):
    srv = Server(capacity=capacity, epoch_iters=epoch_iters)
    handles = []
    for k, r in enumerate(reqs):
        handles.append(srv.submit(r["sys"], r["state"], r["t_final"], pars=r["pars"], tol=1e-15))
        if stagger and k % 2 == 1:
            srv.step()  # later requests arrive while earlier ones are on the device
    srv.run_until_idle()
    return srv, handles


def test_alone_equals_mixed_equals_propagate():
    print("[gpu] serving: mixed vs alone vs propagate_until", flush=True)
    reqs = _requests()
    srv, mixed = _run(reqs)
    assert srv.epochs > 3
    assert all(h.finished is not None for h in mixed)
    lat = srv.latency_percentiles()
    print(f"[gpu] serving: {len(mixed)} requests, {srv.epochs} epochs, p50 {lat['p50']*1e3:.1f} ms "
          f"p99 {lat['p99']*1e3:.1f} ms", flush=True)
    for r, h in zip(reqs, mixed):
        _, alone = _run([r], stagger=False)
        a = alone[0]
        assert np.array_equal(h.result_state, a.result_state)
        assert np.array_equal(h.result_time, a.result_time)
        assert np.array_equal(h.result, a.result)
        ta = hy.taylor_adaptive_batch(r["sys"], r["state"], pars=r["pars"], tol=1e-15)
        ta.propagate_until(r["t_final"])
        assert np.array_equal(ta.state, h.result_state)
        assert [int(p[3]) for p in ta.propagate_res] == h.result["nsteps"].tolist()
        assert np.all(h.result["outcome"] == int(hy.taylor_outcome.time_limit))


def test_capacity_pressure_and_validation():
    print("[gpu] serving: capacity pressure", flush=True)
    reqs = _requests()[:4]
    srv, hs = _run(reqs, capacity=24, epoch_iters=16)
    assert all(h.finished is not None for h in hs)
    with pytest.raises(ValueError):
        srv.submit(hy.model.pendulum(), np.zeros((2, 25)), 1.0)
