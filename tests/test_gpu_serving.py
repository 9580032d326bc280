"""Continuous-batching server: results alone == in the mix == plain propagate_until.

Three systems (pendulum lane form, ragged CR3BP lane form, 3-body group form)
share the server; small epochs and a small capacity force requests to be
admitted while others run and to wait for free slots. Every result must be bit
for bit the one the request gets alone and the one propagate_until gives.
"""

import numpy as np
import pytest

import heyoka_amd as hy
from heyoka_amd import _lib
from heyoka_amd import workloads as W
from heyoka_amd.integrator import get_system
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


def test_eviction_under_hbm_pressure_continues_bitwise():
    # An HBM budget that holds one pool at a time: admitting a request of a second system
    # parks the first pool mid-flight (slot snapshots to host, device memory freed); it is
    # restored into fresh slots once the second pool drains, and every result is still bit
    # for bit the request's propagate_until.
    print("[gpu] serving: eviction and restore", flush=True)
    reqs = _requests()
    cr_req, pend_req, nb_req = reqs[1], reqs[0], reqs[2]
    cap = 32
    one_pool = cap * 8 * (4 + 1 + 16)  # the largest of the pools involved (CR3BP: 4 states, 1 par)
    big = cap * 8 * (18 + 1 + 16)
    srv = Server(capacity=cap, epoch_iters=8, hbm_budget_bytes=max(one_pool, big) + 1)
    h1 = srv.submit(cr_req["sys"], cr_req["state"], cr_req["t_final"], pars=cr_req["pars"], tol=1e-15)
    for _ in range(4):
        srv.step()
    while any(p.in_epoch for p in srv.pools.values() if hasattr(p, "in_epoch")):
        srv.step()
    assert h1.finished is None  # still running when the other systems arrive
    h2 = srv.submit(pend_req["sys"], pend_req["state"], pend_req["t_final"], tol=1e-15)
    h3 = srv.submit(nb_req["sys"], nb_req["state"], nb_req["t_final"], tol=1e-15)
    srv.run_until_idle()
    assert srv.evictions >= 1
    for r, h in ((cr_req, h1), (pend_req, h2), (nb_req, h3)):
        assert h.finished is not None
        ta = hy.taylor_adaptive_batch(r["sys"], r["state"], pars=r["pars"], tol=1e-15)
        ta.propagate_until(r["t_final"])
        assert np.array_equal(ta.state, h.result_state)
        assert [int(p[3]) for p in ta.propagate_res] == h.result["nsteps"].tolist()
        assert [int(p[0]) for p in ta.propagate_res] == h.result["outcome"].tolist()


@pytest.mark.parametrize("form", ["wave", "compact", "group"])
def test_armed_slots_beyond_the_first_claims(form):
    # A pool whose armed slots all lie beyond the positions the resident waves claim
    # first (the lower slots are free): one epoch must still run every armed slot to its
    # final time, as propagate_until does.
    print(f"[gpu] serving: armed slots high in the pool, form={form}", flush=True)
    wl = W.j2var(40) if form != "group" else W.nbody6(40)
    sys_ = wl.sys.sys if hasattr(wl.sys, "n_orig_sv") else wl.sys
    dec = hy.taylor_decompose(sys_)
    system = get_system(dec, 19, compact=form if form != "compact" else True)
    assert system.form == form
    cap = 16384
    b = _lib.Batch(system, cap)
    b.slots_reset()
    slots = np.arange(cap - 40, cap, dtype=np.uint32)
    b.slots_write(_lib.FIELD_STATE, slots, wl.state)
    b.slots_write(_lib.FIELD_TIME_HI, slots, np.zeros((1, 40)))
    b.slots_write(_lib.FIELD_TIME_LO, slots, np.zeros((1, 40)))
    b.slots_arm(slots, np.full(40, 3.0))
    b.slots_launch(100000)
    assert b.slots_wait() == 0
    done, res = b.slots_results(slots)
    assert done.all() and np.all(res["outcome"] == int(hy.taylor_outcome.time_limit))
    ta = hy.taylor_adaptive_batch(wl.sys, wl.state, tol=1e-15, code_form=form)
    ta.propagate_until(3.0)
    assert np.array_equal(b.slots_read(_lib.FIELD_STATE, slots, dec.n_state), ta.state)
    assert res["nsteps"].tolist() == [int(p[3]) for p in ta.propagate_res]
