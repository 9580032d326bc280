"""The twin and quad forms on the GPU: persistent propagations that hand their last
instances to the twin form (two lanes per instance) or to the quad form (four lanes, the
CR3BP's default) end bit for bit where the plain one-launch-per-step lane form ends, for
the ragged CR3BP ensemble (forward and backward in time), a planar two-primary problem with
drag, and a pendulum pulled by two phase-shifted sines (twin only: sin / cos pairs,
parameter bindings); with the hand-off forced at the start, the whole propagation runs in
the tail form.
"""

import numpy as np
import pytest

import heyoka_amd as hy
from heyoka_amd import _lib
from heyoka_amd import workloads as W
from heyoka_amd.integrator import get_system

pytestmark = pytest.mark.gpu


def _two_sines(n, seed=5):
    x, v = hy.make_vars("x", "v")
    sys_ = [(x, v), (v, -(hy.par[0] * hy.sin(x + hy.par[1])) - hy.par[2] * hy.sin(x - hy.par[3]) - 0.1 * v)]
    rng = np.random.default_rng(seed)
    st = np.stack([rng.uniform(-2, 2, n), rng.uniform(-1, 1, n)])
    pa = np.stack([rng.uniform(0.5, 2, n), rng.uniform(-1, 1, n), rng.uniform(0.5, 2, n), rng.uniform(-1, 1, n)])
    return sys_, st, pa, 10 ** rng.uniform(0.5, 2.0, n)


def _planar_drag(n, seed=9):
    # A planar two-primary problem with smooth Gaussian pulls and a speed-dependent drag
    # (quad form; |v| is common to both components); no singularity, so no instance can
    # spiral into a primary with ever smaller steps.
    x, y, vx, vy = hy.make_vars("x", "y", "vx", "vy")
    mu = hy.par[0]
    d1, d2 = x + mu, x - (1.0 - mu)
    f1, f2 = hy.exp(-(d1 ** 2 + y ** 2)), hy.exp(-(d2 ** 2 + y ** 2))
    sp = hy.sqrt(vx ** 2 + vy ** 2)
    sys_ = [(x, vx), (y, vy), (vx, 2.0 * vy + x - (f1 * d1 * (1.0 - mu) + f2 * d2 * mu) - 0.01 * sp * vx),
            (vy, -2.0 * vx + y - (f1 * y * (1.0 - mu) + f2 * y * mu) - 0.01 * sp * vy)]
    rng = np.random.default_rng(seed)
    st = np.stack([rng.uniform(-1.5, 1.5, n), rng.uniform(0.4, 1.5, n) * rng.choice([-1, 1], n),
                   rng.uniform(-0.3, 0.3, n), rng.uniform(-0.3, 0.3, n)])
    return sys_, st, rng.uniform(0.001, 0.05, (1, n)), 10 ** rng.uniform(0.5, 1.8, n)


def _case(name, n):
    if name == "cr3bp":
        wl = W.cr3bp(n, log_t=(1.0, 2.6))
        return wl.sys, wl.state, wl.pars, wl.t_final
    if name == "planar_drag":
        return _planar_drag(n)
    if name == "cr3bp_back":
        wl = W.cr3bp(n, log_t=(1.0, 2.6), seed=11)
        return wl.sys, wl.state, wl.pars, -wl.t_final
    return _two_sines(n)


def _run(system, st, pa, tf, mode, tail=None):
    N = st.shape[1]
    b = _lib.Batch(system, N)
    if tail is not None:
        b.set_twin_tail(tail)
    b.write(_lib.FIELD_STATE, st)
    b.write(_lib.FIELD_PARS, pa)
    b.write(_lib.FIELD_TIME_HI, np.zeros(N))
    b.write(_lib.FIELD_TIME_LO, np.zeros(N))
    r = b.propagate_until(tf, None, None, 0, mode)
    ms, _ = b.last_timing()
    tail_ms = b.last_tail_ms()
    # The tail phase is part of the propagation's kernel time, and absent without a hand-off.
    assert 0.0 <= tail_ms <= ms and (tail_ms > 0.0) == (b.twin_launches() > 0)
    out = (b.read(_lib.FIELD_STATE, st.shape), b.read(_lib.FIELD_TIME_HI, (N,)), r, b.twin_launches())
    b.close()
    return out


@pytest.mark.parametrize("name", ["cr3bp", "two_sines", "cr3bp_twin", "planar_drag", "cr3bp_back"])
@pytest.mark.parametrize("tail", [None, 1 << 30, 0])
def test_twin_handoff_bitwise_equal_to_plain(name, tail):
    print(f"[gpu] twin hand-off: {name}, tail_min={tail}", flush=True)
    n = 6000
    sys_, st, pa, tf = _case(name.replace("_twin", ""), n)
    dec = hy.taylor_decompose(sys_)
    system = _lib.System(dec, 19, quad=False) if name == "cr3bp_twin" else get_system(dec, 19)
    tsrc = system.twin_source()
    assert tsrc != "" and ("#define HYX_QUAD 1" in tsrc) == (name in ("cr3bp", "planar_drag", "cr3bp_back"))
    ps, pt, pr, _ = _run(system, st, pa, tf, _lib.MODE_PLAIN)
    qs, qt, qr, launches = _run(system, st, pa, tf, _lib.MODE_PERSISTENT, tail)
    if tail == 0:
        assert launches == 0
    else:
        assert launches > 0
    assert np.array_equal(pt, qt)
    # Instances that collide with a primary stop early (err_nf_state) in both runs alike.
    done = pr["outcome"] == int(hy.taylor_outcome.time_limit)
    assert done.mean() > 0.5 and np.array_equal(pt[done], tf[done])
    assert np.array_equal(ps[:, done], qs[:, done])
    for k in ("outcome", "min_h", "max_h", "nsteps"):
        assert np.array_equal(pr[k], qr[k]), k
    print(f"[gpu] twin hand-off: {name}: {launches} twin launches, max nsteps {int(qr['nsteps'].max())}", flush=True)


def test_twin_integrator_api_and_oracle():
    # Through the drop-in API (persistent by default) on a small batch, which the twin
    # form runs entirely, against the oracle.
    import pyoracle as O

    print("[gpu] twin via taylor_adaptive_batch vs oracle", flush=True)
    sys_, st, pa, tf = _two_sines(64)
    tf = np.minimum(tf, 20.0)
    ta = hy.taylor_adaptive_batch(sys_, st, pars=pa, tol=1e-15)
    ta.propagate_until(tf)
    assert ta._hyx_batch.twin_launches() > 0
    r = O.propagate_until(hy.taylor_decompose(sys_), 19, st, tf, pars=pa, nthreads=8, lane_batch=1)
    err = np.max(np.abs(ta.state - r["state"]) / np.maximum(np.abs(r["state"]), 1.0))
    assert err < 1e-11, err


def test_quad_integrator_api_and_oracle():
    # The CR3BP through the drop-in API: a batch small enough to run wholly in the quad
    # form, against the oracle on horizons short enough to stay well conditioned.
    import pyoracle as O

    print("[gpu] quad via taylor_adaptive_batch vs oracle", flush=True)
    wl = W.cr3bp(64, log_t=(0.3, 1.0))
    ta = hy.taylor_adaptive_batch(wl.sys, wl.state, pars=wl.pars, tol=1e-15)
    assert "#define HYX_QUAD 1" in ta._hyx_batch.sys.twin_source()
    ta.propagate_until(wl.t_final)
    assert ta._hyx_batch.twin_launches() > 0
    r = O.propagate_until(hy.taylor_decompose(wl.sys), 19, wl.state, wl.t_final, pars=wl.pars, nthreads=8,
                          lane_batch=1)
    err = np.max(np.abs(ta.state - r["state"]) / np.maximum(np.abs(r["state"]), 1.0))
    assert err < 1e-10, err
