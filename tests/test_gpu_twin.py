"""The twin and quad forms on the GPU: persistent propagations that hand their last
instances to the twin form (two lanes per instance) or to the quad form (four lanes, the
CR3BP's default) end bit for bit where the plain one-launch-per-step lane form ends, for
the ragged CR3BP ensemble and a pendulum pulled by two phase-shifted sines (twin only:
sin / cos pairs, parameter bindings); with the hand-off forced at the start, the whole
propagation runs in the tail form.
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


def _case(name, n):
    if name == "cr3bp":
        wl = W.cr3bp(n, log_t=(1.0, 2.6))
        return wl.sys, wl.state, wl.pars, wl.t_final
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
    out = (b.read(_lib.FIELD_STATE, st.shape), b.read(_lib.FIELD_TIME_HI, (N,)), r, b.twin_launches())
    b.close()
    return out


@pytest.mark.parametrize("name", ["cr3bp", "two_sines", "cr3bp_twin"])
@pytest.mark.parametrize("tail", [None, 1 << 30, 0])
def test_twin_handoff_bitwise_equal_to_plain(name, tail):
    print(f"[gpu] twin hand-off: {name}, tail_min={tail}", flush=True)
    n = 6000
    sys_, st, pa, tf = _case(name.replace("_twin", ""), n)
    dec = hy.taylor_decompose(sys_)
    system = _lib.System(dec, 19, quad=False) if name == "cr3bp_twin" else get_system(dec, 19)
    tsrc = system.twin_source()
    assert tsrc != "" and ("#define HYX_QUAD 1" in tsrc) == (name == "cr3bp")
    ps, pt, pr, _ = _run(system, st, pa, tf, _lib.MODE_PLAIN)
    qs, qt, qr, launches = _run(system, st, pa, tf, _lib.MODE_PERSISTENT, tail)
    if tail == 0:
        assert launches == 0
    else:
        assert launches > 0
    assert np.array_equal(pt, qt) and np.array_equal(pt, tf)
    assert np.array_equal(ps, qs)
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
