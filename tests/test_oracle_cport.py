"""The compiled CPU port (the CPU baseline) against the interpreter oracle."""

import time

import numpy as np

import heyoka_amd as hy
import cport
import pyoracle as O
from heyoka_amd import workloads as W


def test_cport_matches_oracle():
    for wl, tf in ((W.cr3bp(64, log_t=(0.5, 1.0)), 8.0), (W.nbody6(16), 1.0), (W.pendulum(), 10.0)):
        sys_ = wl.sys
        dec = hy.taylor_decompose(sys_)
        p = hy.order_from_tol(wl.tol)
        a = cport.propagate_until(dec, p, wl.state, tf, pars=wl.pars, nthreads=4)
        b = O.propagate_until(dec, p, wl.state, tf, pars=wl.pars, nthreads=4)
        np.testing.assert_allclose(a["state"], b["state"], rtol=1e-12, atol=1e-14)
        assert np.max(np.abs(a["nsteps"].astype(np.int64) - b["nsteps"].astype(np.int64))) <= 1


def test_cport_is_faster_than_interpreter():
    wl = W.nbody6(32)
    dec = hy.taylor_decompose(wl.sys)
    cport.build(dec, 19)
    t = time.perf_counter()
    a = cport.propagate_until(dec, 19, wl.state, 2.0, nthreads=1)
    ta = time.perf_counter() - t
    t = time.perf_counter()
    b = O.propagate_until(dec, 19, wl.state, 2.0, nthreads=1, lane_batch=8)
    tb = time.perf_counter() - t
    print(f"cport {a['nsteps'].sum() / ta:.0f} vs interpreter {b['nsteps'].sum() / tb:.0f} instance-steps/s")
    assert ta < tb
