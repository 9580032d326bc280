"""Compiled functions on the GPU: values against numpy, the reference's single-point call
(heyoka/_test_cfunc.py:35-37), copies, and device diagnostics on a batch's resident state."""

import pickle
from copy import copy, deepcopy

import numpy as np
import pytest

import heyoka_amd as hy
from heyoka_amd import workloads as W
from heyoka_amd.expr import evaluate

pytestmark = pytest.mark.gpu


def test_single_and_multi_point():
    print("[gpu] cfunc values", flush=True)
    x, y, z = hy.make_vars("x", "y", "z")
    cf = hy.cfunc([y * (x + z)], [x, y, z])
    assert cf([1, 2, 3])[0] == 8.0
    assert cf([1, 2, 3]) == copy(cf)([1, 2, 3]) == deepcopy(cf)([1, 2, 3]) == pickle.loads(pickle.dumps(cf))([1, 2, 3])
    fn = [hy.sin(x) * hy.exp(-y) + hy.par[1] * hy.time, (x ** 2 + y ** 2 + z ** 2) ** -1.5, hy.par[0] - z]
    cf = hy.cfunc(fn, [x, y, z])
    rng = np.random.default_rng(1)
    N = 1000
    xin = rng.uniform(0.5, 2.0, (3, N))
    pars = rng.uniform(-1, 1, (2, N))
    t = rng.uniform(0, 3, N)
    out = cf(xin, pars=pars, time=t)
    assert out.shape == (3, N)
    env = {"x": xin[0], "y": xin[1], "z": xin[2]}
    for q, f in enumerate(fn):
        np.testing.assert_allclose(out[q], evaluate(f, env, pars, t), rtol=1e-14, atol=1e-15)
    with pytest.raises(ValueError):
        cf(np.zeros((3, 2, 2)))
    with pytest.raises(TypeError):
        cf(xin, pars=pars, time=1.0)


def test_device_diagnostics_on_a_batch():
    # The energy of every instance of a 6-body batch, read from its device-resident state
    # after a propagation, against numpy on the host copy of the same state.
    print("[gpu] cfunc eval_batch", flush=True)
    wl = W.nbody6(256)
    ta = hy.taylor_adaptive_batch(wl.sys, wl.state, tol=1e-15)
    ta.propagate_until(0.5)
    svars = [l for l, _ in wl.sys]
    en = hy.model.nbody_energy(6, masses=list(wl.masses))
    cf = hy.cfunc([en], svars)
    e_dev = cf.eval_batch(ta)[0]
    env = {v.name: ta.state[i] for i, v in enumerate(svars)}
    e_host = evaluate(en, env)
    np.testing.assert_allclose(e_dev, e_host, rtol=1e-13)
    e0 = cf(wl.state)[0]
    assert np.median(np.abs((e_dev - e0) / e0)) < 1e-12
