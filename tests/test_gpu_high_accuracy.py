"""high_accuracy=True (compensated summation in the state update) on the GPU, all code forms,
against the oracle's high-accuracy mode (heyoka/expose_batch_integrators.cpp:228-230 kwarg;
doc/notebooks/Customising the adaptive integrator.ipynb, cell 12)."""

import numpy as np
import pytest

import heyoka_amd as hy
import pyoracle as O
from heyoka_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("code_form", ["unrolled", "compact", "wave"])
def test_pendulum_high_accuracy(code_form):
    print(f"[gpu] high accuracy pendulum {code_form}", flush=True)
    wl = W.pendulum(64, seed=7)
    ta = hy.taylor_adaptive_batch(wl.sys, wl.state, high_accuracy=True, code_form=code_form)
    assert ta.high_accuracy
    ta.propagate_until(20.0)
    ref = O.propagate_until(hy.taylor_decompose(wl.sys), ta.order, wl.state, 20.0, high_accuracy=True, lane_batch=8)
    np.testing.assert_allclose(ta.state, ref["state"], rtol=1e-12, atol=1e-14)
    assert [r[3] for r in ta.propagate_res] == [int(x) for x in ref["nsteps"]]
    # Plain Horner differs from the compensated update only at the ulp level per step.
    tb = hy.taylor_adaptive_batch(wl.sys, wl.state, code_form=code_form)
    tb.propagate_until(20.0)
    assert np.max(np.abs(tb.state - ta.state)) < 1e-11


def test_group_form_high_accuracy():
    print("[gpu] high accuracy nbody6 group", flush=True)
    wl = W.nbody6(32)
    ta = hy.taylor_adaptive_batch(wl.sys, wl.state, tol=1e-15, high_accuracy=True, code_form="group")
    assert ta.code_form == "group"
    ta.propagate_until(1.0)
    ref = O.propagate_until(hy.taylor_decompose(wl.sys), 19, wl.state, 1.0, high_accuracy=True, lane_batch=8,
                            nthreads=4)
    scale = np.max(np.abs(ref["state"]), axis=1, keepdims=True)
    assert np.max(np.abs(ta.state - ref["state"]) / scale) < 1e-11
