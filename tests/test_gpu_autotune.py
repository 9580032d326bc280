"""Autotuner: launch configurations never change results (bitwise), and it picks one."""

import numpy as np
import pytest

import heyoka_amd as hy
from heyoka_amd import autotune as AT
from heyoka_amd import workloads as W
from heyoka_amd.integrator import get_system

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", ["cr3bp", "nbody3"])
def test_launch_configs_bitwise_identical(case, tmp_path, monkeypatch):
    print(f"[gpu] autotune bitwise identity: {case}", flush=True)
    if case == "cr3bp":
        wl = W.cr3bp(500, log_t=(0.5, 1.2))
        sys_, st, pars, tf = wl.sys, wl.state, wl.pars, wl.t_final
    else:
        sys_ = hy.model.nbody(3, masses=[1.0, 0.3, 0.01])
        base = np.array([0, 0, 0, 0, 0, 0, 1, 0, 0.1, 0, 1.1, 0, -2, 0.5, 0, 0.2, -0.6, 0.05], dtype=float)
        st = base[:, None] + np.random.default_rng(3).normal(0, 1e-3, (18, 200))
        pars, tf = None, np.linspace(0.5, 2.0, 200)
    dec = hy.taylor_decompose(sys_)
    system = get_system(dec, 19)
    finals = []
    for w, rf in [(0, 1), (1, 8), (2, 32), (3, 64)]:
        _, fin = AT.measure(system, st, pars, tf, w, rf)
        finals.append(fin)
    for f in finals[1:]:
        assert np.array_equal(f, finals[0])
    monkeypatch.setenv("HYX_CACHE_DIR", str(tmp_path))
    best, table = AT.autotune(dec, 19, st, pars, tf, waves=(0, 2), refills=(1, 16))
    assert best["rate"] > 0 and len(table) == 4
    assert AT.tuned_config(dec, 19) == (best["waves_per_cu"], best["refill_min"])
