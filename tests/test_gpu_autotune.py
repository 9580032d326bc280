"""Autotuner: launch configurations never change results (bitwise), and it picks one."""

import numpy as np
import pytest

import heyoka_amd as hy
from heyoka_amd import autotune as AT
from heyoka_amd import workloads as W
from heyoka_amd.integrator import get_system

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", ["cr3bp", "nbody3", "j2var"])
def test_launch_configs_bitwise_identical(case, tmp_path, monkeypatch):
    print(f"[gpu] autotune bitwise identity: {case}", flush=True)
    if case == "cr3bp":
        wl = W.cr3bp(500, log_t=(0.5, 1.2))
        sys_, st, pars, tf = wl.sys, wl.state, wl.pars, wl.t_final
    elif case == "j2var":
        wl = W.j2var(300)
        sys_, st, pars, tf = wl.sys.sys, wl.state, None, np.linspace(1.0, 4.0, 300)
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


def test_native_autotune_restores_contents_and_caches(tmp_path):
    # hyx_autotune times every candidate from the batch's current contents, restores them,
    # keeps the winner, and reuses a cached result; the propagation after tuning is bitwise
    # the untuned one.
    print("[gpu] native autotune", flush=True)
    from heyoka_amd import _lib

    wl = W.cr3bp(2000, log_t=(0.5, 1.2))
    dec = hy.taylor_decompose(wl.sys)
    system = get_system(dec, 19)

    def fresh():
        b = _lib.Batch(system, wl.state.shape[1])
        b.write(_lib.FIELD_STATE, wl.state)
        b.write(_lib.FIELD_PARS, wl.pars)
        b.write(_lib.FIELD_TIME_HI, np.zeros(wl.state.shape[1]))
        b.write(_lib.FIELD_TIME_LO, np.zeros(wl.state.shape[1]))
        return b

    ref = fresh()
    ref.propagate_until(wl.t_final)
    ref_state = ref.read(_lib.FIELD_STATE, wl.state.shape)
    b = fresh()
    cache = tmp_path / "tune.txt"
    r1 = b.autotune(wl.t_final, cache)
    assert r1["rate"] > 0 and not r1["from_cache"]
    np.testing.assert_array_equal(b.read(_lib.FIELD_STATE, wl.state.shape), wl.state)  # contents restored
    b.propagate_until(wl.t_final)
    np.testing.assert_array_equal(b.read(_lib.FIELD_STATE, wl.state.shape), ref_state)
    r2 = fresh().autotune(wl.t_final, cache)
    assert r2["from_cache"] and (r2["waves_per_cu"], r2["refill_min"]) == (r1["waves_per_cu"], r1["refill_min"])


def test_autotune_instances_per_wavefront(tmp_path, monkeypatch):
    # The wave form's lanes per instance (4, 2 or 1 instances per wavefront) as autotuner
    # candidates: each is its own cache entry, and all of them compute every node with the
    # same recurrence and the group reductions are maxima, so they agree bit for bit.
    print("[gpu] autotune over instances per wavefront (J2 + variational)", flush=True)
    monkeypatch.setenv("HYX_CACHE_DIR", str(tmp_path))
    wl = W.j2var(256)
    dec = hy.taylor_decompose(wl.sys.sys)
    forms = AT.wave_forms(dec, 19)
    assert forms == ["wave16", "wave32", "wave"]
    tf = np.full(256, 2.0)
    best, table = AT.autotune(dec, 19, wl.state, None, tf, forms=tuple(forms), waves=(0,), refills=(1,))
    assert sorted(r["form"] for r in table) == sorted(forms) and best["rate"] > 0
    assert AT.tuned_config(dec, 19, best["form"]) == (best["waves_per_cu"], best["refill_min"])
    finals = [AT.measure(get_system(dec, 19, compact=f), wl.state, None, tf, 0, 1)[1] for f in forms]
    for f in finals[1:]:
        assert np.array_equal(f, finals[0])
