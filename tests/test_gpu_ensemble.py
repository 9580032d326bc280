"""Ensemble propagation on the GPU (assertions of heyoka/_test_ensemble.py:13-178).

Each member of a packed ensemble must end bit for bit where its own
propagate_until / propagate_for takes it, with the same propagate_res.
"""

import numpy as np
import pytest

import heyoka_amd as hy

pytestmark = pytest.mark.gpu


def _setup():
    x, v = hy.make_vars("x", "v")
    sys_ = [(x, v), (v, -9.8 * hy.sin(x))]
    ta = hy.taylor_adaptive_batch(sys_, state=[[0.0] * 4] * 2)
    ics = np.zeros((10, 2, 4))
    for i in range(10):
        ics[i, 0] = [0.05 + i / 100, 0.051 + i / 100, 0.052 + i / 100, 0.053 + i / 100.0]
        ics[i, 1] = [0.025 + i / 100, 0.026 + i / 100, 0.027 + i / 100, 0.028 + i / 100.0]
    return ta, ics


@pytest.mark.parametrize("algo", ["thread", "process"])
def test_ensemble_until_for_bitwise(algo):
    print(f"[gpu] ensemble algorithm={algo}", flush=True)
    ta, ics = _setup()

    def gen(t, idx):
        t.set_time(0.0)
        t.state[:] = ics[idx]
        return t

    kw = {"max_workers": 8} if algo == "thread" else {"max_workers": 8, "chunksize": 3}
    ret = hy.ensemble_propagate_until_batch(ta, 20.0, 10, gen, algorithm=algo, **kw)
    assert len(ret) == 10
    for i in range(10):
        ta.set_time(0.0)
        ta.state[:] = ics[i]
        ta.propagate_until(20.0)
        assert np.allclose(ret[i][0].time, 20.0)
        assert np.array_equal(ta.state, ret[i][0].state)
        assert ret[i][1] is None
        assert np.array_equal(ta.time, ret[i][0].time)
        assert ta.propagate_res == ret[i][0].propagate_res

    def gen10(t, idx):
        t.set_time(10.0)
        t.state[:] = ics[idx]
        return t

    ret = hy.ensemble_propagate_for_batch(ta, 20.0, 10, gen10, algorithm=algo, **kw)
    for i in range(10):
        ta.set_time(10.0)
        ta.state[:] = ics[i]
        ta.propagate_for(20.0)
        assert np.allclose(ret[i][0].time, 30.0)
        assert np.array_equal(ta.state, ret[i][0].state)
        assert ta.propagate_res == ret[i][0].propagate_res


def test_ensemble_with_callback_copies():
    print("[gpu] ensemble with a step callback (deep-copied per member)", flush=True)
    ta, ics = _setup()

    class Counter:
        def __init__(self):
            self.n = 0

        def __call__(self, t):
            self.n += 1
            return True

    cb = Counter()

    def gen(t, idx):
        t.set_time(0.0)
        t.state[:] = ics[idx]
        return t

    ret = hy.ensemble_propagate_until(ta, 1.0, 3, gen, callback=cb)
    assert cb.n == 0
    assert all(r[2] is not cb and r[2].n > 0 for r in ret)
