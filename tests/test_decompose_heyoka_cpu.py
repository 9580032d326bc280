"""Lowering heyoka's own Taylor decomposition (the ``decomposition`` property,
heyoka/expose_batch_integrators.cpp:231: n variables, m elementary operations over
u variables, n right-hand sides) to the op-list, without a GPU.

- Round trip: the decomposition this package exports in heyoka's shape lowers back to
  the identical op-list words and constants for the pendulum, the forced damped
  pendulum, the planar CR3BP, the 6-body problem and J2 + variational equations.
- A decomposition written the way heyoka writes one (n-ary ``prod``, ``pow(u, 2.)``, a
  sin whose cos partner is only a hidden dependency, ``(expression, hidden_deps)``
  pairs) lowers to an op-list whose oracle propagation matches the one of the
  package's own decomposition of the same system.
"""

import numpy as np
import pytest

import heyoka_amd as hy
from heyoka_amd import workloads as W
from heyoka_amd.decompose import Decomposition

import pyoracle as O


def _systems():
    x, v = hy.make_vars("x", "v")
    fdp = [(x, v), (v, hy.cos(hy.time) - hy.par[0] * v - hy.sin(x))]
    return {"P": W.pendulum().sys, "FDP": fdp, "CR": W.cr3bp(4).sys, "S6": W.nbody6(4).sys,
            "J2V": W.j2var(4).sys.sys}


@pytest.mark.parametrize("name", ["P", "FDP", "CR", "S6", "J2V"])
@pytest.mark.parametrize("shape", ["expressions", "pairs"])
def test_round_trip_identical_words(name, shape):
    dec = hy.taylor_decompose(_systems()[name])
    src = dec.as_expressions() if shape == "expressions" else dec.as_heyoka()
    d2 = Decomposition.from_heyoka(src)
    assert d2.state_names == dec.state_names and d2.n_par == dec.n_par
    np.testing.assert_array_equal(d2.to_words(), dec.to_words())
    np.testing.assert_array_equal(d2.consts_array(), dec.consts_array())


def test_hidden_deps_pair_sin_cos():
    x, v = hy.make_vars("x", "v")
    dc = hy.taylor_decompose([(x, v), (v, -9.8 * hy.sin(x))]).as_heyoka()
    kinds = {str(e): d for e, d in dc}
    s_idx = [i for i, (e, _) in enumerate(dc) if str(e).startswith("sin")][0]
    c_idx = [i for i, (e, _) in enumerate(dc) if str(e).startswith("cos")][0]
    assert dc[s_idx][1] == [c_idx] and dc[c_idx][1] == [s_idx]
    assert all(d == [] for e, d in dc if not str(e).startswith(("sin", "cos"))) and kinds


def _u(i):
    return hy.expression(f"u_{i}")


def test_heyoka_style_decomposition_propagates_like_own():
    # x' = v, v' = -(1/2) x^2 * y * sin(x) - v, y' = 0 (a constant-amplitude factor), written
    # as heyoka would: a sin with only a hidden cos, pow(u, 2.), a 3-factor prod, n-ary sum.
    x, v, y = hy.make_vars("x", "v", "y")
    own = hy.taylor_decompose([(x, v), (v, -0.5 * (x ** 2 * y * hy.sin(x)) - v), (y, hy.expression(0.0))])
    dc = [
        (x, []), (v, []), (y, []),
        (hy.sin(_u(0)), [4]),                                  # u_3
        (hy.cos(_u(0)), [3]),                                  # u_4
        (hy.expression._func("pow", _u(0), hy.expression(2.0)), []),  # u_5 = x^2
        (hy.expression._func("prod", _u(5), _u(2), _u(3)), []),        # u_6 = x^2 y sin x
        (hy.expression._func("mul", hy.expression(-0.5), _u(6)), []),  # u_7
        (hy.expression._func("sub", _u(7), _u(1)), []),               # u_8
        (_u(1), []), (_u(8), []), (hy.expression(0.0), []),
    ]
    low = Decomposition.from_heyoka(dc)
    assert low.state_names == ["x", "v", "y"] and len(low.nodes) == 7
    rng = np.random.default_rng(11)
    st = np.stack([rng.uniform(-1, 1, 8), rng.uniform(-0.5, 0.5, 8), rng.uniform(0.5, 1.5, 8)])
    a = O.propagate_until(low, 19, st.copy(), 5.0)
    b = O.propagate_until(own, 19, st.copy(), 5.0)
    assert np.all(a["outcome"] == O.TIME_LIMIT)
    np.testing.assert_allclose(a["state"], b["state"], rtol=1e-13, atol=1e-14)


def test_from_heyoka_errors():
    x, v = hy.make_vars("x", "v")
    with pytest.raises(ValueError, match="not defined before it"):
        Decomposition.from_heyoka([x, v, hy.sin(_u(3)), _u(1), _u(2)])
    with pytest.raises(ValueError, match="nested operand"):
        Decomposition.from_heyoka([x, v, hy.sin(_u(0) + 1.0), _u(1), _u(2)])
    with pytest.raises(NotImplementedError, match="Unsupported elementary operation"):
        Decomposition.from_heyoka([x, v, hy.expression._func("neg", _u(0), _u(1)), _u(1), _u(2)])
