"""Host-side behaviour of the drop-in API, without a GPU.

Assertions follow the reference's own tests (their assertions, not their code):
ensemble argument validation (heyoka/_test_ensemble.py:393-438), model
structure (heyoka/_test_model.py:153-278), var_ode_sys
(heyoka/_test_var_ode_sys.py:14-69, heyoka/_test_var_integrator.py:173-201),
batch constructor validation (heyoka/expose_batch_integrators.cpp:101-157).
"""

import copy
import pickle

import numpy as np
import pytest

import heyoka_amd as hy
from heyoka_amd import model


def gen(ta, i):
    return ta


def test_ensemble_validation_messages():
    with pytest.raises(TypeError, match="The n_iter parameter must be an integer, but an object of type"):
        hy.ensemble_propagate_until(None, 20.0, "a", gen)
    with pytest.raises(ValueError, match="The n_iter parameter must be non-negative"):
        hy.ensemble_propagate_until(None, 20.0, -1, gen)
    msg = ("Cannot perform an ensemble propagate_until/for\\(\\): the final epoch/time interval must be a scalar, "
           "not an iterable object")
    with pytest.raises(TypeError, match=msg):
        hy.ensemble_propagate_until(None, [20.0], 10, gen)
    with pytest.raises(TypeError, match=msg):
        hy.ensemble_propagate_for(None, [20.0], 10, gen)
    with pytest.raises(ValueError, match="the input time grid must be one-dimensional, but instead it has 2 dimensions"):
        hy.ensemble_propagate_grid(None, [[20.0, 20.0]], 10, gen)
    with pytest.raises(TypeError, match='the "max_delta_t" argument must be a scalar, not an iterable object'):
        hy.ensemble_propagate_until(None, 20.0, 10, gen, max_delta_t=[10])
    with pytest.raises(TypeError):
        hy.ensemble_propagate_until(None, 20.0, 10, gen, chunksize=1)
    with pytest.raises(ValueError, match="The parallelisation algorithm must be one of"):
        hy.ensemble_propagate_until(None, 20.0, 10, gen, algorithm="mpi")
    assert hy.ensemble_propagate_until(None, 20.0, 0, gen) == []


def test_models_structure():
    x, v = hy.make_vars("x", "v")
    dyn = model.pendulum()
    assert dyn[0] == (x, v) and dyn[1][0] == v and dyn[1][1] == -hy.sin(x)
    assert model.pendulum(gconst=2.0)[1][1] == -2.0 * hy.sin(x)
    assert model.pendulum(gconst=4.0, length=2.0)[1][1] == -2.0 * hy.sin(x)
    assert model.pendulum_energy() == ((0.5 * v ** 2) + (1.0 - hy.cos(x)))
    assert model.pendulum_energy(gconst=2.0) == ((0.5 * v ** 2) + (2.0 * (1.0 - hy.cos(x))))
    assert model.pendulum_energy(length=2.0, gconst=4.0) == ((2.0 * v ** 2) + (8.0 * (1.0 - hy.cos(x))))
    dyn = model.nbody(2, masses=[0.0, 0.0])
    assert len(dyn) == 12
    for q in (3, 4, 5, 9, 10, 11):
        assert dyn[q][1] == hy.expression(0.0)
    dyn = model.nbody(2, Gconst=5.0)
    for q in (3, 4, 5, 9, 10, 11):
        assert "5.0000000000000" in str(dyn[q][1])
    assert model.nbody_energy(2, masses=[0.0, 0.0]) == hy.expression(0.0)
    assert "5.0000000000000" in str(model.nbody_energy(2, Gconst=5.0))
    xx, px, yy = hy.make_vars("x", "px", "y")
    assert model.cr3bp()[0] == (xx, px + yy)
    assert "0.06250000000" in str(model.cr3bp(mu=1.0 / 2 ** 4)[3][1])
    assert "0.00100000" in str(model.cr3bp_jacobi())


def test_var_ode_sys_properties():
    x, v = hy.make_vars("x", "v")
    va = hy.var_args
    assert va.vars | va.time | va.params == va.all
    assert (va.vars | va.time) & va.time
    orig = [(x, v), (v, -hy.par[0] * hy.sin(x) + hy.time)]
    vs = hy.var_ode_sys(orig, va.vars)
    assert orig == vs.sys[:2] and vs.vargs == [x, v] and vs.n_orig_sv == 2 and vs.order == 1
    vs = hy.var_ode_sys(sys=orig, args=[v, hy.time, x], order=2)
    assert orig == vs.sys[:2] and vs.vargs == [v, hy.time, x] and vs.order == 2
    vs = hy.var_ode_sys(orig, va.vars | va.params)
    assert vs.vargs == [x, v, hy.par[0]]
    for vs2 in (copy.copy(vs), copy.deepcopy(vs), pickle.loads(pickle.dumps(vs))):
        assert orig == vs2.sys[:2] and vs2.vargs == [x, v, hy.par[0]] and vs2.n_orig_sv == 2 and vs2.order == 1
    # Layout of an order-2 system (heyoka/_test_var_integrator.py:190-201).
    vs = hy.var_ode_sys([(x, v), (v, hy.cos(hy.time) - hy.par[0] * v - hy.sin(x))], va.vars, order=2)
    assert vs.get_vslice(0) == slice(0, 2) and vs.get_vslice(0, 1) == slice(1, 2)
    assert vs.get_vslice(1) == slice(2, 6) and vs.get_vslice(1, 1) == slice(4, 6)
    assert [vs.get_mindex(i) for i in range(6)] == [[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 0], [1, 0, 1]]


def test_time_varg_initial_condition():
    # d x / d t0 starts at -f(x0, t0).
    x, v = hy.make_vars("x", "v")
    vs = hy.var_ode_sys([(x, v), (v, -hy.par[0] * hy.sin(x) + hy.time)], [hy.time], order=1)
    ic = vs.initial_derivatives(np.array([0.3, 0.2]), np.array([0.5]), 0.7)
    assert ic[0] == -0.2 and abs(ic[1] - (0.5 * np.sin(0.3) - 0.7)) < 1e-16


def test_batch_constructor_validation():
    x, v = hy.make_vars("x", "v")
    sys_ = [(x, v), (v, -9.8 * hy.sin(x))]
    with pytest.raises(ValueError, match="the expected number of dimensions is 2, but the input array has a "
                                         "dimension of 1"):
        hy.taylor_adaptive_batch(sys_, [0.1, 0.2])
    with pytest.raises(ValueError, match="Invalid parameter vector passed to the constructor of a batch integrator"):
        hy.taylor_adaptive_batch(sys_, [[0.1, 0.2], [0.0, 0.0]], pars=[[1.0, 2.0, 3.0]])
    with pytest.raises(ValueError, match="Invalid time vector passed to the constructor of a batch integrator"):
        hy.taylor_adaptive_batch(sys_, [[0.1, 0.2], [0.0, 0.0]], time=[0.0])
    with pytest.raises(TypeError, match="not recognized/supported"):
        hy.taylor_adaptive_batch(sys_, [[0.1], [0.0]], fp_type=np.float32)
    with pytest.raises(ValueError, match="not a state variable"):
        hy.taylor_decompose([(x, hy.make_vars("z"))])


def test_decomposition_and_flop_model():
    x, v = hy.make_vars("x", "v")
    dec = hy.taylor_decompose([(x, v), (v, -hy.sin(x))])
    ex = dec.as_expressions()
    # n variable definitions, m elementary ops (sin, cos, neg), n right-hand sides.
    assert len(ex) == 2 + 3 + 2
    # Hand count at order 20: order 0 = 3 nodes; sin/cos pair sum_{k=1}^{19} (4k+3) = 817;
    # neg 19; state variables 2*20; step size 2*2; Horner 2*2*20.
    fl = dec.algorithmic_flops(20)
    assert fl["total"] == 3 + 817 + 19 + 40 + 4 + 80 == 963
    assert fl["transcendentals"] == 2
    w = dec.to_words()
    assert w[0] == 0x31585948 and w[1] == 2 and w[3] == 5


def test_lazy_propagate_res_matches_the_list():
    # A batch keeps propagate_res as its result arrays until it is read; the list it then
    # builds is the reference's: (taylor_outcome or int, min_h, max_h, nsteps) per lane,
    # with Python floats and ints, and copies and pickles keep it.
    from heyoka_amd.integrator import _PropRes, _outcome

    r = {"outcome": np.array([-4294967299, -4294967297, -1, 7], dtype=np.int64),
         "min_h": np.array([0.5, 0.25, np.inf, 0.0]), "max_h": np.array([1.0, 2.0, 0.0, 0.0]),
         "nsteps": np.array([3, 4, 0, 9], dtype=np.uint64)}
    eager = [(_outcome(o), float(a), float(b), int(c))
             for o, a, b, c in zip(r["outcome"], r["min_h"], r["max_h"], r["nsteps"])]
    lazy = _PropRes.of(r)
    assert lazy.tolist() == eager
    assert [type(x) for x in lazy.tolist()[0]] == [type(x) for x in eager[0]]
    assert lazy.tolist()[0][0] is hy.taylor_outcome.time_limit and lazy.tolist()[2][0] == -1
    assert _PropRes.of(r, slice(1, 3)).tolist() == eager[1:3]
    assert copy.deepcopy(lazy).tolist() == eager and pickle.loads(pickle.dumps(lazy)).tolist() == eager
