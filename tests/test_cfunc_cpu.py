"""Compiled functions: the Python surface of heyoka.py's cfunc (heyoka/_test_cfunc.py:14-112,
the assertions that do not involve LLVM) and the generated HIP function compiled for the
CPU with g++ (qualifiers defined away), checked against a numpy evaluation."""

import ctypes as ct
import pickle
import subprocess
from copy import copy, deepcopy

import numpy as np
import pytest

import heyoka_amd as hy
from heyoka_amd.expr import evaluate


def test_cfunc_surface():
    with pytest.raises(ValueError):
        hy.cfunc([], [])
    x, y, z, s = hy.make_vars("x", "y", "z", "s")
    cf = hy.cfunc([y * (x + z)], [x, y, z])
    assert cf.vars == [x, y, z] and cf.fn == [y * (x + z)] and len(cf.dc) == 6
    assert deepcopy(cf).vars == [x, y, z] and deepcopy(cf).dc == cf.dc and copy(cf).fn == cf.fn
    assert pickle.loads(pickle.dumps(cf)).dc == cf.dc
    assert hy.cfunc([y * (x + z), x], vars=[y, z, x]).nparams == 0
    assert hy.cfunc([y * (x + z), hy.par[0]], vars=[y, z, x]).nparams == 1
    assert hy.cfunc([y * (x + z) - hy.par[89], hy.par[0]], vars=[y, z, x]).nparams == 90
    assert not hy.cfunc([y * (x + z), x], vars=[y, z, x]).is_time_dependent
    assert hy.cfunc([y * (x + z) + hy.time, x], vars=[y, z, x]).is_time_dependent
    assert hy.cfunc([y * (x + z), x + hy.time], vars=[y, z, x]).is_time_dependent
    assert hy.cfunc([y * (x + z), x + hy.time], vars=[y, z, x, s]).vars == [y, z, x, s]
    assert hy.cfunc([y * (x + z), x + hy.time], vars=[s, y, z, x]).vars == [s, y, z, x]
    with pytest.raises(ValueError):
        hy.cfunc([y * (x + z)], [x, y])  # z is not among the variables


HARNESS = r"""
#include <cmath>
#include <cstdint>
#include <math.h>
#define __device__
#define __forceinline__ inline
using std::sqrt; using std::pow; using std::exp; using std::log; using std::fabs;
#include "prelude.inc"
extern "C" void eval_cpu(const double *x, const double *pv, double t0, double *out)
{
    double xx[HYX_N], pp[HYX_NPAR];
    for (int i = 0; i < HYX_N; ++i) xx[i] = x[i];
    for (int i = 0; i < HYX_NPAR; ++i) pp[i] = pv[i];
    hyx_cfunc_eval(xx, pp, t0, out, 1);
}
"""


def test_generated_cfunc_matches_numpy(tmp_path):
    x, y, z = hy.make_vars("x", "y", "z")
    fn = [y * (x + z), hy.sin(x) * hy.exp(-y) + hy.par[1] * hy.time, (x ** 2 + y ** 2 + z ** 2) ** -1.5,
          hy.sqrt(z) - hy.log(y) / x, hy.cos(z) ** 2, hy.par[0]]
    cf = hy.cfunc(fn, [x, y, z])
    prelude = cf.source().split("// Compiled functions (heyoka's cfunc")[0]
    (tmp_path / "prelude.inc").write_text(prelude)
    (tmp_path / "h.cpp").write_text(HARNESS)
    so = tmp_path / "h.so"
    subprocess.run(["g++", "-O1", "-std=c++17", "-ffp-contract=off", "-shared", "-fPIC", "-o", str(so),
                    str(tmp_path / "h.cpp")], check=True, capture_output=True)
    L = ct.CDLL(str(so))
    P = ct.POINTER(ct.c_double)
    L.eval_cpu.argtypes = [P, P, ct.c_double, P]
    rng = np.random.default_rng(5)
    ptr = lambda a: a.ctypes.data_as(P)
    for _ in range(20):
        xv = np.array([rng.uniform(0.5, 2), rng.uniform(0.5, 2), rng.uniform(0.5, 2)])
        pv = rng.uniform(-1, 1, 2)
        t = rng.uniform(0, 5)
        out = np.zeros(len(fn))
        L.eval_cpu(ptr(xv), ptr(pv), t, ptr(out))
        env = {"x": xv[0], "y": xv[1], "z": xv[2]}
        ref = np.array([float(evaluate(f, env, pv, t)) for f in fn])
        np.testing.assert_allclose(out, ref, rtol=1e-14, atol=1e-15)
