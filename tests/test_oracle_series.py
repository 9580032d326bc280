"""Each recurrence of the oracle against closed-form Taylor series (mpmath).

A state variable y with y' = f(...) exposes f's normalised derivatives through
its Taylor coefficients: y^[k+1] = f^[k] / (k+1). Driving f with arguments whose
series are known and not polynomial (a = a0 e^t from a' = a) or polynomial
(c = c0 + t from c' = 1) and comparing with mpmath.taylor of the closed form
checks every convolution term of every op: a dropped term, a wrong index or a
wrong sign moves the result far outside the tolerance.
"""

import mpmath as mp
import numpy as np
import pytest

import heyoka_amd as hy
import pyoracle as O

A0, C0 = 1.3, 0.7
P = 20

a, c = hy.make_vars("a", "c")

# name -> (expression of a, c and time; closed form in t)
CASES = {
    "mul": (a * c, lambda t: A0 * mp.e**t * (C0 + t)),
    "div_ac": (a / c, lambda t: A0 * mp.e**t / (C0 + t)),
    "div_ca": (c / a, lambda t: (C0 + t) / (A0 * mp.e**t)),
    "div_const": (a / 3.0, lambda t: A0 * mp.e**t / 3),
    "sqrt_a": (hy.sqrt(a), lambda t: mp.sqrt(A0 * mp.e**t)),
    "sqrt_c": (hy.sqrt(c), lambda t: mp.sqrt(C0 + t)),
    "pow_m15": (a ** -1.5, lambda t: (A0 * mp.e**t) ** mp.mpf(-1.5)),
    "pow_c25": (c ** 2.5, lambda t: (C0 + t) ** mp.mpf(2.5)),
    "pow_cm15": (c ** -1.5, lambda t: (C0 + t) ** mp.mpf(-1.5)),
    "square": (a ** 2, lambda t: (A0 * mp.e**t) ** 2),
    "square_c": (hy.square(c * a), lambda t: ((C0 + t) * A0 * mp.e**t) ** 2),
    "exp_c": (hy.exp(c), lambda t: mp.e ** (C0 + t)),
    "exp_a": (hy.exp(a), lambda t: mp.e ** (A0 * mp.e**t)),
    "log_a": (hy.log(a), lambda t: mp.log(A0 * mp.e**t)),
    "log_c": (hy.log(c), lambda t: mp.log(C0 + t)),
    "sin_a": (hy.sin(a), lambda t: mp.sin(A0 * mp.e**t)),
    "cos_a": (hy.cos(a), lambda t: mp.cos(A0 * mp.e**t)),
    "sin_c": (hy.sin(c), lambda t: mp.sin(C0 + t)),
    "sum": (a + c + a * c, lambda t: A0 * mp.e**t + (C0 + t) + A0 * mp.e**t * (C0 + t)),
    "sub_neg": (-(a - 2.0 * c), lambda t: -(A0 * mp.e**t - 2 * (C0 + t))),
    "time_mul": (hy.time * a, lambda t: t * A0 * mp.e**t),
    "cos_time": (hy.cos(hy.time), lambda t: mp.cos(t)),
    "par_mul": (hy.par[0] * a, lambda t: mp.mpf(0.25) * A0 * mp.e**t),
}


def _system():
    names = list(CASES)
    ys = hy.make_vars(*[f"y_{n}" for n in names])
    sys = [(a, a), (c, hy.expression(1.0))] + [(y, CASES[n][0]) for y, n in zip(ys, names)]
    return names, hy.taylor_decompose(sys)


@pytest.mark.parametrize("f128", [False, True])
def test_recurrences_match_closed_forms(f128):
    names, dec = _system()
    n = dec.n_state
    st = np.zeros((n, 1))
    st[0, 0], st[1, 0] = A0, C0
    r = O.step_batch(dec, P, st, pars=[[0.25]], want_tc=True, f128=f128)
    tc = r["tc"][:, :, 0]
    mp.mp.prec = 160
    rtol = 1e-12 if not f128 else 1e-15  # f128 coefficients are rounded to double on output
    for j, nm in enumerate(names):
        coeffs = mp.taylor(CASES[nm][1], 0, P - 1)
        got = [(k + 1) * tc[2 + j, k + 1] for k in range(P)]
        for k in range(P):
            want = float(coeffs[k])
            scale = max(abs(float(x)) for x in coeffs[: k + 1]) + 1e-300
            assert abs(got[k] - want) <= rtol * scale * (k + 1), (nm, k, got[k], want)


def test_state_variable_orders():
    # x' = a with a = a0 e^t: x^[k] = a0 / k! for k >= 1.
    x = hy.make_vars("x")
    dec = hy.taylor_decompose([(a, a), (x, a)])
    r = O.step_batch(dec, P, [[A0], [0.0]], want_tc=True)
    for k in range(1, P + 1):
        assert abs(r["tc"][1, k, 0] - A0 / float(mp.factorial(k))) <= 1e-15 * A0 / float(mp.factorial(k))


def test_kepler_two_body_against_kepler_equation():
    # nbody(2) with one massless body is a Kepler orbit: the closed-form solution
    # pins the whole propagate loop, including relative-error step control.
    from heyoka_amd import model

    sys = model.nbody(2, masses=[1.0, 0.0])
    dec = hy.taylor_decompose(sys)
    e, aa = 0.3, 1.7
    rp = aa * (1 - e)
    vp = np.sqrt((1 + e) / rp)  # mu = 1
    st = np.array([0, 0, 0, 0, 0, 0, rp, 0, 0, 0, vp, 0], dtype=float).reshape(12, 1)
    T = 2 * np.pi * aa ** 1.5
    r = O.propagate_until(dec, 20, st, 3.25 * T)
    # Kepler's equation for the position at t.
    M = 2 * np.pi * 3.25
    E = M
    for _ in range(50):
        E = E - (E - e * np.sin(E) - M) / (1 - e * np.cos(E))
    xk = aa * (np.cos(E) - e)
    yk = aa * np.sqrt(1 - e * e) * np.sin(E)
    np.testing.assert_allclose(r["state"][6:8, 0], [xk, yk], rtol=0, atol=1e-12)
    assert r["outcome"][0] == O.TIME_LIMIT


def test_f128_oracle_agrees_with_f64():
    x, v = hy.make_vars("x", "v")
    dec = hy.taylor_decompose([(x, v), (v, -9.8 * hy.sin(x))])
    r64 = O.propagate_until(dec, 20, [[0.05], [0.025]], 10.0)
    r128 = O.propagate_until(dec, 38, [[0.05], [0.025]], 10.0, f128=True)
    np.testing.assert_allclose(r64["state"], r128["state"], rtol=0, atol=1e-14)
