"""J2 + first-order variational equations on the GPU.

The state transition matrix (the 36 variational states) is checked against
central finite differences of the plain J2 system, against symplecticity
(Phi^T J Phi = J: the J2 field is conservative), and the J2 energy and polar
angular momentum are checked for conservation.
"""

import numpy as np
import pytest

import heyoka_amd as hy
from heyoka_amd import workloads as W

pytestmark = pytest.mark.gpu


def _phi(ta):
    # Layout: for each component i, the derivatives with respect to the 6 initial conditions.
    return ta.state[6:42].reshape(6, 6, -1)


def test_stm_vs_finite_differences_and_symplecticity():
    print("[gpu] j2var STM vs central finite differences", flush=True)
    nb = 8
    wl = W.j2var(nb)
    T = 2 * np.pi * np.sqrt(1.3 ** 3)
    tf = 2.0 * T
    ta = hy.taylor_adaptive_batch(wl.sys, wl.state, tol=1e-15)
    assert ta.is_variational and ta.get_vslice(1) == slice(6, 42)
    ta.propagate_until(tf)
    phi = _phi(ta)
    x0 = wl.state[:6]
    cols = []
    deltas = []
    for j in range(6):
        d = 1e-6 * np.maximum(1.0, np.abs(x0[j]))
        deltas.append(d)
        for sgn in (1.0, -1.0):
            xp = x0.copy()
            xp[j] += sgn * d
            cols.append(xp)
    big = np.concatenate(cols, axis=1)  # [6, 12 * nb]
    tb = hy.taylor_adaptive_batch(hy.model.j2(), big, tol=1e-15)
    tb.propagate_until(tf)
    out = tb.state.reshape(6, 12, nb)
    for j in range(6):
        fd = (out[:, 2 * j, :] - out[:, 2 * j + 1, :]) / (2 * deltas[j])
        err = np.max(np.abs(fd - phi[:, j, :]) / (np.max(np.abs(phi), axis=(0, 1)) + 1e-300))
        assert err < 1e-6, (j, err)
    Jm = np.block([[np.zeros((3, 3)), np.eye(3)], [-np.eye(3), np.zeros((3, 3))]])
    for b in range(nb):
        P = phi[:, :, b]
        assert np.max(np.abs(P.T @ Jm @ P - Jm)) < 1e-9


def test_j2_invariants():
    print("[gpu] j2 energy and h_z drift over 20 periods", flush=True)
    wl = W.j2var(64)
    x = wl.state[:6].copy()
    ta = hy.taylor_adaptive_batch(hy.model.j2(), x, tol=1e-15)
    mu, J2, R = 1.0, 1.0826e-3, 1.0

    def energy(s):
        r = np.sqrt(s[0] ** 2 + s[1] ** 2 + s[2] ** 2)
        return 0.5 * (s[3] ** 2 + s[4] ** 2 + s[5] ** 2) - mu / r + 0.5 * mu * J2 * R * R * (3 * s[2] ** 2 / r ** 2 - 1) / r ** 3

    def hz(s):
        return s[0] * s[4] - s[1] * s[3]

    e0, h0 = energy(x), hz(x)
    ta.propagate_until(20 * 2 * np.pi * 1.5 ** 1.5)
    de = np.max(np.abs(energy(ta.state) - e0) / np.abs(e0))
    dh = np.max(np.abs(hz(ta.state) - h0) / np.maximum(np.abs(h0), 1e-3))
    print(f"[gpu] j2: max relative energy drift {de:.2e}, h_z drift {dh:.2e}", flush=True)
    assert de < 1e-12 and dh < 1e-12
