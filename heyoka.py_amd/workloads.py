"""Seeded synthetic inputs of the batch/ensemble workloads.

Each generator returns the ODE system, the state ``[n, N]`` (instance index
fastest, the reference's batch layout), the parameters ``[n_par, N]``, the
final times ``[N]`` and the tolerance:

- pendulum: x'' = -sin x, N = 4, tol = eps, x0 ~ U(-1, 1), v0 ~ U(-0.5, 0.5)
  (default_rng(0)), propagate_until t = 100;
- nbody6: 6 bodies in 3D (G = 1), masses 1 and U(1e-3, 1) (default_rng(1)),
  near-circular coplanar orbits at radii geomspace(1, 30, 5) with speed
  sqrt((1 + m_i) / r_i) * (1 + U(-1e-3, 1e-3)) and random phases, moved to the
  barycentre; N copies with positions perturbed by N(0, 1e-6) (default_rng(2));
  tol = 1e-15; t = 1000;
- cr3bp: planar CR3BP with mu = par[0] ~ U(0.001, 0.05) (default_rng(3)); half
  of the instances near L4, half near L5, all four components offset by
  U(-0.05, 0.05); horizons 10**U(1, 4); tol = 1e-15;
- j2var: J2 Kepler problem with first-order variational equations (42 states);
  a ~ U(1.05, 1.5), e ~ U(0, 0.1), i ~ U(0, pi), Omega, omega, M ~ U(0, 2 pi)
  (default_rng(4)); 200 orbital periods; tol = 1e-15.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import model
from .var_ode_sys import var_ode_sys, var_args


@dataclass
class Workload:
    name: str
    sys: object
    state: np.ndarray  # [n, N]
    pars: np.ndarray | None  # [n_par, N]
    t_final: np.ndarray  # [N]
    tol: float
    masses: list | None = None


def pendulum(n=4, seed=0):
    rng = np.random.default_rng(seed)
    x0 = rng.uniform(-1, 1, n)
    v0 = rng.uniform(-0.5, 0.5, n)
    return Workload("pendulum", model.pendulum(), np.stack([x0, v0]), None, np.full(n, 100.0),
                    float(np.finfo(float).eps))


def nbody6_base():
    rng = np.random.default_rng(1)
    m = np.concatenate([[1.0], rng.uniform(1e-3, 1.0, 5)])
    radii = np.geomspace(1.0, 30.0, 5)
    pos = np.zeros((6, 3))
    vel = np.zeros((6, 3))
    for i in range(1, 6):
        th = rng.uniform(0, 2 * np.pi)
        r = radii[i - 1]
        sp = np.sqrt((1.0 + m[i]) / r) * (1.0 + rng.uniform(-1e-3, 1e-3))
        pos[i] = [r * np.cos(th), r * np.sin(th), 0.0]
        vel[i] = [-sp * np.sin(th), sp * np.cos(th), 0.0]
    M = m.sum()
    pos -= (m[:, None] * pos).sum(axis=0) / M
    vel -= (m[:, None] * vel).sum(axis=0) / M
    return m, pos, vel


def nbody6(n=65536, t_final=1000.0, seed=2):
    m, pos, vel = nbody6_base()
    rng = np.random.default_rng(seed)
    st = np.zeros((36, n))
    for i in range(6):
        for q in range(3):
            st[6 * i + q] = pos[i, q] + rng.normal(0.0, 1e-6, n)
            st[6 * i + 3 + q] = vel[i, q]
    return Workload("nbody6", model.nbody(6, masses=list(m)), st, None, np.full(n, t_final), 1e-15, masses=list(m))


def cr3bp(n=131072, log_t=(1.0, 4.0), seed=3):
    rng = np.random.default_rng(seed)
    mu = rng.uniform(0.001, 0.05, n)
    sign = np.where(np.arange(n) % 2 == 0, 1.0, -1.0)
    st = np.stack([0.5 - mu, sign * np.sqrt(3.0) / 2, np.zeros(n), np.zeros(n)])
    st += rng.uniform(-0.05, 0.05, (4, n))
    tf = 10.0 ** rng.uniform(log_t[0], log_t[1], n)
    return Workload("cr3bp", model.cr3bp_planar(), st, mu[None, :], tf, 1e-15)


def kepler_to_cartesian(a, e, inc, Om, om, M, mu=1.0):
    E = M.copy()
    for _ in range(60):
        E = E - (E - e * np.sin(E) - M) / (1 - e * np.cos(E))
    cE, sE = np.cos(E), np.sin(E)
    b = a * np.sqrt(1 - e * e)
    xp, yp = a * (cE - e), b * sE
    n_mean = np.sqrt(mu / a ** 3)
    rr = a * (1 - e * cE)
    vxp = -a * n_mean * sE / (1 - e * cE)
    vyp = b * n_mean * cE / (1 - e * cE)
    del rr
    cO, sO, co, so, ci, si = np.cos(Om), np.sin(Om), np.cos(om), np.sin(om), np.cos(inc), np.sin(inc)
    R11 = cO * co - sO * so * ci
    R12 = -cO * so - sO * co * ci
    R21 = sO * co + cO * so * ci
    R22 = -sO * so + cO * co * ci
    R31 = so * si
    R32 = co * si
    pos = np.stack([R11 * xp + R12 * yp, R21 * xp + R22 * yp, R31 * xp + R32 * yp])
    vel = np.stack([R11 * vxp + R12 * vyp, R21 * vxp + R22 * vyp, R31 * vxp + R32 * vyp])
    return pos, vel


def j2var(n=32768, periods=200.0, seed=4):
    rng = np.random.default_rng(seed)
    a = rng.uniform(1.05, 1.5, n)
    e = rng.uniform(0.0, 0.1, n)
    inc = rng.uniform(0.0, np.pi, n)
    Om, om, M = (rng.uniform(0.0, 2 * np.pi, n) for _ in range(3))
    pos, vel = kepler_to_cartesian(a, e, inc, Om, om, M)
    vs = var_ode_sys(model.j2(), var_args.vars, order=1)
    st = np.concatenate([pos, vel, np.repeat(np.array(vs.initial_derivatives())[:, None], n, axis=1)])
    tf = periods * 2 * np.pi * np.sqrt(a ** 3)
    return Workload("j2var", vs, st, None, tf, 1e-15)
