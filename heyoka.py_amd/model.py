"""Dynamical models used by the batch/ensemble workloads.

``nbody``, ``nbody_energy``, ``pendulum``, ``pendulum_energy``, ``cr3bp`` and
``cr3bp_jacobi`` keep the names, argument names, defaults and state layouts of
heyoka.py's models (``heyoka/expose_models.cpp:237-283, 396-400``; structure
tests ``heyoka/_test_model.py:153-278``). ``cr3bp_planar`` (velocity form with
the mass ratio as a runtime parameter) and ``j2`` (Cartesian Kepler problem with
the J2 zonal term) are hand-written here: the reference has no model in that form.
"""

from __future__ import annotations

import numbers

from .expr import expression, make_vars, par, sqrt, sin, cos, square, sum as _sum, _ex

__all__ = [
    "nbody", "nbody_energy", "pendulum", "pendulum_energy", "cr3bp", "cr3bp_jacobi",
    "cr3bp_planar", "cr3bp_planar_jacobi", "j2", "j2_energy", "j2_hz",
]


def _nbody_vars(n):
    out = []
    for i in range(n):
        out.append(make_vars(f"x_{i}", f"y_{i}", f"z_{i}", f"vx_{i}", f"vy_{i}", f"vz_{i}"))
    return out


def _masses(n, masses):
    if masses is None:
        return [1.0] * n
    ms = list(masses)
    if len(ms) != n:
        raise ValueError(f"Inconsistent sizes detected in the specification of an N-body system: the number of "
                         f"particles is {n}, but the number of masses is {len(ms)}")
    return ms


def nbody(n, Gconst=1.0, masses=None):
    """Newtonian N-body problem; state per body ``[x, y, z, vx, vy, vz]``.

    For each pair i < j: r_ij = x_j - x_i, ``rm3 = (|r_ij|^2)^(-3/2)``, and the
    accelerations receive ``G m_j r_ij rm3`` (body i) and ``-G m_i r_ij rm3``
    (body j). Terms with a zero mass are omitted, so massless bodies feel the
    others but do not perturb them.
    """
    if not isinstance(n, numbers.Integral) or n < 2:
        raise ValueError("An N-body system must contain at least 2 bodies")
    ms = _masses(n, masses)
    G = _ex(Gconst)
    bodies = _nbody_vars(n)
    acc = [[[], [], []] for _ in range(n)]
    for i in range(n):
        for j in range(i + 1, n):
            mi_zero = isinstance(ms[i], numbers.Real) and ms[i] == 0
            mj_zero = isinstance(ms[j], numbers.Real) and ms[j] == 0
            if mi_zero and mj_zero:
                continue
            d = [bodies[j][q] - bodies[i][q] for q in range(3)]
            rm3 = (square(d[0]) + square(d[1]) + square(d[2])) ** -1.5
            for q in range(3):
                tmp = d[q] * rm3
                if not mj_zero:
                    acc[i][q].append((G * _ex(ms[j])) * tmp)
                if not mi_zero:
                    acc[j][q].append((-(G * _ex(ms[i]))) * tmp)
    sys = []
    for i in range(n):
        x, y, z, vx, vy, vz = bodies[i]
        sys += [(x, vx), (y, vy), (z, vz)]
        sys += [(v, _sum(acc[i][q])) for q, v in enumerate((vx, vy, vz))]
    return sys


def nbody_energy(n, Gconst=1.0, masses=None):
    """Total energy: sum_i m_i |v_i|^2 / 2 - sum_{i<j} G m_i m_j / |r_ij|."""
    ms = _masses(n, masses)
    G = _ex(Gconst)
    bodies = _nbody_vars(n)
    terms = []
    for i in range(n):
        if isinstance(ms[i], numbers.Real) and ms[i] == 0:
            continue
        x, y, z, vx, vy, vz = bodies[i]
        terms.append((0.5 * _ex(ms[i])) * (square(vx) + square(vy) + square(vz)))
    for i in range(n):
        for j in range(i + 1, n):
            if (isinstance(ms[i], numbers.Real) and ms[i] == 0) or (isinstance(ms[j], numbers.Real) and ms[j] == 0):
                continue
            d = [bodies[j][q] - bodies[i][q] for q in range(3)]
            terms.append(-(G * _ex(ms[i]) * _ex(ms[j])) / sqrt(square(d[0]) + square(d[1]) + square(d[2])))
    return _sum(terms)


def pendulum(gconst=1.0, length=1.0):
    """Simple pendulum ``x' = v, v' = -(g/l) sin(x)``."""
    x, v = make_vars("x", "v")
    k = _ex(gconst) / _ex(length)
    acc = -sin(x) if (k.is_number() and k.value == 1.0) else (-k) * sin(x)
    return [(x, v), (v, acc)]


def pendulum_energy(gconst=1.0, length=1.0):
    """Energy ``l^2 v^2 / 2 + g l (1 - cos x)``."""
    x, v = make_vars("x", "v")
    g, l = _ex(gconst), _ex(length)
    return (0.5 * l * l) * v ** 2 + (g * l) * (1.0 - cos(x))


def cr3bp(mu=1e-3):
    """Spatial circular restricted three-body problem in canonical variables.

    State ``(x, y, z, px, py, pz)``; primaries of mass 1-mu at (-mu, 0, 0) and mu
    at (1-mu, 0, 0); ``x' = px + y``. ``mu`` may be a number or an expression
    (e.g. ``par[0]``).
    """
    mu = _ex(mu)
    x, y, z, px, py, pz = make_vars("x", "y", "z", "px", "py", "pz")
    r1m3 = (square(x + mu) + square(y) + square(z)) ** -1.5
    r2m3 = (square(x - (1.0 - mu)) + square(y) + square(z)) ** -1.5
    om = 1.0 - mu
    return [
        (x, px + y),
        (y, py - x),
        (z, pz),
        (px, py - om * (x + mu) * r1m3 - mu * (x - om) * r2m3),
        (py, -px - om * y * r1m3 - mu * y * r2m3),
        (pz, -om * z * r1m3 - mu * z * r2m3),
    ]


def cr3bp_jacobi(mu=1e-3):
    """Hamiltonian of ``cr3bp`` (conserved): |p|^2/2 + y px - x py - (1-mu)/r1 - mu/r2."""
    mu = _ex(mu)
    x, y, z, px, py, pz = make_vars("x", "y", "z", "px", "py", "pz")
    r1 = sqrt(square(x + mu) + square(y) + square(z))
    r2 = sqrt(square(x - (1.0 - mu)) + square(y) + square(z))
    return (0.5 * (square(px) + square(py) + square(pz)) + y * px - x * py
            - (1.0 - mu) / r1 - mu / r2)


def cr3bp_planar(mu=None):
    """Planar CR3BP in the rotating frame, velocity form (hand-written; not a reference model).

    State ``(x, y, vx, vy)``, primaries at (-mu, 0) and (1-mu, 0), ``mu = par[0]``
    by default:
        x'' =  2 y' + x - (1-mu)(x+mu)/r1^3 - mu(x-1+mu)/r2^3
        y'' = -2 x' + y - (1-mu) y/r1^3    - mu y/r2^3
    """
    mu = par[0] if mu is None else _ex(mu)
    x, y, vx, vy = make_vars("x", "y", "vx", "vy")
    om = 1.0 - mu
    r1m3 = (square(x + mu) + square(y)) ** -1.5
    r2m3 = (square(x - om) + square(y)) ** -1.5
    return [
        (x, vx),
        (y, vy),
        (vx, 2.0 * vy + x - om * (x + mu) * r1m3 - mu * (x - om) * r2m3),
        (vy, -2.0 * vx + y - om * y * r1m3 - mu * y * r2m3),
    ]


def cr3bp_planar_jacobi(mu=None):
    """Jacobi constant C = x^2 + y^2 + 2(1-mu)/r1 + 2 mu/r2 - (vx^2 + vy^2)."""
    mu = par[0] if mu is None else _ex(mu)
    x, y, vx, vy = make_vars("x", "y", "vx", "vy")
    r1 = sqrt(square(x + mu) + square(y))
    r2 = sqrt(square(x - (1.0 - mu)) + square(y))
    return square(x) + square(y) + 2.0 * (1.0 - mu) / r1 + 2.0 * mu / r2 - (square(vx) + square(vy))


def j2(mu=1.0, J2=1.0826e-3, R=1.0):
    """Kepler problem with the J2 zonal harmonic, Cartesian (hand-written; not a reference model).

    a_{x,y} = -mu r_{x,y}/r^3 [1 + 1.5 J2 (R/r)^2 (1 - 5 z^2/r^2)]
    a_z     = -mu z/r^3       [1 + 1.5 J2 (R/r)^2 (3 - 5 z^2/r^2)]
    """
    x, y, z, vx, vy, vz = make_vars("x", "y", "z", "vx", "vy", "vz")
    mu, J2, R = _ex(mu), _ex(J2), _ex(R)
    r2 = square(x) + square(y) + square(z)
    rm3 = r2 ** -1.5
    rm2 = r2 ** -1.0
    k = (1.5 * J2 * R * R) * rm2
    z2r2 = square(z) * rm2
    fxy = -mu * rm3 * (1.0 + k * (1.0 - 5.0 * z2r2))
    fz = -mu * rm3 * (1.0 + k * (3.0 - 5.0 * z2r2))
    return [(x, vx), (y, vy), (z, vz), (vx, fxy * x), (vy, fxy * y), (vz, fz * z)]


def j2_energy(mu=1.0, J2=1.0826e-3, R=1.0):
    """E = |v|^2/2 - mu/r + mu J2 R^2 (3 z^2/r^2 - 1) / (2 r^3)."""
    x, y, z, vx, vy, vz = make_vars("x", "y", "z", "vx", "vy", "vz")
    mu, J2, R = _ex(mu), _ex(J2), _ex(R)
    r = sqrt(square(x) + square(y) + square(z))
    return (0.5 * (square(vx) + square(vy) + square(vz)) - mu / r
            + (0.5 * mu * J2 * R * R) * (3.0 * square(z) / square(r) - 1.0) / (r * r * r))


def j2_hz(**_):
    """Polar component of the angular momentum, conserved by the axisymmetric J2 field."""
    x, y, z, vx, vy, vz = make_vars("x", "y", "z", "vx", "vy", "vz")
    return x * vy - y * vx
