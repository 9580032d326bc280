"""The event-detection oracle (oracle/events.py) pinned on the CPU.

- Root isolation and refinement against polynomials with planted roots
  (brute force: the roots are known exactly) and against numpy's companion-matrix
  roots.
- The pendulum notebook goldens (doc/notebooks/Event detection.ipynb, cells 9, 11,
  17, 21): event times, directions, processing order, and the step count and
  step sizes of a propagation with events (which fixes that the event functions'
  coefficients join the step-size norms).
- The closed form: the pendulum's half period from the complete elliptic
  integral, T / 2 = 2 K(sin^2(theta0 / 2)) / sqrt(g / l).
- The reference's batch event tests (heyoka/test.py:693-1000,
  heyoka/_test_batch_integrator.py:555-600): trigger counts, callback order and
  terminal outcome codes.

Tolerances: the roots of v = 0 are well conditioned (|dv/dt| ~ 0.49), so times
agree to a few ulps; those of v^2 - 1e-12 sit where |dg/dt| ~ 1e-6 while the
polynomial's terms are ~1e-2, so rounding moves them by up to ~1e-12.
"""

import numpy as np
import pytest
from scipy.special import ellipk

import heyoka_amd as hy
import events as OE

x, v = hy.make_vars("x", "v")
PEND = [(x, v), (v, -9.8 * hy.sin(x))]


@pytest.mark.parametrize("seed", range(6))
def test_planted_roots(seed):
    rng = np.random.default_rng(seed)
    k = int(rng.integers(1, 6))
    roots = np.sort(rng.uniform(0.02, 0.98, k))
    if np.min(np.diff(roots), initial=1.0) < 0.01:
        roots = np.linspace(0.1, 0.9, k)
    outside = rng.uniform(1.5, 3.0, int(rng.integers(0, 8)))
    c = np.polynomial.polynomial.polyfromroots(np.concatenate([roots, outside, -outside]))
    found = OE.poly_roots01(c * rng.uniform(0.5, 2.0))
    np.testing.assert_allclose(found, roots, rtol=0, atol=1e-12)
    # numpy's companion-matrix roots agree on the real roots in [0, 1).
    r = np.roots(c[::-1])
    r = np.sort(r[(np.abs(r.imag) < 1e-9) & (r.real >= 0) & (r.real < 1)].real)
    np.testing.assert_allclose(found, r, rtol=0, atol=1e-9)


def test_roots_edge_cases():
    # No root in [0, 1); a root exactly at 0 is reported; a root at 1 is not (it
    # belongs to the next step).
    assert OE.poly_roots01([1.0, 0.5, 0.25]) == []
    assert OE.poly_roots01([0.0, 1.0, -3.0]) == pytest.approx([0.0, 1.0 / 3.0], abs=1e-15)
    assert OE.poly_roots01([-1.0, 1.0]) == []
    assert OE.poly_roots01([0.0] * 5) == []
    # Close roots are separated: (s - 0.5)(s - 0.5 - 1e-9).
    c = np.polynomial.polynomial.polyfromroots([0.5, 0.5 + 1e-9])
    np.testing.assert_allclose(OE.poly_roots01(c), [0.5, 0.5 + 1e-9], atol=1e-13)
    # Backward steps: tau runs from 0 down to h; d_sgn is the sign of dg/dt.
    assert OE.event_roots([0.5, 1.0], -1.0) == [(-0.5, 1)]
    assert OE.event_roots([-0.5, -1.0], -1.0) == [(-0.5, -1)]
    assert OE.event_roots([0.5, 1.0], 1.0) == []


def _nt(*evs):
    return [(e, int(d)) for e, d in evs]


def test_notebook_zero_velocity_times(goldens):
    g = goldens["ev_zero_vel_times"]
    r = OE.propagate(hy, PEND, np.array(g["ic"])[:, None], g["t_end"], nt_events=_nt((v, 0)))
    times = [t for _, t, _ in r["nt"][0]]
    np.testing.assert_allclose(times, g["times"], rtol=4e-16, atol=0)
    # The pendulum is at rest at its amplitude: d_sgn alternates, starting positive.
    assert [d for _, _, d in r["nt"][0]] == [1, -1, 1, -1, 1]
    gp = goldens["ev_positive_times"]
    r = OE.propagate(hy, PEND, np.array(gp["ic"])[:, None], gp["t_end"], nt_events=_nt((v, 1)))
    np.testing.assert_allclose([t for _, t, _ in r["nt"][0]], gp["times"], rtol=4e-16, atol=0)


def test_notebook_two_events(goldens):
    g = goldens["ev_two_events"]
    r = OE.propagate(hy, PEND, np.array(g["ic"])[:, None], g["t_end"], nt_events=_nt((v, 0), (v * v - 1e-12, 0)))
    got = r["nt"][0]
    assert [i for i, _, _ in got] == [i for i, _ in g["events"]]
    for (i, t, _), (gi, gt) in zip(got, g["events"]):
        assert abs(t - gt) <= (4e-16 * abs(gt) if i == 0 else 2e-12), (i, t, gt)
    assert r["outcome"][0] == OE.TIME_LIMIT and r["nsteps"][0] == g["nsteps"]
    np.testing.assert_allclose([r["min_h"][0], r["max_h"][0]], [g["min_h"], g["max_h"]], rtol=1e-15)


def test_half_period_closed_form():
    # T = 4 K(m) / sqrt(9.8), m = sin^2(theta0 / 2): zeros of v every T / 2.
    th0 = 0.05
    half = 2.0 * ellipk(np.sin(th0 / 2) ** 2) / np.sqrt(9.8)
    r = OE.propagate(hy, PEND, np.array([[-th0], [0.0]]), 5.0, nt_events=_nt((v, 0)))
    times = np.array([t for _, t, _ in r["nt"][0]])
    np.testing.assert_allclose(times, half * np.arange(len(times)), rtol=0, atol=2e-14)


def test_reference_batch_counts():
    # heyoka/test.py:693-760: two lanes, nt events v*v - 1e-10 and v over [0, 4]:
    # 12 triggers each, v*v - 1e-10 / v / v*v - 1e-10 around every zero of v.
    ic = np.array([[0.0, 0.001], [0.25, 0.2501]])
    r = OE.propagate(hy, PEND, ic, 4.0, nt_events=_nt((v * v - 1e-10, 0), (v, 0)))
    for b in range(2):
        idx = [i for i, _, _ in r["nt"][b]]
        assert idx.count(0) == 8 and idx.count(1) == 4
        assert idx == [0, 1, 0] * 4
        ts = [t for _, t, _ in r["nt"][b]]
        assert all(np.diff(ts) > 0)
    assert list(r["outcome"]) == [OE.TIME_LIMIT] * 2


def test_reference_terminal_stop_and_cooldown():
    # heyoka/_test_batch_integrator.py:570-600: a terminal event v without callback
    # stops every lane with outcome -1 at the first zero of v, with a cooldown set.
    ic = np.array([[0.0, 0.001], [0.25, 0.2501]])
    r = OE.propagate(hy, PEND, ic, 1e9, t_events=[(v, 0, -1.0, True)], nt_events=_nt((v * v - 1e-6, 0)))
    assert list(r["outcome"]) == [-1, -1]
    assert all(abs(r["state"][1, b]) < 1e-15 for b in range(2))
    assert all(r["cooldowns"][b][0] is not None and r["cooldowns"][b][0][2] > 0 for b in range(2))
    # Continuing from the trigger: the cooldown keeps v = 0 from firing again at once;
    # the next trigger is half a period later.
    t1 = r["t_hi"].copy()
    r2 = OE.propagate(hy, PEND, r["state"], 1e9, t_events=[(v, 0, -1.0, True)], t0=t1, cooldowns=r["cooldowns"])
    assert list(r2["outcome"]) == [-1, -1]
    assert np.all(r2["t_hi"] - t1 > 0.9)
    # Without the cooldown the same restart fires immediately (tau ~ 0).
    r3 = OE.propagate(hy, PEND, r["state"], 1e9, t_events=[(v, 0, 0.0, True)], t0=t1)
    assert np.all(r3["t_hi"] - t1 < 1e-12)
