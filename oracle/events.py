"""CPU oracle for event detection (terminal and non-terminal events).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker, never by the
product path.

The algorithm restated here is the event detection of the heyoka C++ library
(version 7.11.0; it is an external dependency of heyoka.py and not part of the
reference tree), as published in its documentation and exercised by heyoka.py's
tests and notebooks:

- every step, each event function g(x(t), t) gets its Taylor polynomial on the
  step (the same normalised-derivative recurrences as the state; here obtained
  from an augmented system y' = g, whose coefficient k+1 is g^[k] / (k+1));
- the polynomial is rescaled to s in [0, 1] (tau = s h), and its roots in
  [0, 1) are isolated by Descartes' rule of signs with bisection (the
  Vincent-Collins-Akritas scheme: sign variations of (1 + x)^n p(1 / (1 + x))
  on a work list of subintervals) and refined with TOMS 748
  (Alefeld, Potra and Shi 1995; scipy.optimize.toms748);
- d_sgn is the sign of dg/dt at the root; events with a direction keep roots of
  that sign only (doc/notebooks/Event detection.ipynb cells 12-17);
- the earliest terminal root shortens the step; the terminal event then enters a
  cooldown (its own, or 10 * tol * max|x| / |dg/dt|) during which its roots are
  ignored (cells 22-30); non-terminal roots later than the terminal one are
  dropped; callbacks run in chronological order (cell 21);
- the batch test heyoka/test.py:693-1000 and heyoka/_test_batch_integrator.py:555-600
  fix the callback order and the outcome codes (-(i + 1) stopping, i continuing).

The cooldown formula's constant is the one part no reference output pins: its
effect (no re-trigger right after a terminal event) is what the tests check.
"""

from __future__ import annotations

import math

import numpy as np
from scipy.optimize import toms748

import pyoracle as O

SUCCESS, STEP_LIMIT, TIME_LIMIT, ERR_NF_STATE = O.SUCCESS, O.STEP_LIMIT, O.TIME_LIMIT, O.ERR_NF_STATE


def _two_sum(a, b):
    x = a + b
    z = x - a
    return x, (a - (x - z)) + (b - z)


def dfl_add(ah, al, bh, bl):
    """Double-length sum (Knuth two-sum, renormalised)."""
    s, e = _two_sum(ah, bh)
    t, f = _two_sum(al, bl)
    e = e + t
    x = s + e
    e = e - (x - s)
    s = x
    e = e + f
    x = s + e
    return x, e - (x - s)


# ---- polynomial root isolation (Descartes / bisection) and refinement -----------------------
def _taylor_shift(c, a):
    """Coefficients of p(x + a)."""
    q = list(c)
    n = len(q) - 1
    for i in range(n):
        for k in range(n - 1, i - 1, -1):
            q[k] = q[k] + a * q[k + 1]
    return q


def _variations(c, lb, ub):
    """Sign variations of (1 + x)^n p(lb + (ub - lb) / (1 + x)): bounds the roots in (lb, ub)."""
    r = _taylor_shift(c, lb) if lb != 0.0 else list(c)
    w = ub - lb
    n = len(r) - 1
    q = [0.0] * (n + 1)
    wk = 1.0
    for k in range(n + 1):
        q[n - k] = r[k] * wk
        wk *= w
    q = _taylor_shift(q, 1.0)
    v, last = 0, 0
    for x in q:
        s = (x > 0) - (x < 0)
        if s:
            if last and s != last:
                v += 1
            last = s
    return v


def _peval(c, x):
    r = 0.0
    for a in reversed(c):
        r = r * x + a
    return r


def poly_roots01(c, max_depth=44):
    """Roots of sum c[k] s^k for s in [0, 1), ascending (isolation + TOMS 748)."""
    c = [float(x) for x in c]
    if not any(c) or not all(math.isfinite(x) for x in c):
        return []
    roots = []
    if c[0] == 0.0:
        roots.append(0.0)
    work = [(0.0, 1.0, 0)]
    while work:
        lb, ub, d = work.pop()
        v = _variations(c, lb, ub)
        if v == 0:
            continue
        if v == 1:
            fa, fb = _peval(c, lb), _peval(c, ub)
            if fa != 0.0 and fb != 0.0 and (fa > 0) != (fb > 0):
                r = toms748(lambda s: _peval(c, s), lb, ub, xtol=1e-300, rtol=4 * np.finfo(float).eps, maxiter=200)
                roots.append(float(r))
                continue
        if d >= max_depth:
            continue
        mid = 0.5 * (lb + ub)
        if _peval(c, mid) == 0.0:
            roots.append(mid)
        work.append((lb, mid, d + 1))
        work.append((mid, ub, d + 1))
    return sorted(set(r for r in roots if 0.0 <= r < 1.0))


def event_roots(g, h):
    """Roots tau of sum g[k] tau^k with tau from 0 (included) towards h (excluded), in
    chronological order, each with d_sgn = sign(dg/dt) there."""
    if h == 0.0:
        return []
    g = [float(x) for x in g]
    h = float(h)
    n = len(g)
    c = [g[k] * h**k for k in range(n)]
    dg = [k * g[k] for k in range(1, n)]
    out = []
    for s in poly_roots01(c):
        tau = s * h
        d = _peval(dg, tau)
        out.append((tau, int((d > 0) - (d < 0))))
    return out


# ---- event polynomials ---------------------------------------------------------------------
def event_jets(hy, sys_pairs, exprs, order, state, pars, t_hi):
    """Taylor coefficients g[e][k][lane], k = 0..order, of each event function along the
    trajectory starting at state (one column per lane), from the augmented system
    y_e' = g_e integrated one order higher (y_e^[k+1] = g_e^[k] / (k + 1))."""
    n, B = state.shape
    ys = [hy.expression(f"__oracle_ev_{e}") for e in range(len(exprs))]
    aug = list(sys_pairs) + list(zip(ys, exprs))
    dec = hy.taylor_decompose(aug)
    st = np.concatenate([state, np.zeros((len(exprs), B))])
    npar = max(dec.n_par, 1)
    pa = np.zeros((npar, B)) if pars is None else np.concatenate(
        [pars, np.zeros((max(npar - pars.shape[0], 0), B))])[:npar]
    r = O.step_batch(dec, order + 1, st, pars=pa, t_hi=t_hi, max_dt=0.0, want_tc=True)
    tc = r["tc"][n:]  # [ne][order + 2][B]
    ks = np.arange(1, order + 2, dtype=np.float64)[None, :, None]
    return tc[:, 1:, :] * ks  # [ne][order + 1][B]


def propagate(hy, sys_pairs, state, tf, t_events=(), nt_events=(), pars=None, tol=2.220446049250313e-16,
              max_steps=0, max_delta_t=np.inf, t0=0.0, cooldowns=None, high_accuracy=False):
    """propagate_until(tf) with events, step by step, every lane of state [n, B].

    t_events: list of (expression, direction, cooldown, stopping); nt_events: list of
    (expression, direction). Returns dict with state, t_hi, t_lo, outcome, nsteps,
    min_h, max_h, nt (per lane: list of (event index, time, d_sgn) in processing order),
    te (per lane: list of (event index, time, d_sgn)) and cooldowns."""
    order = O_order(tol)
    dec = hy.taylor_decompose(sys_pairs)
    state = np.array(state, dtype=np.float64)
    n, B = state.shape
    npar = max(dec.n_par, 1)
    pars = np.zeros((npar, B)) if pars is None else np.array(pars, dtype=np.float64)
    tes, ntes = list(t_events), list(nt_events)
    exprs = [e[0] for e in tes] + [e[0] for e in ntes]
    nte = len(tes)
    thi = np.broadcast_to(np.asarray(t0, dtype=np.float64), (B,)).copy()
    tlo = np.zeros(B)
    tf = np.broadcast_to(np.asarray(tf, dtype=np.float64), (B,)).copy()
    mdt = np.broadcast_to(np.asarray(max_delta_t, dtype=np.float64), (B,)).copy()
    cds = [[None] * nte for _ in range(B)] if cooldowns is None else [list(c) for c in cooldowns]
    done = np.zeros(B, dtype=bool)
    res = np.full(B, TIME_LIMIT, dtype=np.int64)
    nst = np.zeros(B, dtype=np.int64)
    mn, mx = np.full(B, np.inf), np.zeros(B)
    nt_log = [[] for _ in range(B)]
    te_log = [[] for _ in range(B)]
    it = 0
    while not done.all():
        rem = np.array([dfl_add(tf[b], 0.0, -thi[b], -tlo[b])[0] for b in range(B)])
        fwd = rem >= 0
        lim = np.where(fwd, np.minimum(mdt, rem), np.maximum(-mdt, rem))
        lim = np.where(done, 0.0, lim)
        x0 = state.copy()
        # The state's and the event functions' coefficients; the step size comes from both.
        r0 = O.step_batch(dec, order, state, pars=pars, t_hi=thi, t_lo=tlo, max_dt=0.0, want_tc=True)
        tc = r0["tc"]
        jets = event_jets(hy, sys_pairs, exprs, order, x0, pars, thi) if exprs else np.zeros((0, order + 1, B))
        hs = step_sizes(tc, jets, order)
        hsel = np.where(hs >= np.abs(lim), np.abs(lim), hs)
        hsel = np.where(lim < 0, -hsel, hsel)
        r = O.step_batch(dec, order, state, pars=pars, t_hi=thi, t_lo=tlo, max_dt=hsel, high_accuracy=high_accuracy)
        it += 1
        for b in range(B):
            if done[b]:
                continue
            h = float(r["h"][b])
            assert h == hsel[b]
            oc = TIME_LIMIT if h == lim[b] else SUCCESS
            best, te = abs(h), None
            for e in range(nte):
                expr, d, cd_user, stopping = tes[e]
                for tau, ds in event_roots(jets[e, :, b], h):
                    if abs(tau) >= best:
                        break
                    if d != 0 and ds != d:
                        continue
                    c = cds[b][e]
                    if c is not None and abs(dfl_add(thi[b], tlo[b], -c[0], -c[1])[0] + tau) < c[2]:
                        continue
                    best, te = abs(tau), (e, tau, ds)
                    break
            nts = []
            for j in range(len(ntes)):
                expr, d = ntes[j]
                for tau, ds in event_roots(jets[nte + j, :, b], h):
                    if abs(tau) >= best:
                        break
                    if d != 0 and ds != d:
                        continue
                    nts.append((abs(tau), j, tau, ds))
            hend = h
            if te is not None:
                e, tau, ds = te
                hend = tau
                for q in range(n):
                    acc = tc[q, order, b]
                    for k in range(order - 1, -1, -1):
                        acc = acc * tau + tc[q, k, b]
                    state[q, b] = acc
                r["state"][:, b] = state[:, b]
            else:
                state[:, b] = r["state"][:, b]
            nhi, nlo = dfl_add(thi[b], tlo[b], hend, 0.0)
            for nt in sorted(nts):
                t = dfl_add(thi[b], tlo[b], nt[2], 0.0)[0]
                nt_log[b].append((nt[1], t, nt[3]))
            # Cooldowns: expire, then start the triggered one's.
            for e in range(nte):
                c = cds[b][e]
                if c is not None and not abs(dfl_add(nhi, nlo, -c[0], -c[1])[0]) < c[2]:
                    cds[b][e] = None
            if te is not None:
                e, tau, ds = te
                cd = tes[e][2]
                if cd < 0:
                    g = jets[e, :, b]
                    der = abs(sum(k * g[k] * tau ** (k - 1) for k in range(1, order + 1)))
                    cd = 10.0 * (tol * float(np.max(np.abs(x0[:, b])))) / der if der else math.inf
                    if not math.isfinite(cd):
                        cd = 0.0
                cds[b][e] = (nhi, nlo, cd)
                te_log[b].append((e, nhi, ds))
            thi[b], tlo[b] = nhi, nlo
            nst[b] += 1 if hend != 0.0 else 0
            if not np.all(np.isfinite(state[:, b])):
                res[b], done[b] = ERR_NF_STATE, True
                continue
            step_oc = (-(te[0] + 1)) if te is not None else oc
            # heyoka counts only unclamped steps: a step shortened by a terminal event does not count.
            if te is None and oc == SUCCESS:
                mn[b] = min(mn[b], abs(hend))
                mx[b] = max(mx[b], abs(hend))
            if te is not None and tes[te[0]][3]:
                res[b], done[b] = step_oc, True
            elif te is None and oc == TIME_LIMIT and hend == rem[b]:
                res[b], done[b] = TIME_LIMIT, True
        if max_steps and it >= max_steps:
            res[~done] = STEP_LIMIT
            break
    return dict(state=state, t_hi=thi, t_lo=tlo, outcome=res, nsteps=nst, min_h=mn, max_h=mx, nt=nt_log, te=te_log,
                cooldowns=cds)


def step_sizes(tc, jets, order):
    """Step size of every lane when the event functions' coefficients (jets) join the
    state's (tc) in the norms of the last two orders; the max |x| switch between
    absolute and relative tolerance uses the state only."""
    rows = np.concatenate([tc, jets], axis=0)
    mx = np.max(np.abs(tc[:, 0, :]), axis=0)
    mo = np.max(np.abs(rows[:, order, :]), axis=0)
    mom1 = np.max(np.abs(rows[:, order - 1, :]), axis=0)
    out = np.empty(tc.shape[2])
    rf = O.rhofac(order)
    for b in range(out.size):
        num = 1.0 if mx[b] <= 1.0 else float(mx[b])
        rho_o = (num / float(mo[b])) ** (1.0 / order)
        rho_om1 = (num / float(mom1[b])) ** (1.0 / (order - 1))
        out[b] = (rho_om1 if rho_om1 < rho_o else rho_o) * rf
    return out


def O_order(tol):
    return int(math.ceil(-math.log(tol) / 2.0 + 1.0))
