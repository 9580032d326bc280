"""ctypes front end of the CPU oracle (oracle/build/libtaylor_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker or the timed CPU baseline,
never by the product path.
"""

from __future__ import annotations

import ctypes as ct
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "libtaylor_oracle.so")

SUCCESS = -4294967297
STEP_LIMIT = -4294967298
TIME_LIMIT = -4294967299
ERR_NF_STATE = -4294967300

_lib = None


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc only)."""
    cmd = ["make", "-C", HERE] + (["-B"] if force else [])
    subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ct.CDLL(LIB_PATH)
        P = ct.POINTER
        d, u32, i64, u64 = ct.c_double, ct.c_uint32, ct.c_int64, ct.c_uint64
        for sfx in ("f64", "f128"):
            f = getattr(L, f"ora_{sfx}_step_batch")
            f.restype = ct.c_int
            f.argtypes = [P(u32), P(d), u32, ct.c_int, u32, P(d), P(d), P(d), P(d), P(d), P(d), P(d), P(i64), P(d)]
            f = getattr(L, f"ora_{sfx}_propagate_until")
            f.restype = ct.c_int
            f.argtypes = [P(u32), P(d), u32, ct.c_int, u32, u32, P(d), P(d), P(d), P(d), P(d), P(d), P(d), P(d),
                          u64, P(i64), P(d), P(d), P(u64), ct.c_int]
            f = getattr(L, f"ora_{sfx}_propagate_grid")
            f.restype = ct.c_int
            f.argtypes = [P(u32), P(d), u32, ct.c_int, u32, u32, P(d), P(d), P(d), P(d), P(d), u64, P(d), u64,
                          P(i64), P(d), P(d), P(u64), P(d), ct.c_int]
            f = getattr(L, f"ora_{sfx}_rhofac_value")
            f.restype = d
            f.argtypes = [u32]
        _lib = L
    return _lib


def _p(a, ctype):
    return None if a is None else a.ctypes.data_as(ct.POINTER(ctype))


def _f64(a, shape=None):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    if shape is not None:
        a = np.ascontiguousarray(np.broadcast_to(a, shape))
    return a.copy()


def step_batch(dec, order, state, pars=None, t_hi=None, t_lo=None, max_dt=np.inf, high_accuracy=False,
               f128=False, want_tc=False):
    """One step of every lane. Returns dict(state, state_lo, t_hi, t_lo, h, outcome, tc)."""
    words = np.ascontiguousarray(dec.to_words())
    consts = dec.consts_array()
    state = _f64(state)
    n, B = state.shape
    pars = _f64(pars if pars is not None else np.zeros((max(dec.n_par, 1), B)))
    t_hi = _f64(0.0 if t_hi is None else t_hi, (B,))
    t_lo = _f64(0.0 if t_lo is None else t_lo, (B,))
    mdt = _f64(max_dt, (B,))
    h = np.zeros(B)
    oc = np.zeros(B, dtype=np.int64)
    lo = np.zeros_like(state) if f128 else None
    tc = np.zeros((n, order + 1, B)) if want_tc else None
    f = getattr(lib(), "ora_f128_step_batch" if f128 else "ora_f64_step_batch")
    rc = f(_p(words, ct.c_uint32), _p(consts, ct.c_double), order, int(high_accuracy), B, _p(state, ct.c_double),
           _p(lo, ct.c_double), _p(pars, ct.c_double), _p(t_hi, ct.c_double), _p(t_lo, ct.c_double),
           _p(mdt, ct.c_double), _p(h, ct.c_double), _p(oc, ct.c_int64), _p(tc, ct.c_double))
    if rc != 0:
        raise RuntimeError(f"oracle step_batch failed with code {rc}")
    return dict(state=state, state_lo=lo, t_hi=t_hi, t_lo=t_lo, h=h, outcome=oc, tc=tc)


def propagate_until(dec, order, state, tf, pars=None, t_hi=None, t_lo=None, tf_lo=None, max_delta_t=np.inf,
                    max_steps=0, high_accuracy=False, f128=False, lane_batch=8, nthreads=1):
    """propagate_until for every lane. Returns dict with state (+ state_lo for f128),
    times and the per-lane (outcome, min_h, max_h, nsteps)."""
    words = np.ascontiguousarray(dec.to_words())
    consts = dec.consts_array()
    state = _f64(state)
    n, B = state.shape
    pars = _f64(pars if pars is not None else np.zeros((max(dec.n_par, 1), B)))
    t_hi = _f64(0.0 if t_hi is None else t_hi, (B,))
    t_lo = _f64(0.0 if t_lo is None else t_lo, (B,))
    tf_hi = _f64(tf, (B,))
    tf_lo = _f64(0.0 if tf_lo is None else tf_lo, (B,))
    mdt = _f64(max_delta_t, (B,))
    oc = np.zeros(B, dtype=np.int64)
    mn = np.zeros(B)
    mx = np.zeros(B)
    ns = np.zeros(B, dtype=np.uint64)
    lo = np.zeros_like(state) if f128 else None
    f = getattr(lib(), "ora_f128_propagate_until" if f128 else "ora_f64_propagate_until")
    rc = f(_p(words, ct.c_uint32), _p(consts, ct.c_double), order, int(high_accuracy), B, lane_batch,
           _p(state, ct.c_double), _p(lo, ct.c_double), _p(pars, ct.c_double), _p(t_hi, ct.c_double),
           _p(t_lo, ct.c_double), _p(tf_hi, ct.c_double), _p(tf_lo, ct.c_double), _p(mdt, ct.c_double),
           int(max_steps), _p(oc, ct.c_int64), _p(mn, ct.c_double), _p(mx, ct.c_double), _p(ns, ct.c_uint64),
           int(nthreads))
    if rc != 0:
        raise RuntimeError(f"oracle propagate_until failed with code {rc}")
    return dict(state=state, state_lo=lo, t_hi=t_hi, t_lo=t_lo, outcome=oc, min_h=mn, max_h=mx, nsteps=ns)


def propagate_grid(dec, order, state, grid, pars=None, t_hi=None, t_lo=None, max_delta_t=np.inf, max_steps=0,
                   high_accuracy=False, f128=False, lane_batch=8, nthreads=1):
    """propagate_grid for every lane: grid [n_pts, B] (first row = current time). Returns the
    propagate_until dict plus out [n_pts, n, B] (NaN where a lane did not reach the point)."""
    words = np.ascontiguousarray(dec.to_words())
    consts = dec.consts_array()
    state = _f64(state)
    n, B = state.shape
    grid = _f64(grid)
    assert grid.ndim == 2 and grid.shape[1] == B and grid.shape[0] >= 1
    pars = _f64(pars if pars is not None else np.zeros((max(dec.n_par, 1), B)))
    t_hi = _f64(grid[0] if t_hi is None else t_hi, (B,))
    t_lo = _f64(0.0 if t_lo is None else t_lo, (B,))
    mdt = _f64(max_delta_t, (B,))
    oc = np.zeros(B, dtype=np.int64)
    mn = np.zeros(B)
    mx = np.zeros(B)
    ns = np.zeros(B, dtype=np.uint64)
    out = np.full((grid.shape[0], n, B), np.nan)
    f = getattr(lib(), "ora_f128_propagate_grid" if f128 else "ora_f64_propagate_grid")
    rc = f(_p(words, ct.c_uint32), _p(consts, ct.c_double), order, int(high_accuracy), B, lane_batch,
           _p(state, ct.c_double), _p(pars, ct.c_double), _p(t_hi, ct.c_double), _p(t_lo, ct.c_double),
           _p(grid, ct.c_double), grid.shape[0], _p(mdt, ct.c_double), int(max_steps), _p(oc, ct.c_int64),
           _p(mn, ct.c_double), _p(mx, ct.c_double), _p(ns, ct.c_uint64), _p(out, ct.c_double), int(nthreads))
    if rc != 0:
        raise RuntimeError(f"oracle propagate_grid failed with code {rc}")
    return dict(state=state, t_hi=t_hi, t_lo=t_lo, outcome=oc, min_h=mn, max_h=mx, nsteps=ns, out=out)


def rhofac(order, f128=False):
    return getattr(lib(), "ora_f128_rhofac_value" if f128 else "ora_f64_rhofac_value")(order)
