"""Autotuner for the persistent form.

Searches, for one system on the current GPU:
- waves per CU of the persistent launch (occupancy, i.e. instances in flight),
- the refill threshold (idle lanes that trigger a work-stealing claim),
- optionally the code form (unrolled / compact / group / wave), which changes the
  arithmetic order and therefore results at the ulp level; for the wave form also
  the instances per wavefront ("wave16": 4, "wave32": 2, "wave": 1).
  `wave_forms(dec, order)` lists the candidates a system admits.

Launch parameters never change an instance's arithmetic, so every candidate of
a given code form gives bit-identical results (tests/test_gpu_autotune.py).
Each candidate propagates the same sample of instances for a fixed time and is
scored by instance-steps per second. Results are cached on disk, keyed by the
op-list, the order and the GPU architecture, like the code objects.
"""

from __future__ import annotations

import json
import os
import time

import numpy as np

from . import _lib
from .integrator import get_system


def _cache_path():
    p = _lib.tune_cache_path()
    os.makedirs(os.path.dirname(p), exist_ok=True)
    return p


def _key(dec, order, form):
    words, consts = _lib.oplist_args(dec)
    return _lib.tune_key(words, consts, order, form)


def _label(form, system):
    """The candidate's name: the wave form's lanes per instance and the unrolled form's
    state-row placement are part of it."""
    return system.label


def wave_forms(dec, order):
    """The wave-form variants (instances per wavefront 4, 2, 1) whose LDS rows fit."""
    out = []
    for f in ("wave16", "wave32", "wave"):
        try:
            get_system(dec, order, compact=f)
            out.append(f)
        except _lib.HyxError:
            pass
    return out


def _load():
    try:
        with open(_cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(db):
    p = _cache_path()
    tmp = p + f".tmp{os.getpid()}"
    with open(tmp, "w") as f:
        json.dump(db, f, indent=1)
    os.replace(tmp, p)


def measure(system, state, pars, t_final, waves_per_cu, refill_min, repeats=1):
    """instance-steps/s of one persistent propagation of the sample (device-resident)."""
    N = state.shape[1]
    b = _lib.Batch(system, N)
    b.set_persistent_config(waves_per_cu, refill_min)
    best = 0.0
    final = None
    for _ in range(repeats):
        b.write(_lib.FIELD_STATE, state)
        if pars is not None and system.npar:
            b.write(_lib.FIELD_PARS, pars)
        b.write(_lib.FIELD_TIME_HI, np.zeros(N))
        b.write(_lib.FIELD_TIME_LO, np.zeros(N))
        b.sync()
        t = time.perf_counter()
        r = b.propagate_until(t_final, None, None, 0, _lib.MODE_PERSISTENT)
        dt = time.perf_counter() - t
        best = max(best, float(r["nsteps"].astype(np.float64).sum()) / dt)
        final = b.read(_lib.FIELD_STATE, state.shape)
    return best, final


def autotune(dec, order, state, pars=None, t_final=1.0, forms=(None,), waves=(0, 1, 2, 4), refills=(1, 8, 32),
             use_cache=True, verbose=False):
    """Pick the fastest (form, waves_per_cu, refill_min) on a sample; returns (best, table)."""
    db = _load() if use_cache else {}
    table = []
    best = None
    tf = np.broadcast_to(np.asarray(t_final, dtype=float), (state.shape[1],)).copy()
    for form in forms:
        system = get_system(dec, order, compact=form)
        label = _label(form, system)
        key = _key(dec, order, label)
        if use_cache and key in db:
            cand = dict(db[key], form=label)
            table.append(cand)
            if best is None or cand["rate"] > best["rate"]:
                best = cand
            continue
        local = None
        for w in waves:
            for rf in refills:
                rate, _ = measure(system, state, pars, tf, w, rf)
                row = {"form": label, "waves_per_cu": w, "refill_min": rf, "rate": rate}
                table.append(row)
                if verbose:
                    print(f"[autotune] {row}", flush=True)
                if local is None or rate > local["rate"]:
                    local = row
        db[key] = {k: local[k] for k in ("waves_per_cu", "refill_min", "rate")}
        if best is None or local["rate"] > best["rate"]:
            best = local
    if use_cache:
        _store(db)
    return best, table


def tuned_config(dec, order, form=None):
    """The (waves_per_cu, refill_min) a Batch of the system applies: the user cache's, else
    the shipped table's, else the defaults (0, 1)."""
    system = get_system(dec, order, compact=form)
    rec = _lib.tuned_launch(_key(dec, order, _label(form, system)))
    return (rec[0], rec[1]) if rec else (0, 1)


def tune_workloads(out_path=None, verbose=True):
    """Tune the benchmark workloads (6-body group form; CR3BP over the unrolled form with
    state rows in registers or in LDS and the compact form; J2 + variational over the wave
    variants) on samples of their seeded inputs, into the user cache and, with out_path,
    into a table of the shipped format."""
    from . import workloads as W
    from .decompose import taylor_decompose, order_from_tol

    jobs = []
    nb = W.nbody6(16384)
    jobs.append((nb.sys, nb.tol, nb.state, None, 2.0, (None,)))
    # A bulk sample (every instance to t = 100): a ragged sample's makespan is one
    # straggler's, which hardly depends on the launch shape.
    cr = W.cr3bp(65536)
    jobs.append((cr.sys, cr.tol, cr.state, cr.pars, 100.0, ("unrolled_regs", "unrolled_lds", True)))
    jv = W.j2var(8192)
    jobs.append((jv.sys.sys, jv.tol, jv.state, None, 2.0, ("wave16", "wave32", "wave")))
    table = {}
    for sys_, tol, st, pa, tf, forms in jobs:
        dec = taylor_decompose(sys_)
        order = order_from_tol(tol)
        best, rows = autotune(dec, order, st, pars=pa, t_final=tf, forms=forms, use_cache=False, verbose=verbose)
        db = _load()
        for f in forms:
            system = get_system(dec, order, compact=f)
            lab = _label(f, system)
            cand = [r for r in rows if r["form"] == lab]
            win = max(cand, key=lambda r: r["rate"])
            db[_key(dec, order, lab)] = table[_key(dec, order, lab)] = {
                "waves_per_cu": win["waves_per_cu"], "refill_min": win["refill_min"], "rate": win["rate"],
                "workload": f"{len(st[0])} instances, form {lab}"}
        _store(db)
        if verbose:
            print(f"[autotune] best {best}", flush=True)
    if out_path:
        with open(out_path, "w") as f:
            json.dump(table, f, indent=1, sort_keys=True)
    return table
