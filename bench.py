"""Throughput benchmark: FP64 instance-steps/s of the batch/ensemble Taylor integrator.

Workload (default): the 6-body ensemble of 65,536 perturbed copies (36 states,
FP64, tol = 1e-15, order 19), integrated from t = 0 by consecutive
propagate_until slices of `--dt` time units. One bench step = one slice for the
whole ensemble, device-resident (inputs already in HBM when the timer starts;
the state carries over from slice to slice). instance-steps = the sum over
instances of nsteps in propagate_res (zero-length steps are not counted).

`--config cr3bp` runs the ragged planar CR3BP ensemble (131,072 instances,
horizons 10**U(1, 4)) once per step in the persistent form, and reports the
plain one-launch-per-step form beside it.

Prints one JSON line (see the keys below). The CPU baseline is the oracle
(a C restatement of the same algorithm) timed on this machine's cores over a
bounded sample of the same workload.
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "FP64 instance-steps/sec at tol=1e-15 vs plain GPU & CPU; p50/p99 request latency"
FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 (vector and matrix)
HBM_PEAK_GBS = 8000.0

# HBM bytes per instance per persistent launch of the default workload's kernel, from
# rocprofv3 FETCH_SIZE (x2.000) and WRITE_SIZE (x1.000) counter runs calibrated on a 1 GiB
# device copy, keyed by (workload, code form, the first 16 hex digits of the sha1 of the
# generated kernel source). A launch loads and stores every instance once and keeps the
# Taylor coefficients on chip, so the traffic scales with the instances, not the steps. A
# kernel whose source changed has no measured figure: the line then says null.
MEASURED_TRAFFIC_B_PER_INSTANCE = {
    # 65536 instances, dt=25 slices: FETCH_SIZE 28395.8 KiB (x 2.000), WRITE_SIZE 22984.4 KiB
    # (x 1.000) per launch.
    ("nbody6", "group", "189734a30a4ba451"): 1246.5,
}


def _host_info(threads):
    """CPU model and the cores this process may use (sched_getaffinity)."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        avail = os.cpu_count()
    return {"cpu_model": model, "host_cores_available": avail, "threads_used": threads}


def _default_threads():
    """One worker per available core, at most 16 (the host share of one GPU)."""
    try:
        avail = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        avail = os.cpu_count() or 1
    return max(1, min(16, avail))


def _cold_compile_s(sys_name, form):
    """Seconds to generate and compile a system's kernels with an empty code-object cache
    (a fresh process, so no in-memory cache either): the analogue of heyoka's LLVM JIT
    construction time."""
    import subprocess
    import tempfile

    code = ("import time, sys; sys.path.insert(0, %r)\n"
            "import heyoka_amd as hy\nfrom heyoka_amd import workloads as W, _lib\n"
            "wl = getattr(W, %r)(8)\nsy = wl.sys.sys if hasattr(wl.sys, 'n_orig_sv') else wl.sys\n"
            "dec = hy.taylor_decompose(sy)\nt = time.perf_counter()\n"
            "_lib.System(dec, hy.order_from_tol(wl.tol), compile_only=True, compact=%r)\n"
            "print(time.perf_counter() - t)\n") % (ROOT, sys_name, form)
    with tempfile.TemporaryDirectory() as d:
        env = dict(os.environ, HYX_CACHE_DIR=d)
        out = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=600)
    if out.returncode != 0:
        return None
    return round(float(out.stdout.strip().splitlines()[-1]), 3)


def _cpu_baseline(dec, order, state, pars, t0, t1, n_sample, threads):
    """The compiled CPU port (generated C of the system's recurrences, 8-wide lane loops,
    gcc -O3 -march=native) propagating a sample of the same workload on `threads` host threads."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import cport

    st = np.ascontiguousarray(state[:, :n_sample])
    pa = None if pars is None else np.ascontiguousarray(pars[:, :n_sample])
    tf = np.broadcast_to(np.asarray(t1, dtype=float), (st.shape[1],)).copy() if np.ndim(t1) == 0 else t1[:n_sample]
    cport.build(dec, order)
    t = time.perf_counter()
    r = cport.propagate_until(dec, order, st, tf, pars=pa, t_hi=np.full(st.shape[1], t0), nthreads=threads)
    dt = time.perf_counter() - t
    return float(r["nsteps"].astype(np.float64).sum()) / dt, dt


def _nbody_energy(st, masses):
    n = len(masses)
    pos = st.reshape(n, 6, -1)[:, :3, :]
    vel = st.reshape(n, 6, -1)[:, 3:, :]
    e = sum(0.5 * masses[i] * np.sum(vel[i] ** 2, axis=0) for i in range(n))
    for i in range(n):
        for j in range(i + 1, n):
            e = e - masses[i] * masses[j] / np.sqrt(np.sum((pos[j] - pos[i]) ** 2, axis=0))
    return e


def _cr3bp_jacobi(st, mu):
    x, y, vx, vy = st
    r1 = np.sqrt((x + mu) ** 2 + y ** 2)
    r2 = np.sqrt((x - 1 + mu) ** 2 + y ** 2)
    return x ** 2 + y ** 2 + 2 * (1 - mu) / r1 + 2 * mu / r2 - (vx ** 2 + vy ** 2)


def _accuracy_vs_real128(system, dec, order, wl, n_sample, t_final, threads):
    """GPU error against the real128 oracle on the first instances of the workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    from heyoka_amd import _lib

    st = np.ascontiguousarray(wl.state[:, :n_sample])
    pa = None if wl.pars is None else np.ascontiguousarray(wl.pars[:, :n_sample])
    tf = np.broadcast_to(np.asarray(t_final, dtype=float), (n_sample,)).copy()
    b = _lib.Batch(system, n_sample)
    b.write(_lib.FIELD_STATE, st)
    if pa is not None and system.npar:
        b.write(_lib.FIELD_PARS, pa)
    b.propagate_until(tf, None, None, 0, _lib.MODE_PERSISTENT)
    g = b.read(_lib.FIELD_STATE, st.shape)
    q = O.propagate_until(dec, 2 * order, st, tf, pars=pa, nthreads=threads, lane_batch=1, f128=True)
    r = O.propagate_until(dec, order, st, tf, pars=pa, nthreads=threads, lane_batch=1)
    qs = q["state"] + q["state_lo"]
    scale = np.maximum(np.abs(qs), 1.0)
    eg = np.max(np.abs(g - qs) / scale, axis=0)
    e64 = np.max(np.abs(r["state"] - qs) / scale, axis=0)
    return {"sample": f"first {n_sample} instances to t={float(np.max(tf))}",
            "gpu_vs_real128_median": float(np.median(eg)), "gpu_vs_real128_max": float(np.max(eg)),
            "cpu_f64_vs_real128_median": float(np.median(e64)), "cpu_f64_vs_real128_max": float(np.max(e64))}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="nbody6", choices=["nbody6", "cr3bp", "j2var", "serving", "events"])
    ap.add_argument("--rates", default="1,4", help="serving: Poisson arrival rates (requests/s)")
    ap.add_argument("--duration", type=float, default=4.0, help="serving: seconds of arrivals per rate")
    ap.add_argument("--requests", type=int, default=None, help="serving: requests per rate (default rate x duration)")
    ap.add_argument("--n", type=int, default=None, help="instances (default: the config's)")
    ap.add_argument("--dt", type=float, default=25.0, help="nbody6: time units per bench step")
    ap.add_argument("--code-form", default=None, choices=[None, "compact", "unrolled", "unrolled_lds",
                                                           "unrolled_regs", "group", "wave", "wave32", "wave16"])
    ap.add_argument("--split-chains", action="store_true",
                    help="two accumulators per long convolution sum instead of one (A/B of the code variant)")
    ap.add_argument("--no-pipeline", action="store_true",
                    help="group form: two LDS exchanges per order instead of one (A/B of the code variant)")
    ap.add_argument("--sub", type=int, default=2, choices=[1, 2, 4, 8],
                    help="compact form: lanes per instance (A/B of the code variant)")
    ap.add_argument("--cpu-threads", type=int, default=None, help="default: available cores, at most 16")
    ap.add_argument("--cpu-sample", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="nbody6: only the headline slices (no plain / CR3BP / J2V / ensemble / serving records)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy (to scale rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.cpu_threads is None:
        args.cpu_threads = _default_threads()
    if args.copy_calibration:
        import torch

        x = torch.ones(2 ** 27, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": 2 ** 30, "check": float(y[-1].item())}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: --gpus {args.gpus} is not supported (instances are independent; "
              "a multi-GPU run would shard them, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import heyoka_amd as hy
    from heyoka_amd import _lib
    from heyoka_amd import workloads as W
    from heyoka_amd.integrator import get_system

    if args.config == "serving":
        return serving_bench(args)
    if args.config == "events":
        return events_bench(args)
    if args.config == "nbody6":
        wl = W.nbody6(args.n or 65536)
    elif args.config == "j2var":
        wl = W.j2var(args.n or 32768)
    else:
        wl = W.cr3bp(args.n or 131072)
    sys_ = wl.sys.sys if hasattr(wl.sys, "n_orig_sv") else wl.sys
    dec = hy.taylor_decompose(sys_)
    order = hy.order_from_tol(wl.tol)
    compact = args.code_form
    t_c = time.perf_counter()
    system = (_lib.System(dec, order, compact=compact, split_chains=args.split_chains, pipeline=not args.no_pipeline,
                          sub=args.sub)
              if (args.split_chains or args.no_pipeline or args.sub != 2) else get_system(dec, order, compact=compact))
    t_compile = time.perf_counter() - t_c
    N = wl.state.shape[1]
    flops = dec.algorithmic_flops(order)["total"]
    batch = _lib.Batch(system, N)

    def load():
        batch.write(_lib.FIELD_STATE, wl.state)
        if wl.pars is not None:
            batch.write(_lib.FIELD_PARS, wl.pars)
        batch.write(_lib.FIELD_TIME_HI, np.zeros(N))
        batch.write(_lib.FIELD_TIME_LO, np.zeros(N))

    def one_step(i, mode=_lib.MODE_PERSISTENT):
        if args.config in ("nbody6", "j2var"):
            tf = np.full(N, args.dt * (i + 1))
        else:
            tf = wl.t_final  # cr3bp: each step is the whole ragged propagation from the initial conditions
        r = batch.propagate_until(tf, None, None, 0, mode)
        ms, launches = batch.last_timing()
        return r, ms, launches

    load()
    for i in range(args.warmup):
        if args.config == "cr3bp":
            load()
        one_step(i)
        print(f"[bench] warmup step {i} done", file=sys.stderr, flush=True)
    load()
    batch.sync()
    total_steps = 0
    kernel_ms = 0.0
    launches = 0
    wall = 0.0
    for i in range(args.steps):
        if args.config == "cr3bp":
            load()
        batch.sync()
        t = time.perf_counter()
        r, ms, nl = one_step(i)
        batch.sync()
        wall += time.perf_counter() - t
        total_steps += int(r["nsteps"].astype(np.int64).sum())
        kernel_ms += ms
        launches += nl
        bad = int(np.sum(r["outcome"] != -4294967299))
        print(f"[bench] step {i}: {r['nsteps'].astype(np.float64).sum():.3e} instance-steps in {ms:.1f} ms "
              f"({nl} launches), outcomes not time_limit: {bad}", file=sys.stderr, flush=True)
    value = total_steps / wall
    import hashlib

    src_sha = hashlib.sha1(system.source().encode()).hexdigest()[:16]
    traffic_key = (args.config, system.label, src_sha)
    avg_launch_s = kernel_ms / 1e3 / max(launches, 1)
    flops_per_launch = flops * total_steps / max(launches, 1)
    achieved_tf = flops_per_launch / avg_launch_s / 1e12
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "instance-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": {
                "nbody6": f"nbody6 ensemble: {N} instances, 36 states, tol 1e-15 (order {order}), propagate_until in "
                          f"slices of dt={args.dt} from t=0",
                "j2var": f"J2 + first-order variational ensemble: {N} instances, 42 states, tol 1e-15, propagate_until "
                         f"in slices of dt={args.dt} from t=0",
                "cr3bp": f"cr3bp ragged ensemble: {N} instances, 4 states + par[0], tol 1e-15, horizons 10**U(1,4)",
            }[args.config],
            "instances": N,
            "order": order,
            "code_form": system.form,
            "sum_chains": "split" if args.split_chains else "serial",
            "group_exchanges_per_order": ((2 if args.no_pipeline else 1) if system.form == "group" else None),
            "lanes_per_instance": system.lanes_per_instance,
            "form": "persistent",
            "launch_shape": ({"waves_per_cu": batch.tuned[0], "refill_min": batch.tuned[1], "source": "autotuned"}
                             if batch.tuned else {"waves_per_cu": "occupancy", "refill_min": 1, "source": "default"}),
            "kernel_source_sha16": src_sha,
            "algorithmic_flops_per_instance_step": flops,
            "hiprtc_compile_s": round(t_compile, 3),
        },
        "roofline": {
            "bound": "fp64_valu",
            "achieved": achieved_tf,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved_tf / FP64_PEAK_TFLOPS,
            "traffic": (MEASURED_TRAFFIC_B_PER_INSTANCE[traffic_key] * N
                        if traffic_key in MEASURED_TRAFFIC_B_PER_INSTANCE else None),
            "kernel": "hyx_persist_k",
            "avg_launch_ms": avg_launch_s * 1e3,
        },
        "cpu_baseline": None,
    }
    # Invariant drift of the whole ensemble over the timed horizon.
    final = batch.read(_lib.FIELD_STATE, wl.state.shape)
    if args.config == "nbody6":
        e0, e1 = _nbody_energy(wl.state, wl.masses), _nbody_energy(final, wl.masses)
        d = np.abs((e1 - e0) / e0)
        out["drift"] = {"invariant": "energy", "median_rel": float(np.median(d)), "max_rel": float(np.max(d))}
    elif args.config == "cr3bp":
        c0, c1 = _cr3bp_jacobi(wl.state, wl.pars[0]), _cr3bp_jacobi(final, wl.pars[0])
        d = np.abs((c1 - c0) / c0)
        ok = (r["outcome"] == -4294967299) & np.isfinite(d)
        out["drift"] = {"invariant": "Jacobi constant", "median_rel": float(np.median(d[ok])),
                        "max_rel": float(np.max(d[ok])), "p99_rel": float(np.percentile(d[ok], 99)),
                        "excluded_err_nf_state": int(np.sum(~ok))}
        ns = r["nsteps"].astype(np.int64)
        out["ragged"] = {"nsteps_median": float(np.median(ns)), "nsteps_p99": float(np.percentile(ns, 99)),
                         "nsteps_max": int(ns.max()), "persistent_launches": int(nl)}
        if system.twin_source() and not args.no_extras:
            k = int(np.argmax(ns))
            out["straggler_alone"] = _straggler_alone(system, wl, k, final[:, k].copy(), out["ms_per_step"] / 1e3)
    if not args.no_accuracy:
        ns = {"nbody6": 16, "cr3bp": 64, "j2var": 8}[args.config]
        ta = {"nbody6": 2.0, "cr3bp": 5.0, "j2var": 2.0}[args.config]
        out["accuracy"] = _accuracy_vs_real128(system, dec, order, wl, ns, ta, args.cpu_threads)
    if args.config == "cr3bp" and not args.no_plain:
        load()
        t = time.perf_counter()
        r, ms, nl = one_step(0, _lib.MODE_PLAIN)
        batch.sync()
        plain_wall = time.perf_counter() - t
        plain_value = float(r["nsteps"].astype(np.float64).sum()) / plain_wall
        out["plain_gpu"] = {"value": plain_value, "unit": "instance-steps/s", "launches": nl,
                            "persistent_over_plain": value / plain_value}
    if not args.no_cpu:
        ns = args.cpu_sample or {"nbody6": 32768, "j2var": 32768, "cr3bp": 16384}[args.config]
        if args.config in ("nbody6", "j2var"):
            # J2V takes few steps per time unit: its sample spans ten slices for ~5-10 s of CPU work.
            t_cpu = args.dt * (10 if args.config == "j2var" else 1)
            v, secs = _cpu_baseline(dec, order, wl.state, wl.pars, 0.0, t_cpu, ns, args.cpu_threads)
            sample = (f"{ns} instances of the same ensemble, t in [0, {t_cpu}], {args.cpu_threads} threads, "
                      "compiled port of the recurrences")
        else:
            v, secs = _cpu_baseline(dec, order, wl.state, wl.pars, 0.0, wl.t_final, ns, args.cpu_threads)
            sample = (f"first {ns} instances of the same ensemble to their own horizons, {args.cpu_threads} threads, "
                      "compiled port of the recurrences")
        out["cpu_baseline"] = {"value": v, "unit": "instance-steps/s", "cores": args.cpu_threads, "kind": "port",
                               "sample": sample + f" ({secs:.1f} s)", **_host_info(args.cpu_threads)}
    if args.config == "nbody6" and not args.no_extras:
        out["plain_gpu"] = _plain_slice(batch, load, N, args.dt, value)
        out["cold_compile_s"] = _cold_compile_s("nbody6", None)
        out["cr3bp_ragged"] = _cr3bp_record()
        out["j2var"] = _j2var_record(args.dt)
        out["ensemble_api"] = _ensemble_record(wl)
        out["serving"] = _serving_record()
    print(json.dumps(out), flush=True)


def _log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _timed(batch, tf, mode):
    from heyoka_amd import _lib

    batch.sync()
    t = time.perf_counter()
    r = batch.propagate_until(tf, None, None, 0, mode)
    batch.sync()
    wall = time.perf_counter() - t
    ms, nl = batch.last_timing()
    return r, wall, ms, nl


def _plain_slice(batch, load, N, dt, persistent_value):
    """The plain one-launch-per-step form on one slice of the headline workload."""
    from heyoka_amd import _lib

    load()
    r, wall, ms, nl = _timed(batch, np.full(N, dt), _lib.MODE_PLAIN)
    v = float(r["nsteps"].astype(np.float64).sum()) / wall
    _log(f"nbody6 plain slice: {v:.3e} instance-steps/s ({nl} launches)")
    return {"value": v, "unit": "instance-steps/s", "launches": int(nl), "kernel_ms": ms,
            "workload": f"the first slice (t in [0, {dt}]) of the same ensemble, one launch per step",
            "persistent_over_plain": persistent_value / v}


def _cr3bp_record():
    """The ragged CR3BP ensemble (131,072 instances, horizons 10**U(1, 4)): persistent
    work-stealing launch against the plain one-launch-per-step form, same code object."""
    import heyoka_amd as hy
    from heyoka_amd import _lib, workloads as W
    from heyoka_amd.integrator import get_system

    wl = W.cr3bp(131072)
    dec = hy.taylor_decompose(wl.sys)
    system = get_system(dec, hy.order_from_tol(wl.tol))
    N = wl.state.shape[1]
    b = _lib.Batch(system, N)

    def load():
        b.write(_lib.FIELD_STATE, wl.state)
        b.write(_lib.FIELD_PARS, wl.pars)
        b.write(_lib.FIELD_TIME_HI, np.zeros(N))
        b.write(_lib.FIELD_TIME_LO, np.zeros(N))

    out = {"workload": f"cr3bp ragged ensemble: {N} instances, 4 states + par[0], tol 1e-15, horizons 10**U(1,4), "
                       "whole propagation from the initial conditions",
           "code_form": system.label,
           "launch_shape": ({"waves_per_cu": b.tuned[0], "refill_min": b.tuned[1], "source": "autotuned"}
                            if b.tuned else {"waves_per_cu": "occupancy", "refill_min": 1, "source": "default"})}
    for label, mode in (("persistent", _lib.MODE_PERSISTENT), ("plain", _lib.MODE_PLAIN)):
        load()
        r, wall, ms, nl = _timed(b, wl.t_final, mode)
        ns = r["nsteps"].astype(np.int64)
        out[label] = {"value": float(ns.sum()) / wall, "wall_s": wall, "kernel_ms": ms, "launches": int(nl),
                      "instance_steps": int(ns.sum()), "nsteps_max": int(ns.max()),
                      "err_nf_state": int(np.sum(r["outcome"] == -4294967300))}
        _log(f"cr3bp {label}: {out[label]['value']:.3e} instance-steps/s, {wall:.2f} s, {nl} launches")
        if label == "persistent":
            k = int(np.argmax(ns))
            end_k = b.read(_lib.FIELD_STATE, wl.state.shape)[:, k].copy()
            # From the first tail-form launch to the end: the stretch only the last instances run.
            out[label]["tail_s"] = b.last_tail_ms() / 1e3
    out["persistent_over_plain"] = out["persistent"]["value"] / out["plain"]["value"]
    tsrc = system.twin_source()
    if tsrc:
        out["straggler_alone"] = _straggler_alone(system, wl, k, end_k, out["persistent"]["wall_s"])
    out["tail_form"] = ("quad: 4 lanes per instance" if "#define HYX_QUAD 1" in tsrc
                        else "twin: 2 lanes per instance" if tsrc else None)
    out["tail_launches"] = b.twin_launches()
    b.close()
    # Step latency of one instance running alone (the makespan's tail) in each form; the
    # twin beside the quad from a system built without the quad form.
    lone = {}
    forms = [(system, "lane_persistent", _lib.MODE_PERSISTENT, 0), (system, "plain", _lib.MODE_PLAIN, 0)]
    if tsrc:
        forms.insert(1, (system, "tail_form_persistent", _lib.MODE_PERSISTENT, 1 << 30))
    if "#define HYX_QUAD 1" in tsrc:
        forms.insert(2, (_lib.System(dec, hy.order_from_tol(wl.tol), quad=False), "twin_persistent",
                         _lib.MODE_PERSISTENT, 1 << 30))
    for sysm, label, mode, tail in forms:
        one = _lib.Batch(sysm, 1)
        if sysm.twin_source():
            one.set_twin_tail(tail)
        one.write(_lib.FIELD_STATE, np.ascontiguousarray(wl.state[:, 5:6]))
        one.write(_lib.FIELD_PARS, np.ascontiguousarray(wl.pars[:, 5:6]))
        one.write(_lib.FIELD_TIME_HI, np.zeros(1))
        one.write(_lib.FIELD_TIME_LO, np.zeros(1))
        r, wall, ms, nl = _timed(one, np.full(1, 2000.0), mode)
        lone[label] = ms * 1e3 / float(r["nsteps"][0])
        one.close()
    out["lone_instance_us_per_step"] = lone
    return out


def _straggler_alone(system, wl, k, end_k, ensemble_wall):
    """Ensemble instance k (its longest-running one) propagated alone in the tail form from
    the start: the floor of the ensemble's persistent makespan. It must end where it ended
    inside the ensemble (end_k), bit for bit."""
    from heyoka_amd import _lib

    n = wl.state.shape[0]
    one = _lib.Batch(system, 1)
    one.set_twin_tail(1 << 30)
    one.write(_lib.FIELD_STATE, np.ascontiguousarray(wl.state[:, k:k + 1]))
    one.write(_lib.FIELD_PARS, np.ascontiguousarray(wl.pars[:, k:k + 1]))
    one.write(_lib.FIELD_TIME_HI, np.zeros(1))
    one.write(_lib.FIELD_TIME_LO, np.zeros(1))
    r, wall, ms, nl = _timed(one, wl.t_final[k:k + 1].copy(), _lib.MODE_PERSISTENT)
    same = bool(np.array_equal(one.read(_lib.FIELD_STATE, (n, 1))[:, 0], end_k))
    one.close()
    out = {"instance": k, "nsteps": int(r["nsteps"][0]), "wall_s": wall, "kernel_ms": ms, "launches": int(nl),
           "us_per_step": ms * 1e3 / float(r["nsteps"][0]), "end_state_equals_ensemble": same,
           "share_of_persistent_wall": wall / ensemble_wall}
    _log(f"cr3bp straggler {k} alone: {ms:.0f} ms, {out['us_per_step']:.2f} us/step, same end state: {same}")
    return out


def _j2var_record(dt):
    """J2 + variational ensemble (32,768 instances): one persistent slice of dt."""
    import heyoka_amd as hy
    from heyoka_amd import _lib, workloads as W
    from heyoka_amd.integrator import get_system

    wl = W.j2var(32768)
    dec = hy.taylor_decompose(wl.sys.sys)
    system = get_system(dec, hy.order_from_tol(wl.tol))
    N = wl.state.shape[1]
    b = _lib.Batch(system, N)
    b.write(_lib.FIELD_STATE, wl.state)
    b.write(_lib.FIELD_TIME_HI, np.zeros(N))
    b.write(_lib.FIELD_TIME_LO, np.zeros(N))
    _timed(b, np.full(N, 1.0), _lib.MODE_PERSISTENT)  # warm-up
    r, wall, ms, nl = _timed(b, np.full(N, 1.0 + dt), _lib.MODE_PERSISTENT)
    ns = float(r["nsteps"].astype(np.float64).sum())
    flops = dec.algorithmic_flops(hy.order_from_tol(wl.tol))["total"]
    tf = flops * ns / (ms / 1e3) / 1e12
    _log(f"j2var slice: {ns / wall:.3e} instance-steps/s ({system.label})")
    b.close()
    return {"workload": f"J2 + first-order variational ensemble: {N} instances, 42 states, tol 1e-15, one slice "
                        f"t in [1, {1.0 + dt}]",
            "code_form": system.label, "lanes_per_instance": system.lanes_per_instance,
            "value": ns / wall, "kernel_ms": ms, "algorithmic_flops_per_instance_step": flops,
            "achieved_tflops": tf, "frac_of_fp64_peak": tf / FP64_PEAK_TFLOPS}


def _ensemble_record(wl):
    """The drop-in API on the 6-body ensemble (65,536 instances to t = 100):
    ensemble_propagate_until with 16 members of a 4,096-instance template and a Python
    generator (gen(deepcopy(ta), i)), with 65,536 one-instance members from arrays
    (ensemble_arrays), and one taylor_adaptive_batch of all 65,536 lanes (its state
    copied to the device and back around the call). Host overhead = wall time - kernel
    time."""
    import heyoka_amd as hy

    T = 100.0
    out = {"workload": f"nbody6 ensemble, {wl.state.shape[1]} instances, propagate_until t={T}"}
    B = 4096
    tmpl = hy.taylor_adaptive_batch(wl.sys, wl.state[:, :B].copy(), tol=1e-15)

    def gen(ta, i):
        ta.state[:] = wl.state[:, i * B:(i + 1) * B]
        return ta

    for label, fn in (("generator_16x4096", lambda: hy.ensemble_propagate_until_batch(tmpl, T, 16, gen)),
                      ("arrays_65536x1", lambda: hy.ensemble_propagate_until_batch(
                          hy.taylor_adaptive_batch(wl.sys, wl.state[:, :1].copy(), tol=1e-15), T, wl.state.shape[1],
                          hy.ensemble_arrays(state=wl.state)))):
        from heyoka_amd import ensemble as E

        fn()  # warm-up (first launch of the packed batch size)
        t = time.perf_counter()
        res = fn()
        wall = time.perf_counter() - t
        ms = E.last_run["kernel_ms"]
        ns = E.last_run["instance_steps"]
        out[label] = {"wall_s": wall, "kernel_s": ms / 1e3, "host_overhead_s": wall - ms / 1e3,
                      "host_overhead_frac_of_kernel": (wall - ms / 1e3) / (ms / 1e3), "members": len(res),
                      "instance_steps": ns, "instance_steps_per_s": ns / wall}
        _log(f"ensemble {label}: wall {wall:.2f} s, kernel {ms / 1e3:.2f} s")
    ta = hy.taylor_adaptive_batch(wl.sys, wl.state.copy(), tol=1e-15)
    ta.propagate_until(1.0)  # warm-up: the device batch and the first launch
    t = time.perf_counter()
    ta.propagate_until(T)
    wall = time.perf_counter() - t
    ms, _ = ta._hyx_batch.last_timing()
    ns = int(sum(r[3] for r in ta.propagate_res))
    out["batch_65536"] = {"wall_s": wall, "kernel_s": ms / 1e3, "host_overhead_s": wall - ms / 1e3,
                          "host_overhead_frac_of_kernel": (wall - ms / 1e3) / (ms / 1e3), "instance_steps": ns,
                          "instance_steps_per_s": ns / wall, "t_from": 1.0}
    _log(f"ensemble batch_65536: wall {wall:.2f} s, kernel {ms / 1e3:.2f} s")
    return out


def _serving_record():
    """A short Poisson stream through the continuous-batching server (all four systems)."""
    import argparse as _ap

    a = _ap.Namespace(rates="4", duration=4.0, requests=None)
    return serving_bench(a, emit=False)


def events_bench(args):
    """Event detection at ensemble scale: the CR3BP ensemble (65,536 instances, t in
    [0, 100] (--dt sets the horizon), tol 1e-15) propagated (A) without events, (B0) with a
    terminal event that never triggers (r^2 = 1e4: the per-step detection cost), (B) with a
    terminal escape event (r^2 = 4, no callback: stops on the device), (C) with B plus a
    non-terminal section
    event x = 1/2 - mu (direction positive) whose Python callback logs every crossing:
    instances park after a step with crossings and resume after their callbacks.
    Reports instance-steps/s of each, the crossings per second and the kernel share."""
    import heyoka_amd as hy
    from heyoka_amd import workloads as W

    N = args.n or 65536
    wl = W.cr3bp(N)
    T = args.dt if args.dt != 25.0 else 100.0
    x, y, vx, vy = hy.make_vars("x", "y", "vx", "vy")
    results = {}
    # A0: A with the twin-form tail hand-off off, the lane form throughout like the runs with
    # events (the twin form has no event detection): the baseline of the per-step event cost.
    for label in ("A_no_events", "A0_no_events_lane_form_only", "B0_terminal_never_triggers", "B_terminal_device",
                  "C_terminal_and_section_parked"):
        crossings = [0]

        def cb(ta, t, d_sgn, bidx):
            crossings[0] += 1

        kw = {}
        if label.startswith("B0"):
            # Same detection work per step, no trigger: the per-step cost of event detection.
            kw["t_events"] = [hy.t_event_batch(x * x + y * y - 1e4, direction=hy.event_direction.positive)]
        elif not label.startswith("A"):
            kw["t_events"] = [hy.t_event_batch(x * x + y * y - 4.0, direction=hy.event_direction.positive)]
        if label.startswith("C"):
            kw["nt_events"] = [hy.nt_event_batch(x - 0.5 + hy.par[0], cb, direction=hy.event_direction.positive)]
        ta = hy.taylor_adaptive_batch(wl.sys, wl.state, pars=wl.pars, tol=1e-15, code_form="unrolled", **kw)
        if label.startswith("A0") and ta._hyx_system.twin_source():
            ta._hyx_batch.set_twin_tail(0)
        ta.propagate_until(np.full(N, 1.0))  # warm-up (compiles, first launch)
        ta.set_time(np.zeros(N))
        ta.state[:] = wl.state
        ta.reset_cooldowns()
        crossings[0] = 0
        t0 = time.perf_counter()
        ta.propagate_until(np.full(N, T))
        wall = time.perf_counter() - t0
        ns = int(sum(r[3] for r in ta.propagate_res))
        ms, launches = ta.last_kernel_timing()
        stopped = int(sum(1 for r in ta.propagate_res if int(r[0]) == -1))
        results[label] = {"instance_steps_per_s": ns / wall, "instance_steps": ns, "wall_s": wall,
                          "kernel_ms_last_launch_group": ms, "launches_last_call": launches,
                          "escapes_stopped": stopped, "section_crossings": crossings[0],
                          "crossings_per_s": crossings[0] / wall if crossings[0] else None}
        print(f"[bench] events {label}: {ns / wall:.3e} instance-steps/s, {crossings[0]} crossings, "
              f"{stopped} stopped, {wall:.2f} s", file=sys.stderr, flush=True)
    base = results["A0_no_events_lane_form_only"]["instance_steps_per_s"]
    # 6-body ensemble, group form: the same comparison for a terminal close-approach
    # event between bodies 1 and 2 (|r12| = 0.1, never reached: the per-step cost) and
    # for a stopping event that does trigger (body 1 drifting past x = x0 - 2).
    wl6 = W.nbody6(N)
    nm = hy.taylor_decompose(wl6.sys).state_names
    p = [hy.expression(v) for v in nm]
    r12 = (p[6] - p[12]) ** 2 + (p[7] - p[13]) ** 2 + (p[8] - p[14]) ** 2
    nb = {}
    for label, tes in (("A_no_events", []), ("B0_close_approach_never", [hy.t_event_batch(r12 - 0.01)]),
                       ("B_crossing_stops", [hy.t_event_batch(p[6] - p[0] + 2.0,
                                                              direction=hy.event_direction.negative)])):
        ta = hy.taylor_adaptive_batch(wl6.sys, wl6.state, tol=1e-15, t_events=tes)
        ta.propagate_until(np.full(N, 1.0))
        ta.set_time(np.zeros(N))
        ta.state[:] = wl6.state
        ta.reset_cooldowns()
        t0 = time.perf_counter()
        ta.propagate_until(np.full(N, 10.0))
        wall = time.perf_counter() - t0
        ns = int(sum(r[3] for r in ta.propagate_res))
        ms, launches = ta.last_kernel_timing()
        nb[label] = {"instance_steps_per_s": ns / wall, "instance_steps": ns, "wall_s": wall, "kernel_ms": ms,
                     "code_form": ta.code_form,
                     "stopped": int(sum(1 for r in ta.propagate_res if int(r[0]) == -1))}
        print(f"[bench] events nbody6 {label}: {ns / wall:.3e} instance-steps/s, kernel {ms:.1f} ms, "
              f"{nb[label]['stopped']} stopped ({ta.code_form})", file=sys.stderr, flush=True)
    out = {
        "metric": METRIC,
        "value": results["C_terminal_and_section_parked"]["instance_steps_per_s"],
        "unit": "instance-steps/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 1,
        "ms_per_step": results["C_terminal_and_section_parked"]["wall_s"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"cr3bp ensemble with events: {N} instances, t in [0, {T}], tol 1e-15, unrolled "
                               "form, persistent; terminal escape event r^2 = 4 and a Python-logged section "
                               "x = 1/2 - mu (direction positive)"},
        "events": results,
        "relative_to_no_events_lane_form": {k: v["instance_steps_per_s"] / base for k, v in results.items()},
        "nbody6_group_events": nb,
        "nbody6_kernel_ms_relative_to_no_events": {k: v["kernel_ms"] / nb["A_no_events"]["kernel_ms"]
                                                   for k, v in nb.items()},
    }
    print(json.dumps(out), flush=True)


def serving_bench(args, emit=True):
    """Poisson streams of propagation requests through the continuous-batching server.

    Each request: a system drawn uniformly from {pendulum, cr3bp, nbody6, j2var},
    floor(10**U(0, log10 4096)) instances with seeded initial conditions,
    horizons 10**U(1, 3), tol 1e-15. Per arrival rate: `--requests` requests (else
    rate x duration), p50 / p99 latency (completion - arrival), the completed requests
    per second and the aggregate instance-steps/s. Requests of one kind share one
    system object, as a client of a server would.
    """
    from heyoka_amd import workloads as W
    from heyoka_amd.serving import Server

    rng = np.random.default_rng(5)
    kinds = ["pendulum", "cr3bp", "nbody6", "j2var"]
    gens = {"pendulum": W.pendulum, "cr3bp": W.cr3bp, "nbody6": W.nbody6, "j2var": W.j2var}
    shared = {k: gens[k](1, seed=0).sys for k in kinds}
    results = []
    for rate in [float(r) for r in args.rates.split(",")]:
        n_req = getattr(args, "requests", None) or max(1, int(round(rate * args.duration)))
        reqs = []
        t_arr = 0.0
        for k in range(n_req):
            t_arr += rng.exponential(1.0 / rate)
            kind = kinds[int(rng.integers(0, 4))]
            m = int(10 ** rng.uniform(0, np.log10(4096)))
            wl = gens[kind](m, seed=1000 + k)
            reqs.append((t_arr, kind, wl, 10 ** rng.uniform(1, 3, m)))
        srv = Server(capacity=16384, epoch_iters=256)
        # Compile and create each pool before the stream starts.
        for kind in kinds:
            wl = gens[kind](1, seed=0)
            h = srv.submit(shared[kind], wl.state, np.full(1, 0.5), pars=wl.pars, tol=1e-15)
            srv.run_until_idle()
            assert h.finished is not None
        srv.done.clear()
        srv.instance_steps = 0
        _log(f"serving rate {rate}/s: {len(reqs)} requests")
        t_start = time.perf_counter()
        handles = [srv.submit(shared[kind], wl.state, hor, pars=wl.pars, tol=1e-15, arrival=t_start + ta)
                   for ta, kind, wl, hor in reqs]
        last = time.perf_counter()
        while not all(h.finished is not None for h in handles):
            srv.step()
            now = time.perf_counter()
            if now - last > 5.0:
                last = now
                _log(f"serving rate {rate}/s: {sum(h.finished is not None for h in handles)}/{len(handles)} "
                     f"done after {now - t_start:.1f} s")
            pending = [h.arrival for h in handles if h.finished is None and h.admitted is None]
            busy = any(getattr(p, "running", None) for p in srv.pools.values())
            if pending and min(pending) > now and not busy:
                time.sleep(min(0.01, min(pending) - now))
        t_end = max(h.finished for h in handles)
        lat = np.array([h.latency for h in handles])
        span = t_end - t_start
        row = {"rate": rate, "requests": len(handles), "instances": int(sum(h.n for h in handles)),
               "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
               "completed_per_s": len(handles) / span, "instance_steps_per_s": srv.instance_steps / span,
               "span_s": span, "epochs": srv.epochs}
        results.append(row)
        _log(f"rate {rate}: p50 {row['p50_s']:.3f} s p99 {row['p99_s']:.3f} s, {row['completed_per_s']:.1f} req/s, "
             f"{row['instance_steps_per_s']:.3e} instance-steps/s")
    # The knee: the first rate whose p99 exceeds twice the lowest rate's.
    knee = next((r["rate"] for r in results if r["p99_s"] > 2.0 * results[0]["p99_s"]), None)
    summary = {"workload": "serving: Poisson request stream over pendulum / cr3bp / nbody6 / j2var, sizes "
                           "10**U(0, log10 4096), horizons 10**U(1, 3), tol 1e-15, continuous batching",
               "latency": results, "knee_rate": knee,
               "saturation_instance_steps_per_s": max(r["instance_steps_per_s"] for r in results)}
    if not emit:
        return summary
    best = max(results, key=lambda r: r["instance_steps_per_s"])
    out = {
        "metric": METRIC,
        "value": best["instance_steps_per_s"],
        "unit": "instance-steps/s",
        "n_gpus": 1,
        "steps": len(results),
        "warmup": 0,
        "ms_per_step": None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": summary["workload"], "duration_s": args.duration},
        **{k: v for k, v in summary.items() if k != "workload"},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
