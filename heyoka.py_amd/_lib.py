"""ctypes binding of libhyx.so (the C-ABI declared in include/hyx.h).

ctypes releases the GIL for the duration of every foreign call, as the
reference's binding does around step/propagate
(heyoka/expose_batch_integrators.cpp:253-255, 289-291, 300-302).
"""

from __future__ import annotations

import ctypes as ct
import os
import subprocess
import threading

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libhyx.so")
CSRC = os.path.join(PKG_DIR, "csrc")

FIELD_STATE, FIELD_PARS, FIELD_TIME_HI, FIELD_TIME_LO = 0, 1, 2, 3
MODE_PERSISTENT, MODE_PLAIN, MODE_GRAPH = 0, 1, 2
FLAG_HIGH_ACCURACY, FLAG_COMPILE_ONLY, FLAG_FORCE_COMPACT, FLAG_FORCE_UNROLLED, FLAG_SOURCE_ONLY = 1, 2, 4, 8, 16
FLAG_FORCE_GROUP = 32
FLAG_FORCE_WAVE = 4096
FLAG_WAVE32 = 8192
FLAG_WAVE16 = 16384
FLAG_LDS_ROWS = 32768
FLAG_NO_LDS_ROWS = 65536
FLAG_NO_TWIN = 131072
FLAG_NO_QUAD = 262144
FORMS = {0: "unrolled", 1: "compact", 2: "group", 3: "wave"}
FLAG_GRID = 64
FLAG_SPLIT_CHAINS = 128
FLAG_NO_PIPELINE = 256
FLAG_SUB = {1: 512, 2: 0, 4: 1024, 8: 2048}  # compact form: lanes per instance


class prop_res(ct.Structure):
    _fields_ = [("outcome", ct.c_int64), ("min_h", ct.c_double), ("max_h", ct.c_double), ("nsteps", ct.c_uint64)]


class tune_res(ct.Structure):
    _fields_ = [("waves_per_cu", ct.c_uint32), ("refill_min", ct.c_uint32), ("rate", ct.c_double),
                ("from_cache", ct.c_int)]


class events_desc(ct.Structure):
    _fields_ = [("ops", ct.POINTER(ct.c_uint32)), ("n_words", ct.c_uint32), ("consts", ct.POINTER(ct.c_double)),
                ("n_consts", ct.c_uint32), ("n_te", ct.c_uint32), ("n_nt", ct.c_uint32),
                ("direction", ct.POINTER(ct.c_int32)), ("cooldown", ct.POINTER(ct.c_double)),
                ("stopping", ct.POINTER(ct.c_int32)), ("tol", ct.c_double)]


class HyxError(RuntimeError):
    pass


_lib = None
_lock = threading.Lock()


def build(force: bool = False) -> str:
    """Compile libhyx.so in-tree with its Makefile (hipcc, gfx950)."""
    cmd = ["make", "-C", CSRC] + (["-B"] if force else [])
    subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    return LIB_PATH


def lib():
    """Load libhyx.so. It is never built implicitly: a missing library is an error."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')"
            )
        L = ct.CDLL(LIB_PATH)
        P = ct.POINTER
        u32, u64, i64, d, vp, cp = ct.c_uint32, ct.c_uint64, ct.c_int64, ct.c_double, ct.c_void_p, ct.c_char_p
        sig = {
            "hyx_sys_create": (ct.c_int, [P(u32), u32, P(d), u32, u32, u32, P(vp)]),
            "hyx_sys_precompile": (ct.c_int, [P(u32), u32, P(d), u32, u32, u32, cp]),
            "hyx_sys_order": (ct.c_int, [vp]),
            "hyx_sys_dim": (ct.c_int, [vp]),
            "hyx_sys_npar": (ct.c_int, [vp]),
            "hyx_sys_is_compact": (ct.c_int, [vp]),
            "hyx_sys_code_form": (ct.c_int, [vp]),
            "hyx_sys_lanes_per_instance": (ct.c_int, [vp]),
            "hyx_sys_twin_source": (ct.c_size_t, [vp, ct.c_char_p, ct.c_size_t]),
            "hyx_batch_set_twin_tail": (ct.c_int, [vp, ct.c_uint32]),
            "hyx_batch_twin_launches": (u64, [vp]),
            "hyx_sys_source": (ct.c_size_t, [vp, ct.c_char_p, ct.c_size_t]),
            "hyx_sys_kernel_info": (ct.c_int, [vp, P(ct.c_int), P(ct.c_int)]),
            "hyx_sys_destroy": (None, [vp]),
            "hyx_batch_create": (ct.c_int, [vp, u64, P(vp)]),
            "hyx_batch_destroy": (None, [vp]),
            "hyx_batch_size": (u64, [vp]),
            "hyx_batch_write": (ct.c_int, [vp, ct.c_int, P(d)]),
            "hyx_batch_read": (ct.c_int, [vp, ct.c_int, P(d)]),
            "hyx_batch_device_ptr": (ct.c_int, [vp, ct.c_int, P(vp)]),
            "hyx_step": (ct.c_int, [vp, P(d), ct.c_int, P(i64), P(d)]),
            "hyx_read_tc": (ct.c_int, [vp, P(d)]),
            "hyx_propagate_until": (ct.c_int, [vp, P(d), P(d), P(d), u64, ct.c_int, P(prop_res)]),
            # Device pointers throughout (caller-owned arrays), then a hipStream_t.
            "hyx_propagate_until_noalloc": (ct.c_int, [vp, u64, vp, vp, vp, vp, vp, vp, vp, u64, vp, vp, vp, vp,
                                                       vp]),
            "hyx_propagate_grid": (ct.c_int, [vp, P(d), u64, P(d), u64, ct.c_int, P(prop_res), P(d)]),
            "hyx_batch_last_timing": (ct.c_int, [vp, P(d), P(u64)]),
            "hyx_batch_last_tail_ms": (ct.c_int, [vp, P(d)]),
            "hyx_batch_set_persistent_config": (ct.c_int, [vp, u32, u32]),
            "hyx_batch_set_queue": (ct.c_int, [vp, P(u32)]),
            "hyx_batch_set_iter_budget": (ct.c_int, [vp, u64]),
            "hyx_batch_sync": (ct.c_int, [vp]),
            "hyx_slots_reset": (ct.c_int, [vp]),
            "hyx_slots_write": (ct.c_int, [vp, ct.c_int, P(u32), u32, P(d)]),
            "hyx_slots_read": (ct.c_int, [vp, ct.c_int, P(u32), u32, P(d)]),
            "hyx_slots_arm": (ct.c_int, [vp, P(u32), u32, P(d), P(d), P(d)]),
            "hyx_slots_launch": (ct.c_int, [vp, u64]),
            "hyx_slots_wait": (ct.c_int, [vp, P(u32)]),
            "hyx_slots_poll": (ct.c_int, [vp, P(ct.c_int)]),
            "hyx_autotune": (ct.c_int, [vp, P(d), ct.c_char_p, P(tune_res)]),
            "hyx_slots_snapshot_words": (u64, [vp]),
            "hyx_slots_snapshot": (ct.c_int, [vp, P(u32), u32, P(u64)]),
            "hyx_slots_restore": (ct.c_int, [vp, P(u32), u32, P(u64)]),
            "hyx_cfunc_create": (ct.c_int, [P(u32), u32, P(d), u32, u32, u32, u32, P(vp)]),
            "hyx_cfunc_source": (ct.c_size_t, [vp, ct.c_char_p, ct.c_size_t]),
            "hyx_cfunc_eval": (ct.c_int, [vp, P(d), P(d), u32, P(d), u64, P(d)]),
            "hyx_cfunc_eval_batch": (ct.c_int, [vp, vp, P(d)]),
            "hyx_cfunc_destroy": (None, [vp]),
            "hyx_slots_results": (ct.c_int, [vp, P(u32), u32, P(u32), P(prop_res)]),
            "hyx_sys_create_events": (ct.c_int, [P(u32), u32, P(d), u32, u32, u32, P(events_desc), P(vp)]),
            "hyx_sys_n_events": (ct.c_int, [vp, P(u32), P(u32)]),
            "hyx_events_read": (ct.c_int, [vp, P(i64), P(u32), P(d), P(i64), P(u32)]),
            "hyx_batch_set_event_pause": (ct.c_int, [vp, ct.c_int]),
            "hyx_events_parked": (ct.c_int, [vp, P(u32), P(u32)]),
            "hyx_events_gather": (ct.c_int, [vp, P(u32), u32, P(i64), P(u32), P(d), P(i64), P(d), P(d)]),
            "hyx_events_resume": (ct.c_int, [vp, P(u32), u32, P(i64)]),
            "hyx_propagate_resume": (ct.c_int, [vp, u64, ct.c_int, P(prop_res), P(d)]),
            "hyx_batch_set_clog": (ct.c_int, [vp, u32]),
            "hyx_clog_drain": (ct.c_int, [vp, P(u32), u32, P(u32), P(d), P(d)]),
            "hyx_cooldowns_read": (ct.c_int, [vp, P(d)]),
            "hyx_cooldowns_write": (ct.c_int, [vp, P(d)]),
            "hyx_last_error": (ct.c_char_p, []),
            "hyx_device_count": (ct.c_int, []),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
        return _lib


EXPORTED = (
    "hyx_sys_create", "hyx_sys_precompile", "hyx_sys_order", "hyx_sys_dim", "hyx_sys_npar", "hyx_sys_is_compact",
    "hyx_sys_code_form", "hyx_sys_lanes_per_instance", "hyx_sys_twin_source", "hyx_batch_set_twin_tail",
    "hyx_batch_twin_launches",
    "hyx_sys_source",
    "hyx_sys_kernel_info", "hyx_sys_destroy", "hyx_batch_create", "hyx_batch_destroy", "hyx_batch_size",
    "hyx_batch_write", "hyx_batch_read", "hyx_batch_device_ptr", "hyx_step", "hyx_read_tc", "hyx_propagate_until",
    "hyx_propagate_until_noalloc",
    "hyx_propagate_grid",
    "hyx_batch_last_timing", "hyx_batch_last_tail_ms", "hyx_batch_set_persistent_config", "hyx_batch_set_queue", "hyx_batch_set_iter_budget",
    "hyx_batch_sync", "hyx_slots_reset", "hyx_slots_write", "hyx_slots_read", "hyx_slots_arm", "hyx_slots_launch",
    "hyx_slots_wait", "hyx_slots_poll", "hyx_slots_results", "hyx_autotune",
    "hyx_slots_snapshot_words", "hyx_slots_snapshot", "hyx_slots_restore",
    "hyx_cfunc_create", "hyx_cfunc_source", "hyx_cfunc_eval", "hyx_cfunc_eval_batch", "hyx_cfunc_destroy",
    "hyx_sys_create_events", "hyx_sys_n_events", "hyx_events_read", "hyx_cooldowns_read", "hyx_cooldowns_write",
    "hyx_batch_set_event_pause", "hyx_events_parked", "hyx_events_gather", "hyx_events_resume",
    "hyx_propagate_resume", "hyx_batch_set_clog", "hyx_clog_drain",
    "hyx_last_error", "hyx_device_count",
)


def check(rc: int):
    if rc != 0:
        msg = lib().hyx_last_error().decode(errors="replace")
        raise HyxError(f"hyx error {rc}: {msg}")


def dptr(a: np.ndarray, ctype=ct.c_double):
    return None if a is None else a.ctypes.data_as(ct.POINTER(ctype))


def oplist_args(dec):
    words = np.ascontiguousarray(dec.to_words(), dtype=np.uint32)
    consts = np.ascontiguousarray(dec.consts_array(), dtype=np.float64)
    return words, consts


def _flags(high_accuracy=False, compact=None):
    """compact: None (automatic), True / "compact", False / "unrolled" ("unrolled_lds" / "unrolled_regs":
    state rows in LDS / in registers), "group", or "wave" (64 lanes per instance; "wave32" / "wave16": 32
    or 16)."""
    f = FLAG_HIGH_ACCURACY if high_accuracy else 0
    if compact is True or compact == "compact":
        f |= FLAG_FORCE_COMPACT
    elif compact is False or compact == "unrolled":
        f |= FLAG_FORCE_UNROLLED
    elif compact == "unrolled_lds":
        f |= FLAG_FORCE_UNROLLED | FLAG_LDS_ROWS
    elif compact == "unrolled_regs":
        f |= FLAG_FORCE_UNROLLED | FLAG_NO_LDS_ROWS
    elif compact == "group":
        f |= FLAG_FORCE_GROUP
    elif compact in ("wave", "wave32", "wave16"):
        f |= FLAG_FORCE_WAVE | {"wave": 0, "wave32": FLAG_WAVE32, "wave16": FLAG_WAVE16}[compact]
    return f


def precompile(dec, order: int, high_accuracy: bool = False, arch: str = "gfx950", compact=None, grid=False,
               quad: bool = True):
    """Compile the system's kernels (grid=True: its propagate_grid kernels; quad=False: the twin
    rather than the quad tail form) into the code-object cache (no GPU needed)."""
    words, consts = oplist_args(dec)
    flags = _flags(high_accuracy, compact) | (FLAG_GRID if grid else 0) | (0 if quad else FLAG_NO_QUAD)
    check(lib().hyx_sys_precompile(dptr(words, ct.c_uint32), len(words), dptr(consts), len(dec.consts), order, flags,
                                   arch.encode()))


# ---- tuned launch shapes ---------------------------------------------------------------
# The autotuner (autotune.py) records per system and code form the fastest persistent
# launch shape (waves per CU, refill threshold) in the user cache
# ($HYX_CACHE_DIR or code_cache/autotune.json); tuned_gfx950.json beside this file holds
# the shapes measured for the benchmark workloads and ships with the package. Every Batch
# applies the shape recorded for its system. Launch shapes never change results.
TUNED_PATH = os.path.join(PKG_DIR, "tuned_gfx950.json")
_tuned_cache = {}


def tune_cache_path() -> str:
    d = os.environ.get("HYX_CACHE_DIR") or os.path.join(PKG_DIR, "code_cache")
    return os.path.join(d, "autotune.json")


def tune_key(words, consts, order: int, label: str) -> str:
    import hashlib

    h = hashlib.sha1(np.ascontiguousarray(words).tobytes() + np.ascontiguousarray(consts).tobytes()).hexdigest()[:16]
    return f"{h}-p{int(order)}-{label}-gfx950"


def _tuned_db(path):
    import json

    try:
        mt = os.stat(path).st_mtime_ns
    except OSError:
        return {}
    hit = _tuned_cache.get(path)
    if hit is None or hit[0] != mt:
        try:
            with open(path) as f:
                db = json.load(f)
        except (OSError, ValueError):
            db = {}
        _tuned_cache[path] = hit = (mt, db)
    return hit[1]


def tuned_launch(key: str):
    """The recorded (waves_per_cu, refill_min, rate) for a system key, or None: the user
    cache first, then the shipped table."""
    for path in (tune_cache_path(), TUNED_PATH):
        rec = _tuned_db(path).get(key)
        if rec:
            return int(rec["waves_per_cu"]), int(rec["refill_min"]), float(rec.get("rate", 0.0))
    return None


def tuned_rates(keys: dict) -> dict:
    """{label: recorded rate} for the labelled keys, all from one table: the user cache if it
    records any of them, else the shipped table (rates measured on different samples do not
    compare)."""
    for path in (tune_cache_path(), TUNED_PATH):
        db = _tuned_db(path)
        rates = {lab: float(db[k].get("rate", 0.0)) for lab, k in keys.items() if k in db}
        if rates:
            return rates
    return {}


def device_count() -> int:
    return int(lib().hyx_device_count())


class EventSpec:
    """Events of a system: the event op-list (rows = the system's state variables, then
    padding; right-hand sides = event functions, terminal ones first), directions,
    terminal cooldowns (< 0: automatic) and the integrator tolerance."""

    def __init__(self, dec, n_te: int, n_nt: int, directions, cooldowns, tol: float, stopping=None):
        self.dec, self.n_te, self.n_nt, self.tol = dec, int(n_te), int(n_nt), float(tol)
        self.directions = np.ascontiguousarray(directions, dtype=np.int32)
        self.cooldowns = np.ascontiguousarray(cooldowns if len(cooldowns) else [0.0], dtype=np.float64)
        st = [1] * self.n_te if stopping is None else [int(bool(x)) for x in stopping]
        self.stopping = np.ascontiguousarray(st if st else [1], dtype=np.int32)

    def key(self):
        w, c = oplist_args(self.dec)
        return (w.tobytes(), c.tobytes(), self.n_te, self.n_nt, self.directions.tobytes(), self.cooldowns.tobytes(),
                self.stopping.tobytes(), self.tol)


class System:
    """Owner of a hyx_sys handle."""

    def __init__(self, dec, order: int, high_accuracy: bool = False, compile_only: bool = False, compact=None,
                 source_only: bool = False, split_chains: bool = False, twin: bool = True,
                 pipeline: bool = True, sub: int = 2, events: EventSpec | None = None, quad: bool = True):
        L = lib()
        words, consts = oplist_args(dec)
        flags = (_flags(high_accuracy, compact) | (FLAG_COMPILE_ONLY if compile_only else 0)
                 | (FLAG_SOURCE_ONLY if source_only else 0) | (FLAG_SPLIT_CHAINS if split_chains else 0)
                 | (0 if twin else FLAG_NO_TWIN) | (0 if quad else FLAG_NO_QUAD)
                 | (0 if pipeline else FLAG_NO_PIPELINE) | FLAG_SUB[int(sub)])
        h = ct.c_void_p()
        if events is None:
            check(L.hyx_sys_create(dptr(words, ct.c_uint32), len(words), dptr(consts), len(dec.consts), order, flags,
                                   ct.byref(h)))
        else:
            ew, ec = oplist_args(events.dec)
            self._ev_keep = (ew, ec, events.directions, events.cooldowns, events.stopping)
            ed = events_desc(dptr(ew, ct.c_uint32), len(ew), dptr(ec), len(events.dec.consts), events.n_te,
                             events.n_nt, dptr(events.directions, ct.c_int32), dptr(events.cooldowns),
                             dptr(events.stopping, ct.c_int32), events.tol)
            check(L.hyx_sys_create_events(dptr(words, ct.c_uint32), len(words), dptr(consts), len(dec.consts), order,
                                          flags, ct.byref(ed), ct.byref(h)))
        self.events = events
        self.h = h
        self.order = order
        self.n = dec.n_state
        self.npar = dec.n_par
        self._L = L
        self.compact = bool(L.hyx_sys_is_compact(self.h))
        self.form = FORMS[int(L.hyx_sys_code_form(self.h))]
        self.lanes_per_instance = int(L.hyx_sys_lanes_per_instance(self.h))
        # The autotuner's name for this code object: the wave variants carry their lanes per
        # instance, the unrolled form where its state rows live.
        if self.form == "wave":
            self.label = {64: "wave", 32: "wave32", 16: "wave16"}[self.lanes_per_instance]
        elif self.form == "unrolled":
            self.label = "unrolled_lds" if "#define HYX_LDS_STATE 1" in self.source() else "unrolled"
        else:
            self.label = self.form
        self.tune_key = tune_key(words, consts, order, self.label) if events is None and not high_accuracy else None

    def source(self) -> str:
        n = self._L.hyx_sys_source(self.h, None, 0)
        buf = ct.create_string_buffer(n + 1)
        self._L.hyx_sys_source(self.h, buf, n + 1)
        return buf.value.decode()

    def twin_source(self) -> str:
        """The tail form's source: the quad form (four lanes per instance) where the system's
        state variables pair, else the twin form (two); "" when the system has neither."""
        n = self._L.hyx_sys_twin_source(self.h, None, 0)
        if n == 0:
            return ""
        buf = ct.create_string_buffer(n + 1)
        self._L.hyx_sys_twin_source(self.h, buf, n + 1)
        return buf.value.decode()

    def kernel_info(self):
        v, s = ct.c_int(), ct.c_int()
        check(self._L.hyx_sys_kernel_info(self.h, ct.byref(v), ct.byref(s)))
        return {"vgprs": v.value, "scratch_bytes": s.value}

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value:
            self._L.hyx_sys_destroy(h)
            self.h = None


class Batch:
    """Owner of a hyx_batch handle: N device-resident instances of one System."""

    def __init__(self, system: System, n: int):
        self.sys = system
        self.N = int(n)
        self._L = system._L
        h = ct.c_void_p()
        check(self._L.hyx_batch_create(system.h, self.N, ct.byref(h)))
        self.h = h
        # The tuned persistent launch shape of this system, if one is recorded.
        self.tuned = tuned_launch(system.tune_key) if getattr(system, "tune_key", None) else None
        if self.tuned is not None:
            self.set_persistent_config(self.tuned[0], self.tuned[1])

    def write(self, field: int, arr: np.ndarray):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        check(self._L.hyx_batch_write(self.h, field, dptr(a)))

    def read(self, field: int, shape) -> np.ndarray:
        out = np.empty(shape, dtype=np.float64)
        if out.size:
            check(self._L.hyx_batch_read(self.h, field, dptr(out)))
        return out

    def device_ptr(self, field: int) -> int:
        p = ct.c_void_p()
        check(self._L.hyx_batch_device_ptr(self.h, field, ct.byref(p)))
        return p.value

    def step(self, max_dt: np.ndarray | None, write_tc: bool):
        oc = np.empty(self.N, dtype=np.int64)
        h = np.empty(self.N, dtype=np.float64)
        md = None if max_dt is None else np.ascontiguousarray(max_dt, dtype=np.float64)
        check(self._L.hyx_step(self.h, dptr(md), int(write_tc), dptr(oc, ct.c_int64), dptr(h)))
        return oc, h

    def _cap(self):
        cap = ct.c_uint32()
        check(self._L.hyx_events_read(self.h, None, None, None, None, ct.byref(cap)))
        return cap.value

    def events_read(self):
        """The last step's events: (te [N] (4 * index + d_sgn + 1, or -1), nt_count [N],
        nt_time [cap, N], nt_code [cap, N])."""
        cap = self._cap()
        te = np.empty(self.N, dtype=np.int64)
        cnt = np.empty(self.N, dtype=np.uint32)
        tm = np.empty((cap, self.N), dtype=np.float64)
        code = np.empty((cap, self.N), dtype=np.int64)
        check(self._L.hyx_events_read(self.h, dptr(te, ct.c_int64), dptr(cnt, ct.c_uint32), dptr(tm),
                                      dptr(code, ct.c_int64), None))
        return te, cnt, tm, code

    def set_event_pause(self, on: bool):
        check(self._L.hyx_batch_set_event_pause(self.h, int(bool(on))))

    def events_parked(self) -> np.ndarray:
        lanes = np.empty(self.N, dtype=np.uint32)
        n = ct.c_uint32()
        check(self._L.hyx_events_parked(self.h, dptr(lanes, ct.c_uint32), ct.byref(n)))
        return lanes[: n.value].copy()

    def events_gather(self, lanes, with_tc=True):
        """For parked instances: (te, nt_count, nt_time [cap, k], nt_code [cap, k], h, tc [n, p+1, k])."""
        sl = np.ascontiguousarray(lanes, dtype=np.uint32)
        k = len(sl)
        cap = self._cap()
        te = np.empty(k, dtype=np.int64)
        cnt = np.empty(k, dtype=np.uint32)
        tm = np.empty((cap, k), dtype=np.float64)
        code = np.empty((cap, k), dtype=np.int64)
        h = np.empty(k, dtype=np.float64)
        tc = np.empty((self.sys.n, self.sys.order + 1, k), dtype=np.float64) if with_tc else None
        check(self._L.hyx_events_gather(self.h, dptr(sl, ct.c_uint32), k, dptr(te, ct.c_int64), dptr(cnt, ct.c_uint32),
                                        dptr(tm), dptr(code, ct.c_int64), dptr(h), dptr(tc)))
        return te, cnt, tm, code, h, tc

    def events_resume(self, lanes, stop):
        sl = np.ascontiguousarray(lanes, dtype=np.uint32)
        st = np.ascontiguousarray(stop, dtype=np.int64)
        check(self._L.hyx_events_resume(self.h, dptr(sl, ct.c_uint32), len(sl), dptr(st, ct.c_int64)))

    def set_clog(self, cap: int):
        check(self._L.hyx_batch_set_clog(self.h, int(cap)))
        self._clog_cap = int(cap)

    def clog_drain(self, lanes):
        """Empty the step logs of the listed instances: (nrec [k], tc [cap, n, p+1, k], tt [cap, 3, k])."""
        sl = np.ascontiguousarray(lanes, dtype=np.uint32)
        k, cap = len(sl), self._clog_cap
        nrec = np.empty(k, dtype=np.uint32)
        tc = np.empty((cap, self.sys.n, self.sys.order + 1, k), dtype=np.float64)
        tt = np.empty((cap, 3, k), dtype=np.float64)
        check(self._L.hyx_clog_drain(self.h, dptr(sl, ct.c_uint32), k, dptr(nrec, ct.c_uint32), dptr(tc), dptr(tt)))
        return nrec, tc, tt

    def propagate_resume(self, max_steps=0, mode=MODE_PERSISTENT, grid_out=None):
        res = (prop_res * self.N)()
        check(self._L.hyx_propagate_resume(self.h, int(max_steps), int(mode), res, dptr(grid_out)))
        arr = np.frombuffer(res, dtype=np.dtype([("outcome", "<i8"), ("min_h", "<f8"), ("max_h", "<f8"),
                                                 ("nsteps", "<u8")]), count=self.N)
        return arr.copy()

    def cooldowns(self) -> np.ndarray:
        """[n_te, 3, N]: trigger time hi, lo and length (< 0: no cooldown)."""
        out = np.empty((self.sys.events.n_te, 3, self.N), dtype=np.float64)
        if out.size:
            check(self._L.hyx_cooldowns_read(self.h, dptr(out)))
        return out

    def set_cooldowns(self, cd):
        a = np.ascontiguousarray(cd, dtype=np.float64)
        if a.size:
            check(self._L.hyx_cooldowns_write(self.h, dptr(a)))

    def read_tc(self) -> np.ndarray:
        out = np.empty((self.sys.n, self.sys.order + 1, self.N), dtype=np.float64)
        check(self._L.hyx_read_tc(self.h, dptr(out)))
        return out

    def propagate_until(self, tf_hi, tf_lo=None, max_delta_t=None, max_steps=0, mode=MODE_PERSISTENT,
                        want_res=True):
        th = np.ascontiguousarray(tf_hi, dtype=np.float64)
        tl = None if tf_lo is None else np.ascontiguousarray(tf_lo, dtype=np.float64)
        md = None if max_delta_t is None else np.ascontiguousarray(max_delta_t, dtype=np.float64)
        res = (prop_res * self.N)() if want_res else None
        check(self._L.hyx_propagate_until(self.h, dptr(th), dptr(tl), dptr(md), int(max_steps), int(mode), res))
        if not want_res:
            return None
        arr = np.frombuffer(res, dtype=np.dtype([("outcome", "<i8"), ("min_h", "<f8"), ("max_h", "<f8"),
                                                 ("nsteps", "<u8")]), count=self.N)
        return arr.copy()

    def propagate_grid(self, grid, max_delta_t=None, max_steps=0, mode=MODE_PERSISTENT):
        """grid: [n_pts, N] times. Returns (prop_res records, out [n_pts, n, N])."""
        g = np.ascontiguousarray(grid, dtype=np.float64)
        if g.ndim != 2 or g.shape[1] != self.N:
            raise ValueError(f"the grid must have shape (n_pts, {self.N})")
        md = None if max_delta_t is None else np.ascontiguousarray(max_delta_t, dtype=np.float64)
        res = (prop_res * self.N)()
        out = np.empty((g.shape[0], self.sys.n, self.N), dtype=np.float64)
        check(self._L.hyx_propagate_grid(self.h, dptr(g), g.shape[0], dptr(md), int(max_steps), int(mode), res,
                                         dptr(out)))
        arr = np.frombuffer(res, dtype=np.dtype([("outcome", "<i8"), ("min_h", "<f8"), ("max_h", "<f8"),
                                                 ("nsteps", "<u8")]), count=self.N)
        return arr.copy(), out

    def autotune(self, tf_hi, cache_path=None):
        """Native autotuner: the fastest persistent (waves_per_cu, refill_min) for the batch's
        current contents propagated to tf_hi; the batch keeps it. Returns a dict."""
        th = np.ascontiguousarray(np.broadcast_to(np.asarray(tf_hi, dtype=np.float64), (self.N,)))
        r = tune_res()
        check(self._L.hyx_autotune(self.h, dptr(th), None if cache_path is None else str(cache_path).encode(),
                                   ct.byref(r)))
        return {"waves_per_cu": r.waves_per_cu, "refill_min": r.refill_min, "rate": r.rate,
                "from_cache": bool(r.from_cache)}

    def last_timing(self):
        ms, n = ct.c_double(), ct.c_uint64()
        check(self._L.hyx_batch_last_timing(self.h, ct.byref(ms), ct.byref(n)))
        return ms.value, n.value

    def last_tail_ms(self) -> float:
        """Of the last persistent propagation: ms from the first tail-form launch to the end."""
        ms = ct.c_double()
        check(self._L.hyx_batch_last_tail_ms(self.h, ct.byref(ms)))
        return ms.value

    def set_persistent_config(self, waves_per_cu: int = 0, refill_min: int = 1):
        check(self._L.hyx_batch_set_persistent_config(self.h, int(waves_per_cu), int(refill_min)))

    def set_queue(self, order):
        if order is None:
            check(self._L.hyx_batch_set_queue(self.h, None))
        else:
            o = np.ascontiguousarray(order, dtype=np.uint32)
            check(self._L.hyx_batch_set_queue(self.h, dptr(o, ct.c_uint32)))

    def set_twin_tail(self, tail_min: int):
        """Hand-off threshold of persistent propagations to the tail form (instances in flight;
        default eight per CU, 0: never)."""
        check(self._L.hyx_batch_set_twin_tail(self.h, int(tail_min)))

    def twin_launches(self) -> int:
        return int(self._L.hyx_batch_twin_launches(self.h))

    def set_iter_budget(self, iters: int):
        check(self._L.hyx_batch_set_iter_budget(self.h, int(iters)))

    def sync(self):
        check(self._L.hyx_batch_sync(self.h))

    # ---- slots (serving) ------------------------------------------------------------------
    def slots_reset(self):
        check(self._L.hyx_slots_reset(self.h))

    def slots_write(self, field: int, slots, src):
        sl = np.ascontiguousarray(slots, dtype=np.uint32)
        a = np.ascontiguousarray(src, dtype=np.float64)
        check(self._L.hyx_slots_write(self.h, field, dptr(sl, ct.c_uint32), len(sl), dptr(a)))

    def slots_read(self, field: int, slots, rows: int) -> np.ndarray:
        sl = np.ascontiguousarray(slots, dtype=np.uint32)
        out = np.empty((rows, len(sl)), dtype=np.float64)
        if out.size:
            check(self._L.hyx_slots_read(self.h, field, dptr(sl, ct.c_uint32), len(sl), dptr(out)))
        return out

    def slots_arm(self, slots, tf_hi, tf_lo=None, max_delta_t=None):
        sl = np.ascontiguousarray(slots, dtype=np.uint32)
        th = np.ascontiguousarray(tf_hi, dtype=np.float64)
        tl = None if tf_lo is None else np.ascontiguousarray(tf_lo, dtype=np.float64)
        md = None if max_delta_t is None else np.ascontiguousarray(max_delta_t, dtype=np.float64)
        check(self._L.hyx_slots_arm(self.h, dptr(sl, ct.c_uint32), len(sl), dptr(th), dptr(tl), dptr(md)))

    def slots_launch(self, iter_budget: int):
        check(self._L.hyx_slots_launch(self.h, int(iter_budget)))

    def slots_wait(self) -> int:
        n = ct.c_uint32()
        check(self._L.hyx_slots_wait(self.h, ct.byref(n)))
        return n.value

    def slots_snapshot(self, slots) -> np.ndarray:
        """Every word the propagation of these slots depends on (host copy, uint64)."""
        sl = np.ascontiguousarray(slots, dtype=np.uint32)
        w = int(self._L.hyx_slots_snapshot_words(self.h))
        out = np.empty(w * len(sl), dtype=np.uint64)
        check(self._L.hyx_slots_snapshot(self.h, dptr(sl, ct.c_uint32), len(sl), dptr(out, ct.c_uint64)))
        return out

    def slots_restore(self, slots, blob):
        sl = np.ascontiguousarray(slots, dtype=np.uint32)
        b = np.ascontiguousarray(blob, dtype=np.uint64)
        check(self._L.hyx_slots_restore(self.h, dptr(sl, ct.c_uint32), len(sl), dptr(b, ct.c_uint64)))

    def slots_poll(self) -> bool:
        """True when the running epoch (if any) has finished; never blocks."""
        f = ct.c_int()
        check(self._L.hyx_slots_poll(self.h, ct.byref(f)))
        return bool(f.value)

    def slots_results(self, slots):
        sl = np.ascontiguousarray(slots, dtype=np.uint32)
        done = np.zeros(len(sl), dtype=np.uint32)
        res = (prop_res * max(len(sl), 1))()
        check(self._L.hyx_slots_results(self.h, dptr(sl, ct.c_uint32), len(sl), dptr(done, ct.c_uint32), res))
        arr = np.frombuffer(res, dtype=np.dtype([("outcome", "<i8"), ("min_h", "<f8"), ("max_h", "<f8"),
                                                 ("nsteps", "<u8")]), count=len(sl)).copy()
        return done.astype(bool), arr

    def close(self):
        """Free the batch's device memory now (also done when the object is collected)."""
        h = getattr(self, "h", None)
        if h is not None and h.value:
            self._L.hyx_batch_destroy(h)
            self.h = None

    __del__ = close
