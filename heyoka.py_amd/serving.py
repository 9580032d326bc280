"""Continuous-batching propagation server.

Requests are ensembles (state ``[n, m]``, final times, optional parameters) of
any supported system. The server keeps one device-resident pool per compiled
system (a batch of ``capacity`` instance slots on its own HIP stream) and runs
a loop of bounded epochs:

1. admit waiting requests, first come first served, into free slots of their
   system's pool (creating the pool if the HBM budget allows): scatter their
   state, parameters and time into the slots and arm them with their final
   times;
2. start an epoch on every pool with armed slots and no epoch in flight
   (asynchronous launches on separate streams, so pools of different systems
   run concurrently), each slot advancing by at most ``epoch_iters`` steps;
3. poll the pools without blocking; harvest each pool as soon as its own epoch
   ends: gather the results of requests whose slots are all done, free the
   slots and record their latency.

A request that arrives while others are running is admitted at the next epoch
boundary, as soon as slots are free; running slots stay on the device between
epochs. When a new system's pool does not fit the HBM budget, the pool holding the
youngest work is parked: its running requests are snapshotted to host memory
(every word their propagation depends on, hyx_slots_snapshot) and its device memory
is freed; parked requests are restored (hyx_slots_restore, into any free slots) as
soon as their pool fits again, before new admissions, and continue bit for bit. A slot's arithmetic does not depend on the slot, the pool's other
occupants or the epoch boundaries, so a result is bit-for-bit the same alone
and in a mix (tests/test_gpu_serving.py).
"""

from __future__ import annotations

import itertools
import threading
import time as _time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .decompose import taylor_decompose, order_from_tol
from .expr import _ex
from .integrator import get_system, dfl_add


@dataclass
class Request:
    id: int
    key: tuple
    n: int
    state: np.ndarray
    pars: np.ndarray | None
    t0: np.ndarray
    tf: np.ndarray
    arrival: float
    done_event: threading.Event = field(default_factory=threading.Event)
    slots: np.ndarray | None = None
    admitted: float | None = None
    finished: float | None = None
    result_state: np.ndarray | None = None
    result_time: np.ndarray | None = None
    result: np.ndarray | None = None  # propagate_res columns

    @property
    def latency(self) -> float:
        return self.finished - self.arrival

    def wait(self, timeout=None):
        return self.done_event.wait(timeout)


class _Pool:
    def __init__(self, system, dec, capacity):
        self.system = system
        self.dec = dec
        self.capacity = capacity
        self.batch = _lib.Batch(system, capacity)
        self.batch.slots_reset()
        self.free = deque(range(capacity))
        self.running: list[Request] = []
        self.npar = system.npar
        self.in_epoch = False  # an epoch of this pool is on the GPU
        self.epoch_start = 0.0

    def bytes(self):
        n, npar = self.dec.n_state, max(self.npar, 1)
        return self.capacity * 8 * (n + npar + 16)


class Server:
    """Continuous-batching server over one GPU."""

    def __init__(self, capacity: int = 16384, epoch_iters: int = 512, hbm_budget_bytes: int = 32 << 30,
                 evict: bool = True):
        self.capacity = int(capacity)
        self.epoch_iters = int(epoch_iters)
        self.hbm_budget = int(hbm_budget_bytes)
        self.evict = bool(evict)
        self.parked: dict = {}  # system key -> [(request, slot snapshot)]
        self.evictions = 0
        self.pools: dict = {}
        self.waiting: deque = deque()
        self.done: list[Request] = []
        self._ids = itertools.count()
        self._lock = threading.Lock()
        self._stop = threading.Event()
        self._thread = None
        self.instance_steps = 0
        self.busy_time = 0.0
        self.epochs = 0
        self._systems: dict = {}

    # ---- submission ---------------------------------------------------------------------
    def submit(self, sys, state, t_final, pars=None, tol=1e-15, t0=0.0, arrival=None) -> Request:
        st = np.ascontiguousarray(np.asarray(state, dtype=np.float64))
        if st.ndim != 2:
            raise ValueError("The state of a request must be a 2-D array [n, m]")
        m = st.shape[1]
        if m == 0 or m > self.capacity:
            raise ValueError(f"A request must hold between 1 and {self.capacity} instances")
        order = order_from_tol(tol)
        # Requests usually repeat a system object: its decomposition and compiled code are
        # looked up by identity (the cache holds the object, so its id stays valid).
        hit = self._systems.get((id(sys), order))
        if hit is None:
            pairs = sys.sys if hasattr(sys, "n_orig_sv") else [(_ex(a), _ex(b)) for a, b in sys]
            dec = taylor_decompose(pairs)
            hit = self._systems[(id(sys), order)] = (sys, dec, get_system(dec, order))
        _, dec, system = hit
        if st.shape[0] != dec.n_state:
            raise ValueError(f"The state has {st.shape[0]} rows but the system has {dec.n_state} variables")
        key = (id(system),)
        pa = None
        if dec.n_par:
            if pars is None:
                raise ValueError("The system has runtime parameters: pars must be given")
            pa = np.ascontiguousarray(np.asarray(pars, dtype=np.float64).reshape(-1, m)[: dec.n_par])
        t0v = np.broadcast_to(np.asarray(t0, dtype=np.float64), (m,)).copy()
        tfv = np.broadcast_to(np.asarray(t_final, dtype=np.float64), (m,)).copy()
        req = Request(next(self._ids), key, m, st, pa, t0v, tfv,
                      arrival=_time.perf_counter() if arrival is None else arrival)
        with self._lock:
            if key not in self.pools:
                self.pools[key] = (system, dec)  # pool created lazily at admission
            self.waiting.append(req)
        return req

    # ---- scheduling ---------------------------------------------------------------------
    def _used(self):
        return sum(q.bytes() for q in self.pools.values() if isinstance(q, _Pool))

    def _pool(self, key, evict=False):
        """The pool of a system, created if the HBM budget allows; with evict=True another
        pool may be parked (its running requests snapshotted to host memory and its device
        memory freed) to make room."""
        p = self.pools[key]
        if isinstance(p, _Pool):
            return p
        system, dec = p
        cand = _Pool.__new__(_Pool)
        cand.capacity, cand.dec, cand.npar = self.capacity, dec, system.npar
        if self._used() + cand.bytes() > self.hbm_budget:
            if not evict:
                return None
            # Free idle pools first, then park the pool holding the youngest work (its
            # earliest request arrived last), so the longest-waiting requests keep their slots.
            victims = [(0 if not q.running else 1, -min(r.arrival for r in q.running) if q.running else 0.0, k)
                       for k, q in self.pools.items() if isinstance(q, _Pool) and k != key and not q.in_epoch]
            victims.sort()
            for busy, _, k in victims:
                if self._used() + cand.bytes() <= self.hbm_budget:
                    break
                if busy:
                    self._park(k)
                else:
                    q = self.pools[k]
                    self.pools[k] = (q.system, q.dec)
                    q.batch.close()
            if self._used() + cand.bytes() > self.hbm_budget:
                return None
        pool = _Pool(system, dec, self.capacity)
        self.pools[key] = pool
        return pool

    def _park(self, key):
        """Snapshot every running request of a pool (hyx_slots_snapshot) and free the pool."""
        pool = self.pools[key]
        for req in pool.running:
            self.parked.setdefault(key, []).append((req, pool.batch.slots_snapshot(req.slots)))
            req.slots = None
        self.pools[key] = (pool.system, pool.dec)
        pool.batch.close()
        self.evictions += 1

    def _unpark(self):
        """Restore parked requests into their system's pool when it can be (re)created."""
        for key in list(self.parked):
            pool = self._pool(key)
            if pool is None or pool.in_epoch:
                continue
            items = self.parked[key]
            keep = []
            for req, blob in items:
                if len(pool.free) < req.n:
                    keep.append((req, blob))
                    continue
                slots = np.array([pool.free.popleft() for _ in range(req.n)], dtype=np.uint32)
                pool.batch.slots_restore(slots, blob)
                req.slots = slots
                pool.running.append(req)
            if keep:
                self.parked[key] = keep
            else:
                del self.parked[key]

    def _release_idle(self):
        """Free pools with nothing running while parked work waits for HBM."""
        if not self.parked:
            return
        for k, q in list(self.pools.items()):
            if isinstance(q, _Pool) and not q.running and not q.in_epoch and k not in self.parked:
                self.pools[k] = (q.system, q.dec)
                q.batch.close()

    def _admit(self, now):
        with self._lock:
            pending = list(self.waiting)
            self.waiting.clear()
        keep = []
        blocked = set()
        for req in pending:
            if req.arrival > now or req.key in blocked:
                keep.append(req)
                continue
            pool = self._pool(req.key, evict=self.evict and not self.parked)
            if pool is None or pool.in_epoch or len(pool.free) < req.n:
                # Admission waits for the pool's epoch boundary (its slots are written and
                # armed between epochs), keeping first-come-first-served order within a system.
                blocked.add(req.key)
                keep.append(req)
                continue
            slots = np.array([pool.free.popleft() for _ in range(req.n)], dtype=np.uint32)
            b = pool.batch
            b.slots_write(_lib.FIELD_STATE, slots, req.state)
            if req.pars is not None:
                b.slots_write(_lib.FIELD_PARS, slots, req.pars)
            b.slots_write(_lib.FIELD_TIME_HI, slots, req.t0[None, :])
            b.slots_write(_lib.FIELD_TIME_LO, slots, np.zeros((1, req.n)))
            b.slots_arm(slots, req.tf)
            req.slots = slots
            req.admitted = _time.perf_counter()
            pool.running.append(req)
        with self._lock:
            self.waiting.extendleft(reversed(keep))

    def _harvest(self, p):
        still = []
        for req in p.running:
            done, res = p.batch.slots_results(req.slots)
            if not done.all():
                still.append(req)
                continue
            req.result = res
            req.result_state = p.batch.slots_read(_lib.FIELD_STATE, req.slots, p.dec.n_state)
            req.result_time = p.batch.slots_read(_lib.FIELD_TIME_HI, req.slots, 1)[0]
            req.finished = _time.perf_counter()
            self.instance_steps += int(res["nsteps"].astype(np.int64).sum())
            p.free.extend(int(s) for s in req.slots)
            self.done.append(req)
            req.done_event.set()
        p.running = still

    def step(self) -> int:
        """One scheduling round: admit, start an epoch on every busy pool that is idle,
        then harvest every pool whose epoch has ended. Pools run independently: a pool
        whose epochs are long (a large system) does not hold back the others' harvests
        and admissions. Returns the number of requests still waiting or running."""
        now = _time.perf_counter()
        self._release_idle()
        self._unpark()
        self._admit(now)
        pools = [p for p in self.pools.values() if isinstance(p, _Pool)]
        for p in pools:
            if p.running and not p.in_epoch:
                p.batch.slots_launch(self.epoch_iters)
                p.in_epoch = True
                p.epoch_start = _time.perf_counter()
        inflight = [p for p in pools if p.in_epoch]
        ended = [p for p in inflight if p.batch.slots_poll()]
        if inflight and not ended:
            _time.sleep(0.0002)
        for p in ended:
            p.batch.slots_wait()
            p.in_epoch = False
            self.busy_time += _time.perf_counter() - p.epoch_start
            self.epochs += 1
            self._harvest(p)
        with self._lock:
            nwait = len(self.waiting)
        nparked = sum(len(v) for v in self.parked.values())
        return nwait + nparked + sum(len(p.running) for p in self.pools.values() if isinstance(p, _Pool))

    def run_until_idle(self, max_rounds: int = 10 ** 9):
        for _ in range(max_rounds):
            if self.step() == 0:
                return
            with self._lock:
                only_future = self.waiting and all(r.arrival > _time.perf_counter() for r in self.waiting)
            busy = any(isinstance(p, _Pool) and (p.running or p.in_epoch) for p in self.pools.values())
            if only_future and not busy:
                _time.sleep(max(0.0, min(r.arrival for r in self.waiting) - _time.perf_counter()))

    # ---- background loop ----------------------------------------------------------------
    def start(self):
        def loop():
            while not self._stop.is_set():
                if self.step() == 0:
                    _time.sleep(0.0005)

        self._thread = threading.Thread(target=loop, daemon=True)
        self._thread.start()

    def stop(self):
        self._stop.set()
        if self._thread is not None:
            self._thread.join()

    # ---- metrics ------------------------------------------------------------------------
    def latency_percentiles(self, qs=(50, 99)):
        lat = np.array([r.latency for r in self.done])
        return {f"p{q}": float(np.percentile(lat, q)) for q in qs} if lat.size else {}
