"""hyx_propagate_until_noalloc through the C-ABI, against hyx_propagate_until.

The caller owns every input and output (device arrays from hipMalloc here); the entry
point's scratch is the library's fixed per-GPU array. Results must equal those of the
allocating entry point bit for bit (the same kernels): on two calls in a row, the second
with fewer instances (the scratch keeps the first call's contents), for the lane form with
parameters and explicit tf_lo / max_delta_t, the group form without parameters, and the
compact form, whose Taylor tape lives in the scratch. A call whose scratch would exceed the
fixed size fails with a clear error before launching anything.
"""

import ctypes as ct

import numpy as np
import pytest

import heyoka_amd as hy
from heyoka_amd import _lib
from heyoka_amd import workloads as W
from heyoka_amd.integrator import get_system

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
        try:
            H = ct.CDLL(name)
            break
        except OSError:
            continue
    H.hipMalloc.argtypes = [ct.POINTER(ct.c_void_p), ct.c_size_t]
    H.hipMalloc.restype = ct.c_int
    H.hipFree.argtypes = [ct.c_void_p]
    H.hipFree.restype = ct.c_int
    H.hipMemcpy.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_size_t, ct.c_int]
    H.hipMemcpy.restype = ct.c_int
    H.hipDeviceSynchronize.restype = ct.c_int
    return H


class _Dev:
    """Device copies of host arrays, owned by the test (the caller)."""

    def __init__(self, H):
        self.H, self.ptrs = H, []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = ct.c_void_p()
        assert self.H.hipMalloc(ct.byref(p), max(a.nbytes, 8)) == 0
        self.ptrs.append(p)
        assert self.H.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
        return p

    def empty(self, nbytes):
        p = ct.c_void_p()
        assert self.H.hipMalloc(ct.byref(p), max(nbytes, 8)) == 0
        self.ptrs.append(p)
        return p

    def get(self, p, shape, dtype):
        out = np.empty(shape, dtype=dtype)
        assert self.H.hipMemcpy(out.ctypes.data, p, out.nbytes, D2H) == 0
        return out

    def free(self):
        for p in self.ptrs:
            self.H.hipFree(p)
        self.ptrs = []


def _reference(system, st, pars, tf_hi, tf_lo, mdt):
    N = st.shape[1]
    b = _lib.Batch(system, N)
    b.write(_lib.FIELD_STATE, st)
    if pars is not None:
        b.write(_lib.FIELD_PARS, pars)
    b.write(_lib.FIELD_TIME_HI, np.zeros(N))
    b.write(_lib.FIELD_TIME_LO, np.zeros(N))
    r = b.propagate_until(tf_hi, tf_lo, mdt, 0, _lib.MODE_PERSISTENT)
    out = dict(state=b.read(_lib.FIELD_STATE, st.shape), t_hi=b.read(_lib.FIELD_TIME_HI, (N,)),
               t_lo=b.read(_lib.FIELD_TIME_LO, (N,)), outcome=r["outcome"], min_h=r["min_h"], max_h=r["max_h"],
               nsteps=r["nsteps"])
    b.close()
    return out


def _noalloc(H, system, st, pars, tf_hi, tf_lo, mdt, max_steps=0):
    L = _lib.lib()
    N = st.shape[1]
    d = _Dev(H)
    try:
        ps = d.put(st)
        pp = d.put(pars) if pars is not None else None
        th, tl = d.put(np.zeros(N)), d.put(np.zeros(N))
        pf = d.put(tf_hi)
        pfl = d.put(tf_lo) if tf_lo is not None else None
        pm = d.put(mdt) if mdt is not None else None
        oc, mn, mx, ns = (d.empty(N * 8) for _ in range(4))
        rc = L.hyx_propagate_until_noalloc(system.h, N, ps, pp, th, tl, pf, pfl, pm, max_steps, oc, mn, mx, ns,
                                           None)
        assert rc == 0, L.hyx_last_error()
        assert H.hipDeviceSynchronize() == 0
        return dict(state=d.get(ps, st.shape, np.float64), t_hi=d.get(th, (N,), np.float64),
                    t_lo=d.get(tl, (N,), np.float64), outcome=d.get(oc, (N,), np.int64),
                    min_h=d.get(mn, (N,), np.float64), max_h=d.get(mx, (N,), np.float64),
                    nsteps=d.get(ns, (N,), np.uint64))
    finally:
        d.free()


def _same(a, b):
    for k in ("state", "t_hi", "t_lo", "outcome", "min_h", "max_h", "nsteps"):
        assert np.array_equal(a[k], b[k]), k


def test_lane_form_two_calls_second_smaller():
    print("[gpu] noalloc: cr3bp lane form, 5000 then 1234 instances", flush=True)
    H = _hip()
    wl = W.cr3bp(5000, log_t=(0.5, 1.7))
    system = get_system(hy.taylor_decompose(wl.sys), 19)
    tf = wl.t_final
    a = _noalloc(H, system, wl.state, wl.pars, tf, None, None)
    _same(a, _reference(system, wl.state, wl.pars, tf, None, None))
    assert np.all(a["outcome"] == int(hy.taylor_outcome.time_limit)) and np.array_equal(a["t_hi"], tf)
    # Second call, fewer instances, explicit tf_lo and max_delta_t.
    n2 = 1234
    st2, pa2 = np.ascontiguousarray(wl.state[:, 1000:1000 + n2]), np.ascontiguousarray(wl.pars[:, 1000:1000 + n2])
    tf2, tl2, md2 = tf[1000:1000 + n2] * 0.5, np.full(n2, 1e-17), np.full(n2, 0.05)
    b = _noalloc(H, system, st2, pa2, tf2, tl2, md2)
    _same(b, _reference(system, st2, pa2, tf2, tl2, md2))
    assert np.all(b["max_h"] <= 0.05)


def test_group_form_without_parameters_and_step_limit():
    print("[gpu] noalloc: nbody6 group form, 700 then 333 instances", flush=True)
    H = _hip()
    wl = W.nbody6(700)
    system = get_system(hy.taylor_decompose(wl.sys), 19)
    assert system.form == "group" and system.npar == 0
    tf = np.full(700, 2.0)
    _same(_noalloc(H, system, wl.state, None, tf, None, None), _reference(system, wl.state, None, tf, None, None))
    st2 = np.ascontiguousarray(wl.state[:, :333])
    tf2 = np.full(333, 1.0)
    _same(_noalloc(H, system, st2, None, tf2, None, None), _reference(system, st2, None, tf2, None, None))


def test_compact_form_tape_in_scratch():
    print("[gpu] noalloc: j2var compact form (tape in the scratch), 300 then 100 instances", flush=True)
    H = _hip()
    wl = W.j2var(300)
    system = get_system(hy.taylor_decompose(wl.sys.sys), 19, compact=True)
    assert system.form == "compact"
    tf = np.full(300, 3.0)
    _same(_noalloc(H, system, wl.state, None, tf, None, None), _reference(system, wl.state, None, tf, None, None))
    st2 = np.ascontiguousarray(wl.state[:, 200:])
    tf2 = np.full(100, 1.5)
    _same(_noalloc(H, system, st2, None, tf2, None, None), _reference(system, st2, None, tf2, None, None))


def test_scratch_limit_rejected():
    print("[gpu] noalloc: beyond the fixed scratch", flush=True)
    H = _hip()
    L = _lib.lib()
    system = get_system(hy.taylor_decompose(hy.model.pendulum()), 19)
    d = _Dev(H)
    try:
        p = d.empty(64)
        N = 4_000_000  # about 36 bytes per instance with default tf_lo, max_delta_t and parameter rows
        rc = L.hyx_propagate_until_noalloc(system.h, N, p, None, p, p, p, None, None, 0, p, p, p, p, None)
        assert rc == -1
        msg = L.hyx_last_error().decode()
        assert "HYX_NOALLOC_SCRATCH_BYTES" in msg and str(N) in msg, msg
    finally:
        d.free()
