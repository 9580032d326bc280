"""The generated Taylor-step function, compiled for the CPU, against the oracle.

libhyx.so generates HIP source for hyx_taylor_step<WTC>() from the op-list. The
same function body is plain C++ apart from a few device qualifiers, so it is
compiled here with g++ (qualifiers defined away, fma via the C library) and one
step is compared with the oracle's: Taylor coefficients, h and the new state.
This checks the code generator on the CPU, for both code forms, without a GPU.
"""

import ctypes as ct
import os
import subprocess

import numpy as np
import pytest

import heyoka_amd as hy
import pyoracle as O
from heyoka_amd import _lib
from heyoka_amd import workloads as W

HARNESS = r"""
#include <cmath>
#include <cstdint>
#include <math.h>  // sincos (GNU)
#define __device__
#define __forceinline__ inline
#define __constant__ static const
#define HYX_CPU_EMU 1
using std::sqrt; using std::pow; using std::exp; using std::log; using std::fabs; using std::isfinite;
// LDS state rows of the lane form: lane 0 of a 64-lane column layout.
static double hyx_xs_emu[64 * 64 * 64];
static inline double *hyx_xs_lane() { return hyx_xs_emu; }

#include "prelude.inc"
extern "C" int step_cpu(const double *x_in, const double *pv_in, double t0, double max_dt, double *x_out,
                        double *h, int *nf, double *tc, double *tape)
{
    double x[HYX_N];
    double pv[HYX_NPAR];
    for (int i = 0; i < HYX_N; ++i) x[i] = x_in[i];
    for (int i = 0; i < HYX_NPAR; ++i) pv[i] = pv_in[i];
    bool nfb = false;
    hyx_taylor_step<true>(x, pv, t0, max_dt, *h, nfb, tc, 1, tape, 1);
    for (int i = 0; i < HYX_N; ++i) x_out[i] = x[i];
    *nf = nfb;
    return 0;
}
"""


def _build(src, tmp_path, tag):
    prelude = src.split("// Hand-written CDNA4 kernels")[0]
    prelude = prelude.replace("typedef unsigned long long hyx_u64;", "typedef unsigned long long hyx_u64;\n")
    d = tmp_path / tag
    d.mkdir()
    (d / "prelude.inc").write_text(prelude)
    (d / "h.cpp").write_text(HARNESS)
    so = d / "h.so"
    subprocess.run(["g++", "-O1", "-std=c++17", "-ffp-contract=off", "-shared", "-fPIC", "-o", str(so), str(d / "h.cpp")],
                   check=True, capture_output=True)
    L = ct.CDLL(str(so))
    P = ct.POINTER(ct.c_double)
    L.step_cpu.argtypes = [P, P, ct.c_double, ct.c_double, P, P, ct.POINTER(ct.c_int), P, P]
    return L


CASES = {
    "pendulum": (lambda: W.pendulum(), 0),
    "fdp": (None, 0),
    "cr3bp": (lambda: W.cr3bp(4), 0),
    "nbody6": (lambda: W.nbody6(2), 0),
    "j2var": (lambda: W.j2var(2), 0),
}


def _case(name):
    if name == "fdp":
        x, v = hy.make_vars("x", "v")
        sys_ = [(x, v), (v, hy.cos(hy.time) - hy.par[0] * v - hy.sin(x))]
        return sys_, np.array([[0.03], [1.88]]), np.array([[0.13]]), 0.7, 2.220446049250313e-16
    wl = CASES[name][0]()
    sys_ = wl.sys.sys if hasattr(wl.sys, "n_orig_sv") else wl.sys
    pars = None if wl.pars is None else wl.pars[:, :1]
    return sys_, wl.state[:, :1], pars, 0.0, wl.tol


@pytest.mark.parametrize("compact", [False, True])
@pytest.mark.parametrize("name", list(CASES))
def test_generated_step_matches_oracle(name, compact, tmp_path):
    # One lane per instance here; the compact form's sub-lane split is emulated below.
    if name in ("nbody6", "j2var") and not compact:
        pytest.skip("large systems use the compact form; their unrolled step takes minutes to compile")
    sys_, st, pars, t0, tol = _case(name)
    dec = hy.taylor_decompose(sys_)
    p = hy.order_from_tol(tol)
    s = _lib.System(dec, p, compact=compact, source_only=True, sub=1)
    L = _build(s.source(), tmp_path, f"{name}_{int(compact)}")
    n = dec.n_state
    xin = np.ascontiguousarray(st[:, 0], dtype=np.float64)
    pv = np.ascontiguousarray(pars[:, 0] if pars is not None else np.zeros(1), dtype=np.float64)
    xout = np.zeros(n)
    h = ct.c_double()
    nf = ct.c_int()
    tc = np.zeros(n * (p + 1))
    tape = np.zeros(dec.n_nodes * (p + 1))
    ptr = lambda a: a.ctypes.data_as(ct.POINTER(ct.c_double))
    L.step_cpu(ptr(xin), ptr(pv), t0, np.inf, ptr(xout), ct.byref(h), ct.byref(nf), ptr(tc), ptr(tape))
    r = O.step_batch(dec, p, st, pars=pars, t_hi=[t0], want_tc=True)
    tco = r["tc"][:, :, 0]
    tcg = tc.reshape(n, p + 1)
    scale = np.maximum(np.abs(tco), np.max(np.abs(tco), axis=1, keepdims=True) * 1e-3) + 1e-300
    rel = np.max(np.abs(tcg - tco) / scale)
    assert rel < 1e-11, (name, compact, rel, np.unravel_index(np.argmax(np.abs(tcg - tco) / scale), tco.shape))
    assert abs(h.value - r["h"][0]) <= 1e-13 * abs(r["h"][0])
    np.testing.assert_allclose(xout, r["state"][:, 0], rtol=1e-13, atol=1e-15)
    assert nf.value == 0


GROUP_HARNESS = r"""
#include <cmath>
#include <cstdint>
#include <math.h>
#include <pthread.h>
#define __device__
#define __forceinline__ inline
#define __constant__ static const
using std::sqrt; using std::pow; using std::exp; using std::log; using std::fabs; using std::isfinite;
static pthread_barrier_t bar;
static thread_local int tl_g;
static void __syncthreads() { pthread_barrier_wait(&bar); }
#define HYX_CPU_EMU 1
// The lanes of a wavefront run in lock step on the GPU; threads need a barrier at
// every exchange point, including the one between the owners' loads and the
// clusters' stores of the pipelined schedule.
static void hyx_wsync() { pthread_barrier_wait(&bar); }
#define HYX_EMU_SYNC() hyx_wsync()
static double xch_d[64];
static int xch_i[64];
static double __shfl_xor(double v, int off) { xch_d[tl_g] = v; __syncthreads(); double r = xch_d[tl_g ^ off]; __syncthreads(); return r; }
static int __shfl_xor(int v, int off) { xch_i[tl_g] = v; __syncthreads(); int r = xch_i[tl_g ^ off]; __syncthreads(); return r; }
#include "prelude.inc"
static double fl[HYX_FLAT];
struct Arg { int g; const double *x; const double *pv; double t0, mdt; double *xout; double *h; int *nf; double *tc; };
static void *run(void *p)
{
    Arg *a = (Arg *)p;
    const int g = a->g;
    tl_g = g;
    hyx_gtab t;
    hyx_gtab_load(t, g);
    double xo[HYX_OWN], pv[HYX_NPAR];
    for (int o = 0; o < HYX_OWN; ++o) xo[o] = (t.vo[o] >= 0) ? a->x[t.vo[o]] : 0.0;
    for (int q = 0; q < HYX_NPAR; ++q) pv[q] = a->pv[q];
    double h; bool nf;
    hyx_gstep<true>(xo, pv, a->t0, a->mdt, h, nf, g, fl, t, a->tc, 1);
    for (int o = 0; o < HYX_OWN; ++o) if (t.vo[o] >= 0) a->xout[t.vo[o]] = xo[o];
    a->h[g] = h; a->nf[g] = nf;
    return nullptr;
}
extern "C" int gstep_cpu(const double *x, const double *pv, double t0, double mdt, double *xout, double *h, int *nf, double *tc)
{
    for (int q = 0; q < HYX_FLAT; ++q) fl[q] = 0.0;
    pthread_barrier_init(&bar, nullptr, HYX_G);
    pthread_t th[HYX_G]; Arg args[HYX_G];
    for (int g = 0; g < HYX_G; ++g) { args[g] = Arg{g, x, pv, t0, mdt, xout, h, nf, tc}; pthread_create(&th[g], nullptr, run, &args[g]); }
    for (int g = 0; g < HYX_G; ++g) pthread_join(th[g], nullptr);
    pthread_barrier_destroy(&bar);
    return HYX_G;
}
"""


@pytest.mark.parametrize("pipeline", [True, False])
@pytest.mark.parametrize("name", ["nbody6", "nbody3"])
def test_group_step_matches_oracle(name, pipeline, tmp_path):
    # The group form emulated on the CPU: one thread per lane of a group, a
    # barrier for __syncthreads and a shared array for __shfl_xor.
    if name == "nbody6":
        wl = W.nbody6(2)
        sys_, st = wl.sys, wl.state[:, :1]
    else:
        sys_ = hy.model.nbody(3, masses=[1.0, 0.3, 0.01])
        st = np.array([0, 0, 0, 0, 0, 0, 1, 0, 0.1, 0, 1.1, 0, -2, 0.5, 0, 0.2, -0.6, 0.05], dtype=float)[:, None]
    dec = hy.taylor_decompose(sys_)
    p = 19
    s = _lib.System(dec, p, compact="group", source_only=True, pipeline=pipeline)
    assert s.form == "group"
    src = s.source()
    assert ("#define HYX_PIPELINED 1" in src) == pipeline
    prelude = src.split("// Hand-written CDNA4 kernels")[0]
    d = tmp_path / name
    d.mkdir()
    (d / "prelude.inc").write_text(prelude)
    (d / "g.cpp").write_text(GROUP_HARNESS)
    so = d / "g.so"
    subprocess.run(["g++", "-O1", "-std=c++17", "-ffp-contract=off", "-shared", "-fPIC", "-pthread", "-o", str(so),
                    str(d / "g.cpp")], check=True, capture_output=True)
    L = ct.CDLL(str(so))
    P = ct.POINTER(ct.c_double)
    L.gstep_cpu.argtypes = [P, P, ct.c_double, ct.c_double, P, P, ct.POINTER(ct.c_int), P]
    n = dec.n_state
    x = np.ascontiguousarray(st[:, 0])
    pv = np.zeros(1)
    xout = np.zeros(n)
    h = np.zeros(64)
    nf = np.zeros(64, dtype=np.int32)
    tc = np.zeros(n * (p + 1))
    ptr = lambda a: a.ctypes.data_as(ct.POINTER(ct.c_double))
    G = L.gstep_cpu(ptr(x), ptr(pv), 0.0, np.inf, ptr(xout), ptr(h), nf.ctypes.data_as(ct.POINTER(ct.c_int)), ptr(tc))
    assert np.all(h[:G] == h[0]) and not nf[:G].any()
    r = O.step_batch(dec, p, st, want_tc=True)
    tco = r["tc"][:, :, 0]
    tcg = tc.reshape(n, p + 1)
    scale = np.max(np.abs(tco), axis=1, keepdims=True) + 1e-300
    assert np.max(np.abs(tcg - tco) / scale) < 1e-11
    assert abs(h[0] - r["h"][0]) <= 1e-13 * abs(r["h"][0])
    np.testing.assert_allclose(xout, r["state"][:, 0], rtol=1e-13, atol=1e-15)


SUB_HARNESS = r"""
#include <cmath>
#include <cstdint>
#include <math.h>
#include <pthread.h>
#define __device__
#define __forceinline__ inline
#define __constant__ static const
using std::sqrt; using std::pow; using std::exp; using std::log; using std::fabs; using std::isfinite;
static pthread_barrier_t bar;
static thread_local int tl_lane;
static double xch_d[64];
static unsigned __lane_id() { return (unsigned)tl_lane; }
static double __shfl_xor(double v, int off) { xch_d[tl_lane] = v; pthread_barrier_wait(&bar); double r = xch_d[tl_lane ^ off]; pthread_barrier_wait(&bar); return r; }
#include "prelude.inc"
struct Arg { int lane; const double *x; const double *pv; double t0, mdt; double *xout; double *h; int *nf; double *tc; double *tape; };
static void *run(void *p)
{
    Arg *a = (Arg *)p;
    tl_lane = a->lane;
    double x[HYX_N], pv[HYX_NPAR];
    for (int i = 0; i < HYX_N; ++i) x[i] = a->x[i];
    for (int i = 0; i < HYX_NPAR; ++i) pv[i] = a->pv[i];
    double h; bool nf;
    hyx_taylor_step<true>(x, pv, a->t0, a->mdt, h, nf, a->tc + (size_t)a->lane * HYX_N * (HYX_P + 1), 1, a->tape, 1);
    for (int i = 0; i < HYX_N; ++i) a->xout[a->lane * HYX_N + i] = x[i];
    a->h[a->lane] = h; a->nf[a->lane] = nf;
    return nullptr;
}
extern "C" int sub_step_cpu(const double *x, const double *pv, double t0, double mdt, double *xout, double *h, int *nf,
                            double *tc, double *tape)
{
    pthread_barrier_init(&bar, nullptr, HYX_SUB);
    pthread_t th[HYX_SUB]; Arg args[HYX_SUB];
    for (int l = 0; l < HYX_SUB; ++l) { args[l] = Arg{l, x, pv, t0, mdt, xout, h, nf, tc, tape}; pthread_create(&th[l], nullptr, run, &args[l]); }
    for (int l = 0; l < HYX_SUB; ++l) pthread_join(th[l], nullptr);
    pthread_barrier_destroy(&bar);
    return HYX_SUB;
}
"""


@pytest.mark.parametrize("sub", [2, 4, 8])
@pytest.mark.parametrize("name", ["cr3bp", "j2var"])
def test_compact_sublanes_match_oracle(name, sub, tmp_path):
    # The compact form with `sub` lanes per instance, one thread per lane: every
    # lane must end with the same (bit-identical) step, equal to the oracle's.
    sys_, st, pars, t0, tol = _case(name)
    dec = hy.taylor_decompose(sys_)
    p = hy.order_from_tol(tol)
    src = _lib.System(dec, p, compact=True, source_only=True, sub=sub).source()
    assert f"#define HYX_SUB {sub}" in src
    prelude = src.split("// Hand-written CDNA4 kernels")[0]
    d = tmp_path / f"{name}_{sub}"
    d.mkdir()
    (d / "prelude.inc").write_text(prelude)
    (d / "s.cpp").write_text(SUB_HARNESS)
    so = d / "s.so"
    subprocess.run(["g++", "-O1", "-std=c++17", "-ffp-contract=off", "-shared", "-fPIC", "-pthread", "-o", str(so),
                    str(d / "s.cpp")], check=True, capture_output=True)
    L = ct.CDLL(str(so))
    P = ct.POINTER(ct.c_double)
    L.sub_step_cpu.argtypes = [P, P, ct.c_double, ct.c_double, P, P, ct.POINTER(ct.c_int), P, P]
    n = dec.n_state
    x = np.ascontiguousarray(st[:, 0], dtype=np.float64)
    pv = np.ascontiguousarray(pars[:, 0] if pars is not None else np.zeros(1), dtype=np.float64)
    xout = np.zeros(sub * n)
    h = np.zeros(sub)
    nf = np.zeros(sub, dtype=np.int32)
    tc = np.zeros(sub * n * (p + 1))
    tape = np.zeros(dec.n_nodes * (p + 1))
    ptr = lambda a: a.ctypes.data_as(ct.POINTER(ct.c_double))
    L.sub_step_cpu(ptr(x), ptr(pv), t0, np.inf, ptr(xout), ptr(h), nf.ctypes.data_as(ct.POINTER(ct.c_int)), ptr(tc),
                   ptr(tape))
    xo = xout.reshape(sub, n)
    assert np.all(xo == xo[0]) and np.all(h == h[0]) and not nf.any()
    r = O.step_batch(dec, p, st, pars=pars, t_hi=[t0], want_tc=True)
    assert abs(h[0] - r["h"][0]) <= 1e-13 * abs(r["h"][0])
    np.testing.assert_allclose(xo[0], r["state"][:, 0], rtol=1e-13, atol=1e-15)
    tcg = tc.reshape(sub, n, p + 1)[0]
    tco = r["tc"][:, :, 0]
    scale = np.maximum(np.abs(tco), np.max(np.abs(tco), axis=1, keepdims=True) * 1e-3) + 1e-300
    assert np.max(np.abs(tcg - tco) / scale) < 1e-11


def _mixed_case():
    # Every recurrence kind of the wave form: sqrt, div, exp, log, pow, sin / cos, time,
    # parameters and constants, linear nodes.
    x, v, y = hy.make_vars("x", "v", "y")
    sys_ = [(x, v),
            (v, -hy.sqrt(x * x + 1.5) / (1.0 + y * y) + hy.exp(-x) * hy.par[0] + hy.log(2.0 + y * y) * hy.time),
            (y, hy.cos(x) * hy.par[1] - (v * v + 0.5) ** 1.5 + hy.time)]
    return sys_, np.array([[0.3], [0.1], [0.7]]), np.array([[0.4], [1.3]]), 0.5, 1e-15


def _chains_case():
    # Derived chains of every shape the wave form fuses: products with constants of
    # products with constants, negations, constant + node, sums of a node and constants,
    # chains of different shapes from nodes of one kind and one level (merged templates
    # with steps some lanes skip), and a chain whose end is a right-hand side.
    x, y, z = hy.make_vars("x", "y", "z")
    a = x * y
    b = y * z
    c = x * z
    sys_ = [(x, -(5.0 * (2.0 * a)) + 3.0 * c),
            (y, (1.0 + 0.5 * b) * x - (-(3.0 * c)) + 2.0),
            (z, (2.0 * (-1.5 * (0.25 * a))) * hy.sin(y) - 0.5 * b)]
    return sys_, np.array([[0.3], [-0.2], [0.4]]), None, 0.1, 1e-15


@pytest.mark.parametrize("form,split", [("wave", False), ("wave32", False), ("wave16", False), ("wave32", True)])
@pytest.mark.parametrize("name", ["pendulum", "cr3bp", "j2var", "mixed", "chains"])
def test_wave_step_matches_oracle(name, form, split, tmp_path):
    # The wave form emulated on the CPU with the group harness: one thread per lane of
    # a group, a barrier at every wavefront-scope fence and __shfl_xor exchange. With
    # split, the long convolution sums run as two interleaved accumulators.
    sys_, st, pars, t0, tol = (_mixed_case() if name == "mixed" else _chains_case() if name == "chains"
                               else _case(name))
    dec = hy.taylor_decompose(sys_)
    p = hy.order_from_tol(tol)
    s = _lib.System(dec, p, compact=form, source_only=True, split_chains=split)
    if name == "chains":
        assert "const double dv1 = " in s.source()  # chains of at least two derived steps
    if split:
        assert "double e = " in s.source()
    assert s.form == "wave"
    prelude = s.source().split("// Hand-written CDNA4 kernels")[0]
    d = tmp_path / f"{name}_{form}_{int(split)}"
    d.mkdir()
    (d / "prelude.inc").write_text(prelude)
    (d / "g.cpp").write_text(GROUP_HARNESS)
    so = d / "g.so"
    subprocess.run(["g++", "-O1", "-std=c++17", "-ffp-contract=off", "-shared", "-fPIC", "-pthread", "-o", str(so),
                    str(d / "g.cpp")], check=True, capture_output=True)
    L = ct.CDLL(str(so))
    P = ct.POINTER(ct.c_double)
    L.gstep_cpu.argtypes = [P, P, ct.c_double, ct.c_double, P, P, ct.POINTER(ct.c_int), P]
    n = dec.n_state
    x = np.ascontiguousarray(st[:, 0], dtype=np.float64)
    pv = np.ascontiguousarray(pars[:, 0] if pars is not None else np.zeros(1), dtype=np.float64)
    xout = np.zeros(n)
    h = np.zeros(64)
    nf = np.zeros(64, dtype=np.int32)
    tc = np.zeros(n * (p + 1))
    ptr = lambda a: a.ctypes.data_as(ct.POINTER(ct.c_double))
    G = L.gstep_cpu(ptr(x), ptr(pv), t0, np.inf, ptr(xout), ptr(h), nf.ctypes.data_as(ct.POINTER(ct.c_int)), ptr(tc))
    assert G == {"wave": 64, "wave32": 32, "wave16": 16}[form]
    assert np.all(h[:G] == h[0]) and not nf[:G].any()
    r = O.step_batch(dec, p, st, pars=pars, t_hi=[t0], want_tc=True)
    tco = r["tc"][:, :, 0]
    tcg = tc.reshape(n, p + 1)
    scale = np.maximum(np.abs(tco), np.max(np.abs(tco), axis=1, keepdims=True) * 1e-3) + 1e-300
    assert np.max(np.abs(tcg - tco) / scale) < 1e-11
    assert abs(h[0] - r["h"][0]) <= 1e-13 * abs(r["h"][0])
    np.testing.assert_allclose(xout, r["state"][:, 0], rtol=1e-13, atol=1e-15)


def test_wave_schedule_j2var():
    # The wave form's static schedule of the J2 + variational system at 32 lanes per
    # instance. A linear node or a product with a constant whose only node operand is
    # another node is computed by that node's lane right after it (a derived step, no level
    # of its own), and nodes move within their slack so that a level holds few slots: the
    # 20 dependency levels per order become 14, in 15 slots. The arithmetic is checked
    # against the oracle by test_wave_step_matches_oracle.
    import re

    dec = hy.taylor_decompose(W.j2var(8).sys.sys)
    src = _lib.System(dec, 19, compact="wave32", source_only=True).source()
    assert int(re.search(r"#define HYX_WAVE_SLOTS (\d+)", src).group(1)) == 15
    assert "const double dv0 = " in src
    order10 = src.split("// order 10\n")[1].split("// order 11\n")[0]
    assert order10.count("hyx_wsync();") == 14 + 2  # levels, plus the state rows' two exchanges


def _bank_passes(src):
    """Serialised LDS passes per instance-step of the group form's table-driven loads: a
    64-bit read is served 16 lanes per pass, a double at word a in banks 2a, 2a+1 of 32, so
    lanes with distinct addresses equal mod 16 serialise."""
    import re

    def tab(name):
        m = re.search(r"hyx_tab_%s\[\d+\] = \{([^}]*)\}" % name, src)
        return [int(x) for x in m.group(1).split(",")] if m else None

    D = {k: int(v) for k, v in re.findall(r"#define (HYX_\w+) (\d+)\n", src)}
    G, XB, OWN, NIN, P = D["HYX_G"], D["HYX_XB"], D["HYX_OWN"], D["HYX_NIN"], D["HYX_P"]

    def cost(addrs):
        banks = {}
        for a in addrs:
            banks.setdefault(a % 16, set()).add(a)
        return max(len(s) for s in banks.values()) - 1

    total = 0
    vc = tab("vc")
    for oo in range(OWN):
        total += cost([XB + vc[g * OWN + oo] for g in range(G)]) * (P + 3)
    lis = tab("lis")
    nls = len(lis) // (G * NIN)
    for q in range(NIN * nls):
        total += cost([XB + lis[g * NIN * nls + q] for g in range(G)]) * P
    for oo in range(OWN):
        for nm, off, w in (("tio%d" % oo, 0, P), ("tis%d" % oo, XB, P + 1)):
            t = tab(nm)
            cnt = len(t) // G if len(t) >= G else 0
            for j in range(cnt):
                total += cost([off + t[g * cnt + j] for g in range(G)]) * w
    return total, D, tab


def test_group_lds_layout_spreads_the_exchange_loads():
    # The 6-body group form: the state columns and output slots are placed so that the
    # 16 lanes of a group hit distinct banks; the identity layout (columns = variable
    # indices, outputs at q * 17 + lane) serialises 257 extra passes per instance-step.
    wl = W.nbody6(2)
    dec = hy.taylor_decompose(wl.sys)
    s = _lib.System(dec, 19, compact="group", source_only=True)
    src = s.source()
    passes, D, tab = _bank_passes(src)
    assert passes <= 100
    n, nx, xb, G, own, nout = D["HYX_N"], D["HYX_NX"], D["HYX_XB"], D["HYX_G"], D["HYX_OWN"], D["HYX_NOUT"]
    vo, vc, wp = tab("vo"), tab("vc"), tab("wp")
    # every variable has its own column; empty slots share the zero column
    cols = {v: c for v, c in zip(vo, vc) if v >= 0}
    assert sorted(cols) == list(range(n)) and len(set(cols.values())) == n
    zero_cols = {c for v, c in zip(vo, vc) if v < 0}
    assert len(zero_cols) <= 1 and not zero_cols & set(cols.values())
    assert all(0 <= c < nx for c in vc)
    # one lane permutation for the output rows, every slot below the x rows
    assert sorted(wp) == list(range(G))
    for oo in range(own):
        t = tab("tio%d" % oo)
        assert all(0 <= a < xb for a in t)
