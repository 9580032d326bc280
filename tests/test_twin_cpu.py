"""The twin and quad forms on the CPU: the same step as the unrolled form, bit for bit.

The twin form (two lanes per instance, the isomorphic right-hand-side terms split
between them, partner values exchanged) and the quad form (four lanes: each primary's
pull split again into its x and y components, the state held in lane-relative order)
are compiled here with g++, one thread per lane and a barrier exchange standing in for
the DPP broadcasts; the unrolled form's step is compiled the same way with one thread.
All are built without implicit fusion, as on the GPU. For the planar CR3BP (twin and
quad) and a pendulum pulled by two phase-shifted sines (twin: sin / cos pairs,
parameter bindings), every lane must end with exactly the lane form's state, step size
and Taylor coefficients.
"""

import ctypes as ct
import subprocess

import numpy as np
import pytest

import heyoka_amd as hy
from heyoka_amd import _lib
from heyoka_amd import workloads as W

COMMON = r"""
#include <cmath>
#include <cstdint>
#include <math.h>
#include <pthread.h>
#define __device__
#define __forceinline__ inline
#define __constant__ static const
#define HYX_CPU_EMU 1
using std::sqrt; using std::pow; using std::exp; using std::log; using std::fabs; using std::isfinite;
"""

TWIN = COMMON + r"""
#define NL 2
static pthread_barrier_t bar;
static thread_local int tl_lane;
static double xch[NL];
static double bcast(double v, int from) { xch[tl_lane] = v; pthread_barrier_wait(&bar); double r = xch[from]; pthread_barrier_wait(&bar); return r; }
static bool hyx_twin_lane_b() { return (tl_lane & 1) != 0; }
static double hyx_twin_from_a(double v) { return bcast(v, 0); }
static double hyx_twin_from_b(double v) { return bcast(v, 1); }
#include "prelude.inc"
struct Arg { int lane; const double *x; const double *pv; double t0, mdt; double *xout; double *h; int *nf; double *tc; };
static void *run(void *p)
{
    Arg *a = (Arg *)p;
    tl_lane = a->lane;
    double x[HYX_N], pv[HYX_NPAR];
    for (int i = 0; i < HYX_N; ++i) x[i] = a->x[i];
    for (int i = 0; i < HYX_NPAR; ++i) pv[i] = a->pv[i];
    double h; bool nf;
    hyx_taylor_step<true>(x, pv, a->t0, a->mdt, h, nf, a->tc + (size_t)a->lane * HYX_N * (HYX_P + 1), 1, nullptr, 1);
    for (int i = 0; i < HYX_N; ++i) a->xout[a->lane * HYX_N + i] = x[i];
    a->h[a->lane] = h; a->nf[a->lane] = nf;
    return nullptr;
}
extern "C" int step2(const double *x, const double *pv, double t0, double mdt, double *xout, double *h, int *nf, double *tc)
{
    pthread_barrier_init(&bar, nullptr, NL);
    pthread_t th[NL]; Arg args[NL];
    for (int l = 0; l < NL; ++l) { args[l] = Arg{l, x, pv, t0, mdt, xout, h, nf, tc}; pthread_create(&th[l], nullptr, run, &args[l]); }
    for (int l = 0; l < NL; ++l) pthread_join(th[l], nullptr);
    pthread_barrier_destroy(&bar);
    return 0;
}
"""

# Quad lanes q = 2 * primary + component: from_a / from_b read primary A's / B's lane of
# the same component, hyx_quad_swap the other component's lane of the same primary.
QUAD = TWIN.replace("#define NL 2", "#define NL 4").replace(
    """static bool hyx_twin_lane_b() { return (tl_lane & 1) != 0; }
static double hyx_twin_from_a(double v) { return bcast(v, 0); }
static double hyx_twin_from_b(double v) { return bcast(v, 1); }""",
    """static unsigned hyx_quad_lane() { return (unsigned)tl_lane; }
static double hyx_twin_from_a(double v) { return bcast(v, tl_lane & 1); }
static double hyx_twin_from_b(double v) { return bcast(v, 2 + (tl_lane & 1)); }
static double hyx_quad_swap(double v) { return bcast(v, tl_lane ^ 1); }""")

LANE = COMMON + r"""
static double hyx_xs_emu[64 * 64 * 64];
static inline double *hyx_xs_lane() { return hyx_xs_emu; }
#include "prelude.inc"
extern "C" int step1(const double *xi, const double *pvi, double t0, double mdt, double *xout, double *h, int *nf, double *tc)
{
    double x[HYX_N], pv[HYX_NPAR];
    for (int i = 0; i < HYX_N; ++i) x[i] = xi[i];
    for (int i = 0; i < HYX_NPAR; ++i) pv[i] = pvi[i];
    bool nfb;
    hyx_taylor_step<true>(x, pv, t0, mdt, *h, nfb, tc, 1, nullptr, 1);
    for (int i = 0; i < HYX_N; ++i) xout[i] = x[i];
    *nf = nfb;
    return 0;
}
"""


def _build(src, harness, d, name, fn):
    prelude = src.split("// Hand-written CDNA4 kernels")[0]
    d.mkdir(exist_ok=True)
    (d / "prelude.inc").write_text(prelude)
    (d / f"{name}.cpp").write_text(harness)
    so = d / f"{name}.so"
    subprocess.run(["g++", "-O1", "-std=c++17", "-ffp-contract=off", "-shared", "-fPIC", "-pthread", "-o", str(so),
                    str(d / f"{name}.cpp")], check=True, capture_output=True)
    L = ct.CDLL(str(so))
    P = ct.POINTER(ct.c_double)
    getattr(L, fn).argtypes = [P, P, ct.c_double, ct.c_double, P, P, ct.POINTER(ct.c_int), P]
    return getattr(L, fn)


def _two_sines():
    x, v = hy.make_vars("x", "v")
    return [(x, v), (v, -(hy.par[0] * hy.sin(x + hy.par[1])) - hy.par[2] * hy.sin(x - hy.par[3]) - 0.1 * v)]


def _planar(force, drag=False):
    # A planar two-primary problem in the rotating frame with another central force law
    # (and optionally a speed-dependent drag, whose |v| is common to both components).
    x, y, vx, vy = hy.make_vars("x", "y", "vx", "vy")
    mu = hy.par[0]
    d1, d2 = x + mu, x - (1.0 - mu)
    f1, f2 = force(d1 ** 2 + y ** 2), force(d2 ** 2 + y ** 2)
    ax = 2.0 * vy + x - (f1 * d1 * (1.0 - mu) + f2 * d2 * mu)
    ay = -2.0 * vx + y - (f1 * y * (1.0 - mu) + f2 * y * mu)
    if drag:
        sp = hy.sqrt(vx ** 2 + vy ** 2)
        ax, ay = ax - 0.01 * sp * vx, ay - 0.01 * sp * vy
    return [(x, vx), (y, vy), (vx, ax), (vy, ay)]


PLANAR = {"planar_r5": lambda: _planar(lambda r2: r2 ** -2.5),
          "planar_exp": lambda: _planar(lambda r2: hy.exp(-r2)),
          "planar_drag": lambda: _planar(lambda r2: r2 ** -1.5, drag=True)}


def _cases(name):
    rng = np.random.default_rng(17)
    if name in PLANAR:
        n = 12
        st = np.stack([rng.uniform(-1.5, 1.5, n), rng.uniform(0.3, 1.5, n) * rng.choice([-1, 1], n),
                       rng.uniform(-0.5, 0.5, n), rng.uniform(-0.5, 0.5, n)])
        st[1, 0] = -0.0  # a signed zero on the y lanes' bare coordinate
        return PLANAR[name](), st, rng.uniform(0.001, 0.05, (1, n))
    if name == "cr3bp":
        wl = W.cr3bp(24)
        st, pa = wl.state, wl.pars
        # a few close to the primaries too
        st = np.concatenate([st, np.array([[-0.02, 0.97, 0.9], [0.05, 0.02, -0.03], [0.3, 0.1, 0.0],
                                           [-0.4, 0.2, 0.5]])], axis=1)
        pa = np.concatenate([pa, np.array([[0.01, 0.04, 0.001]])], axis=1)
        return wl.sys, st, pa
    st = np.stack([rng.uniform(-2, 2, 16), rng.uniform(-1, 1, 16)])
    pa = np.stack([rng.uniform(0.5, 2, 16), rng.uniform(-1, 1, 16), rng.uniform(0.5, 2, 16), rng.uniform(-1, 1, 16)])
    return _two_sines(), st, pa


@pytest.mark.parametrize("name,lanes,ha", [("cr3bp", 2, False), ("two_sines", 2, False), ("cr3bp", 4, False),
                                           ("planar_r5", 4, False), ("planar_exp", 4, False),
                                           ("planar_drag", 4, False), ("cr3bp", 4, True)])
def test_twin_step_is_the_lane_step_bitwise(name, lanes, ha, tmp_path):
    sys_, st, pa = _cases(name)
    dec = hy.taylor_decompose(sys_)
    p = 19
    s = _lib.System(dec, p, compact=False, source_only=True, quad=lanes == 4, high_accuracy=ha)
    tsrc = s.twin_source()
    assert tsrc and f"#define HYX_SUB {lanes}" in tsrc and "hyx_twin_from_b" in tsrc
    assert ("hyx_quad_swap" in tsrc) == (lanes == 4)
    assert "#pragma clang fp contract(off)" in s.source()
    f2 = _build(tsrc, TWIN if lanes == 2 else QUAD, tmp_path / "twin", "twin", "step2")
    f1 = _build(s.source(), LANE, tmp_path / "lane", "lane", "step1")
    n = dec.n_state
    ptr = lambda a: a.ctypes.data_as(ct.POINTER(ct.c_double))
    for b in range(st.shape[1]):
        x = np.ascontiguousarray(st[:, b])
        pv = np.ascontiguousarray(pa[:, b])
        for mdt in (np.inf, 0.01, -np.inf):
            x1, h1, nf1, tc1 = np.zeros(n), np.zeros(1), ct.c_int(), np.zeros(n * (p + 1))
            f1(ptr(x), ptr(pv), 0.25, mdt, ptr(x1), ptr(h1), ct.byref(nf1), ptr(tc1))
            L = lanes
            x2, h2, nf2, tc2 = np.zeros(L * n), np.zeros(L), np.zeros(L, dtype=np.int32), np.zeros(L * n * (p + 1))
            f2(ptr(x), ptr(pv), 0.25, mdt, ptr(x2), ptr(h2), nf2.ctypes.data_as(ct.POINTER(ct.c_int)), ptr(tc2))
            for lane in range(L):
                assert np.array_equal(x2[lane * n:(lane + 1) * n], x1), (b, lane)
                assert h2[lane] == h1[0] and nf2[lane] == nf1.value
                assert np.array_equal(tc2[lane * n * (p + 1):(lane + 1) * n * (p + 1)], tc1), (b, lane)


def test_quad_form_applies_where_the_state_pairs():
    # The two-sines pendulum has a twin form but no pairing of its state variables (x' = v
    # is not isomorphic to v' = ...): its tail form stays the twin.
    s = _lib.System(hy.taylor_decompose(_two_sines()), 19, compact=False, source_only=True)
    assert "#define HYX_SUB 2" in s.twin_source()
    s = _lib.System(hy.taylor_decompose(W.cr3bp(4).sys), 19, compact=False, source_only=True)
    assert "#define HYX_SUB 4" in s.twin_source()


def test_no_quad_without_a_state_pairing():
    # A perturbation on the x component only (or a different pull law per component) breaks
    # the x <-> y pairing: the tail form stays the twin.
    x, y, vx, vy = hy.make_vars("x", "y", "vx", "vy")
    mu = hy.par[0]
    d1, d2 = x + mu, x - (1.0 - mu)
    f1, f2 = (d1 ** 2 + y ** 2) ** -1.5, (d2 ** 2 + y ** 2) ** -1.5
    ax = 2.0 * vy + x - (f1 * d1 * (1.0 - mu) + f2 * d2 * mu)
    ay = -2.0 * vx + y - (f1 * y * (1.0 - mu) + f2 * y * mu)
    for sys_ in ([(x, vx), (y, vy), (vx, ax + 0.01 * x * x), (vy, ay)],
                 [(x, vx), (y, vy), (vx, ax), (vy, -2.0 * vx + y - (f1 * y * (1.0 - mu) + f2 * y * y * mu))]):
        s = _lib.System(hy.taylor_decompose(sys_), 19, compact=False, source_only=True)
        assert "#define HYX_SUB 2" in s.twin_source()


def test_no_twin_without_isomorphic_terms():
    for sys_ in (W.pendulum().sys, hy.model.j2()):
        s = _lib.System(hy.taylor_decompose(sys_), 19, compact=False, source_only=True)
        assert s.twin_source() == "" and "fp contract(off)" not in s.source()
