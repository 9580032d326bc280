// Event detection inside the Taylor step (lane forms), compiled after
// hyx_kcommon.hip when a system has events. The generated event prelude defines
//   HYX_NEV, HYX_NTE            events, terminal ones first
//   hyx_ev_dir[], hyx_ev_cd[]   direction of every event, cooldown of every terminal event
//   HYX_TOL                     the integrator's tolerance
//   hyx_ev_coeff(e, k, tc, ld)  coefficient k of event e, from its rows in tc
//   hyx_ev_norms(...)           compact form: the event jet from the tape, folded into
//                               the step-size norms of the last two orders
// The unrolled step forms the event jet in registers itself, folds it into its
// norms, runs the fast exclusion test per event and writes the coefficients to tc
// only for the events (bits of hyx_evm) the test does not clear.
//
// Semantics follow heyoka's events (heyoka/taylor_expose_events.cpp:195-300; the
// batch tests heyoka/test.py:693-1000 and heyoka/_test_batch_integrator.py:555-600;
// doc/notebooks/Event detection.ipynb):
//   - the roots of an event's Taylor polynomial g(tau) in the step, tau from 0
//     (included) towards h (excluded), are the event's triggers; d_sgn is the
//     sign of dg/dt there, and an event with a direction keeps only the roots of
//     that sign;
//   - the earliest terminal trigger truncates the step: the state is evaluated
//     at its tau, and the event enters a cooldown (its own value, or
//     10 * tol * max|x| / |dg/dt| when negative), during which its roots are
//     ignored; cooldowns expire once the time has moved that far away;
//   - non-terminal triggers before the end of the (truncated) step are recorded
//     for the host, which runs their callbacks in chronological order.
//
// In propagations with a.ev_pause set, a lane whose step needs the host (a
// non-terminal trigger, or a terminal one whose event has a callback) is parked
// at the end of that step (status HYX_ST_PARKED); the host runs the callbacks and
// resumes it. Without a.ev_pause every terminal trigger stops the lane.
//
// Root isolation (per lane, only when a fast exclusion test fails): the
// polynomial is rescaled to s in [0, 1] (tau = s h) and dyadic subintervals are
// visited left to right; on each, Descartes' rule of signs counts the sign
// variations of (1 + x)^P r(1 / (1 + x)), r the polynomial moved onto the
// subinterval. 0 variations: no root; 1: exactly one, refined by the Illinois
// variant of regula falsi on g(tau); more: bisect. The traversal keeps only the
// current level and index, so no stack and no per-lane scratch are needed.

#define HYX_EV_MAX_LVL ((44 < 960 / HYX_P) ? 44 : (960 / HYX_P))
#define HYX_EV_MAX_VISITS 4096

// Coefficients of q(x + a) from those of q (repeated synthetic division).
__device__ __forceinline__ void hyx_ev_shift(double (&q)[HYX_P + 1], double a)
{
#pragma unroll
    for (int i = 0; i < HYX_P; ++i)
#pragma unroll
        for (int k = HYX_P - 1; k >= i; --k) q[k] = __builtin_fma(a, q[k + 1], q[k]);
}

__device__ __forceinline__ double hyx_ev_eval(const double (&g)[HYX_P + 1], double x)
{
    double r = g[HYX_P];
#pragma unroll
    for (int k = HYX_P - 1; k >= 0; --k) r = __builtin_fma(r, x, g[k]);
    return r;
}

// dg/dtau.
__device__ __forceinline__ double hyx_ev_deval(const double (&g)[HYX_P + 1], double x)
{
    double r = (double)HYX_P * g[HYX_P];
#pragma unroll
    for (int k = HYX_P - 1; k >= 1; --k) r = __builtin_fma(r, x, (double)k * g[k]);
    return r;
}

__device__ __forceinline__ int hyx_ev_sgn(double v) { return (v > 0.0) - (v < 0.0); }

// Sign variations of (1 + x)^P r(1 / (1 + x)), r(u) = c(a + w u): bounds the
// number of roots of c in (a, a + w), with the same parity. *ca = c(a).
__device__ __forceinline__ int hyx_ev_variations(const double (&c)[HYX_P + 1], double a, double w, double &ca)
{
    double r[HYX_P + 1];
#pragma unroll
    for (int k = 0; k <= HYX_P; ++k) r[k] = c[k];
    if (a != 0.0) hyx_ev_shift(r, a);
    ca = r[0];
    double q[HYX_P + 1];
    double wk = 1.0;
#pragma unroll
    for (int k = 0; k <= HYX_P; ++k) {
        q[HYX_P - k] = r[k] * wk;
        wk *= w;
    }
    hyx_ev_shift(q, 1.0);
    int v = 0, last = 0;
#pragma unroll
    for (int k = 0; k <= HYX_P; ++k) {
        const int s = hyx_ev_sgn(q[k]);
        if (s != 0) {
            v += (last != 0 && s != last) ? 1 : 0;
            last = s;
        }
    }
    return v;
}

// The root of g between x0 and x1 (f0 = g(x0), f1 = g(x1) of opposite signs,
// neither zero): Illinois iteration, bisection when the secant point leaves the
// bracket; stops at adjacent doubles or an exact zero.
__device__ __forceinline__ double hyx_ev_refine(const double (&g)[HYX_P + 1], double x0, double x1, double f0,
                                                double f1)
{
    int side = 0;
    for (int it = 0; it < 160; ++it) {
        const double m = 0.5 * (x0 + x1);
        if (m == x0 || m == x1) break;
        double x = x0 - f0 * (x1 - x0) / (f1 - f0);
        const double lo = fmin(x0, x1), hi = fmax(x0, x1);
        if (!(x > lo && x < hi)) x = m;
        const double fx = hyx_ev_eval(g, x);
        if (fx == 0.0) return x;
        if ((fx > 0.0) == (f1 > 0.0)) {
            x1 = x;
            f1 = fx;
            if (side == 1) f0 *= 0.5;
            side = 1;
        } else {
            x0 = x;
            f0 = fx;
            if (side == -1) f1 *= 0.5;
            side = -1;
        }
    }
    return (fabs(f0) <= fabs(f1)) ? x0 : x1;
}

// Roots of g(tau) for tau from 0 (included) towards h (excluded), in
// chronological order: on_root(tau, d_sgn) returns true to stop the scan.
template <class F>
__device__ __forceinline__ void hyx_ev_scan(const double (&g)[HYX_P + 1], double h, F &&on_root)
{
    double c[HYX_P + 1];
    double hk = 1.0, mag = 0.0, lo = 0.0, hi = 0.0;
#pragma unroll
    for (int k = 0; k <= HYX_P; ++k) {
        c[k] = g[k] * hk;
        hk *= h;
        mag += fabs(c[k]);
        if (k > 0) {
            lo += fmin(c[k], 0.0);
            hi += fmax(c[k], 0.0);
        }
    }
    if (!(mag > 0.0) || !isfinite(mag)) return;
    // Fast exclusion: q(s) lies in c0 + [lo, hi] over s in [0, 1] (with a rounding margin).
    const double slack = (4.0 * (HYX_P + 1)) * 1.1102230246251565e-16 * mag;
    if (c[0] + lo > slack || c[0] + hi < -slack) return;
    double last = __builtin_nan("");
    if (c[0] == 0.0) {
        last = 0.0;
        if (on_root(0.0, hyx_ev_sgn(g[1]))) return;
    }
    int lvl = 0;
    hyx_u64 idx = 0;
    bool fresh = false; // arrived at this interval from its left neighbour
    for (int visit = 0; visit < HYX_EV_MAX_VISITS; ++visit) {
        const double w = __builtin_ldexp(1.0, -lvl), a = (double)idx * w;
        double ca;
        const int v = hyx_ev_variations(c, a, w, ca);
        if (fresh && ca == 0.0) {
            const double tau = a * h;
            if (tau != last) {
                last = tau;
                if (on_root(tau, hyx_ev_sgn(hyx_ev_deval(g, tau)))) return;
            }
        }
        bool split = v > 1;
        if (v == 1) {
            // One root inside: refine when the end values bracket it; an exact zero at an
            // end (reported as the left end of an interval) or a rounding-level sign
            // disagreement is resolved by bisecting further.
            const double ta = a * h, tb = (a + w) * h;
            const double fa = hyx_ev_eval(g, ta), fb = hyx_ev_eval(g, tb);
            if (fa != 0.0 && fb != 0.0 && (fa > 0.0) != (fb > 0.0)) {
                const double tau = hyx_ev_refine(g, ta, tb, fa, fb);
                if (tau != last && fabs(tau) < fabs(h)) {
                    last = tau;
                    if (on_root(tau, hyx_ev_sgn(hyx_ev_deval(g, tau)))) return;
                }
            } else {
                split = true;
            }
        }
        const bool descend = split && lvl < HYX_EV_MAX_LVL;
        if (descend) {
            ++lvl;
            idx <<= 1;
            fresh = false;
            continue;
        }
        while (lvl > 0 && (idx & 1ull)) {
            idx >>= 1;
            --lvl;
        }
        if (lvl == 0) return;
        ++idx;
        fresh = true;
    }
}

// What the root scan of one step found: the step actually taken (h, or the offset of
// the earliest accepted terminal trigger), that event (-1: none) with its direction
// and cooldown, and the number of non-terminal triggers.
struct hyx_ev_hit {
    double hend, cd;
    int te, te_dsgn;
    unsigned ntc;
};

// The root scan of the events whose bit is set in evm (only on steps the fast
// exclusion test cannot clear). tc, evcd, nt_time, nt_code: the instance's columns
// (stride ld). Measured: taking it out of line does not lower the per-step cost
// and makes trigger steps slower, so it stays inline.
__device__ __forceinline__ hyx_ev_hit hyx_ev_scan_all(const double *tc, const double *evcd, double *nt_time,
                                                              hyx_i64 *nt_code, unsigned nt_cap, hyx_u64 ld,
                                                              double thi, double tlo, double h,
                                                              unsigned long long evm, double mx)
{
    hyx_ev_hit r;
    r.hend = h;
    r.cd = 0.0;
    r.te = -1;
    r.te_dsgn = 0;
    r.ntc = 0;
    double g[HYX_P + 1];
    double best = fabs(h);
    for (int e = 0; e < HYX_NTE; ++e) {
        if (!((evm >> e) & 1ull)) continue;
#pragma unroll
        for (int k = 0; k <= HYX_P; ++k) g[k] = hyx_ev_coeff(e, k, tc, ld);
        const double cdl = evcd[((hyx_u64)e * 3 + 2) * ld];
        const double dt0 =
            (cdl >= 0.0) ? hyx_dfl_sub_hi(thi, tlo, evcd[((hyx_u64)e * 3) * ld], evcd[((hyx_u64)e * 3 + 1) * ld]) : 0.0;
        const int dir = hyx_ev_dir[e];
        hyx_ev_scan(g, h, [&](double tau, int ds) {
            if (fabs(tau) >= best) return true;
            if (dir != 0 && ds != dir) return false;
            if (cdl >= 0.0 && fabs(dt0 + tau) < cdl) return false;
            best = fabs(tau);
            r.te = e;
            r.te_dsgn = ds;
            r.hend = tau;
            return true;
        });
    }
    for (int e = HYX_NTE; e < HYX_NEV; ++e) {
        if (!((evm >> e) & 1ull)) continue;
#pragma unroll
        for (int k = 0; k <= HYX_P; ++k) g[k] = hyx_ev_coeff(e, k, tc, ld);
        const int dir = hyx_ev_dir[e];
        hyx_ev_scan(g, h, [&](double tau, int ds) {
            if (fabs(tau) >= best) return true;
            if (dir != 0 && ds != dir) return false;
            if (r.ntc < nt_cap) {
                double th = thi, tl = tlo;
                hyx_dfl_add(th, tl, tau, 0.0);
                nt_time[(hyx_u64)r.ntc * ld] = th;
                nt_code[(hyx_u64)r.ntc * ld] = 4 * (e - HYX_NTE) + ds + 1;
            }
            ++r.ntc;
            return false;
        });
    }
    if (r.te >= 0) {
#pragma unroll
        for (int k = 0; k <= HYX_P; ++k) g[k] = hyx_ev_coeff(r.te, k, tc, ld);
        double cd = hyx_ev_cd[r.te];
        if (cd < 0.0) {
            cd = 10.0 * (HYX_TOL * mx) / fabs(hyx_ev_deval(g, r.hend));
            if (!isfinite(cd)) cd = 0.0;
        }
        r.cd = cd;
    }
    return r;
}

// Event detection after a step of size h from (thi, tlo). Bit e of evm is set for
// the events whose fast exclusion test the step could not clear; their
// coefficients are in tc (column i). Records the triggers, updates the cooldowns
// and returns the step actually taken: h, or the offset hend of the earliest
// terminal trigger (te / te_dsgn: that event, -1 if none). mx: max |x| at the start
// of the step (the automatic cooldown); ntc_out: the non-terminal triggers found,
// whose times and codes go to a.nt_time / a.nt_code (the first a.nt_cap of them).
// Every lane holding the instance computes the same values and writes the same
// records. The caller re-evaluates the state at hend when te >= 0.
// Terminal events whose cooldown is running, from the instance's cooldown columns (a
// bit per event); the instance keeps the mask in registers between steps.
__device__ __forceinline__ unsigned long long hyx_ev_cdmask(const hyx_kargs &a, hyx_u64 i)
{
    unsigned long long m = 0ull;
    for (int e = 0; e < HYX_NTE; ++e)
        if (a.evcd[((hyx_u64)e * 3 + 2) * a.ld + i] >= 0.0) m |= 1ull << e;
    return m;
}

__device__ __forceinline__ double hyx_ev_detect(const hyx_kargs &a, hyx_u64 i, double thi, double tlo, double h,
                                                unsigned long long evm, double mx, int &te, int &te_dsgn,
                                                unsigned &ntc_out, unsigned long long &cdact, bool always_record)
{
    const hyx_u64 ld = a.ld;
    hyx_ev_hit r;
    r.hend = h;
    r.cd = 0.0;
    r.te = -1;
    r.te_dsgn = 0;
    r.ntc = 0;
    if (h != 0.0 && evm != 0ull)
        r = hyx_ev_scan_all(a.tc + i, a.evcd + i, a.nt_time + i, a.nt_code + i, a.nt_cap, ld, thi, tlo, h, evm, mx);
    te = r.te;
    te_dsgn = r.te_dsgn;
    // Cooldowns (only those running, per cdact): expire those the step moved past,
    // then start the triggered one's.
    if (cdact != 0ull || r.te >= 0) {
        double nhi = thi, nlo = tlo;
        hyx_dfl_add(nhi, nlo, r.hend, 0.0);
        for (int e = 0; e < HYX_NTE; ++e) {
            if (!((cdact >> e) & 1ull)) continue;
            const hyx_u64 o = (hyx_u64)e * 3 * ld + i;
            const double cdl = a.evcd[o + 2 * ld];
            if (!(cdl >= 0.0)) {
                cdact &= ~(1ull << e);
            } else if (!(fabs(hyx_dfl_sub_hi(nhi, nlo, a.evcd[o], a.evcd[o + ld])) < cdl)) {
                a.evcd[o + 2 * ld] = -1.0;
                cdact &= ~(1ull << e);
            }
        }
        if (r.te >= 0) {
            const hyx_u64 o = (hyx_u64)r.te * 3 * ld + i;
            a.evcd[o] = nhi;
            a.evcd[o + ld] = nlo;
            a.evcd[o + 2 * ld] = r.cd;
            cdact |= 1ull << r.te;
        }
    }
    // Nothing can be found on a step the exclusion test cleared: the records are read
    // after step() (always written) or for parked instances (which had triggers).
    if (always_record || evm != 0ull) {
        a.te[i] = (r.te >= 0) ? (hyx_i64)(4 * r.te + r.te_dsgn + 1) : -1;
        a.ntc[i] = r.ntc;
    }
    ntc_out = r.ntc;
    return r.hend;
}

// Lane forms: the step left the coefficients of the steps that need a root scan in
// tc (state rows included), so max |x| and the state at a terminal trigger come
// from there.
__device__ __forceinline__ double hyx_ev_step(const hyx_kargs &a, hyx_u64 i, const double *pv, double thi,
                                              double tlo, double h, double (&x)[HYX_N], bool &nf, int &te,
                                              int &te_dsgn, unsigned &ntc_out, unsigned long long evm,
                                              unsigned long long &cdact, bool always_record)
{
    (void)pv;
    const double *tc = a.tc + i;
    const hyx_u64 ld = a.ld;
    double mx = 0.0;
    if (evm != 0ull) {
#pragma unroll 1
        for (int q = 0; q < HYX_N; ++q) {
            const double a0 = fabs(tc[(hyx_u64)q * (HYX_P + 1) * ld]);
            mx = (a0 > mx) ? a0 : mx;
        }
    }
    const double hend = hyx_ev_detect(a, i, thi, tlo, h, evm, mx, te, te_dsgn, ntc_out, cdact, always_record);
    if (te >= 0) {
        // The state at the trigger, from the step's coefficients.
        bool bad = false;
#pragma unroll 1
        for (int q = 0; q < HYX_N; ++q) {
            const double *cq = tc + (hyx_u64)q * (HYX_P + 1) * ld;
#ifdef HYX_HIGH_ACC
            double r = cq[0], cp = 0.0, hk = 1.0;
            for (int k = 1; k <= HYX_P; ++k) {
                hk *= hend;
                const double y = cq[(hyx_u64)k * ld] * hk - cp;
                const double tt = r + y;
                cp = (tt - r) - y;
                r = tt;
            }
#else
            double r = cq[(hyx_u64)HYX_P * ld];
            for (int k = HYX_P - 1; k >= 0; --k) r = __builtin_fma(r, hend, cq[(hyx_u64)k * ld]);
#endif
            x[q] = r;
            bad = bad || !isfinite(r);
        }
        nf = bad;
    }
    return hend;
}
