// Lane form: one instance per lane (hand-written CDNA4 kernels, compiled with
// hipRTC after hyx_kcommon.hip). The generated prelude defines
//   HYX_N, HYX_NPAR (>= 1), HYX_P        state size, parameter count, order
//   hyx_taylor_step<WTC>(x, pv, t0, max_dt, h, nf, tc, tcs, tape, tps)
//       one adaptive Taylor step of the instance held by the calling lane:
//       coefficient recurrences, step size from the last two orders, Horner
//       update of x in place; WTC writes the state's Taylor coefficients.
//
// Built with HYX_GRID defined (a second code object, compiled on the first
// propagate_grid of a system), the propagate kernels also write each step's
// Taylor coefficients to a.tc and evaluate them at the grid times that fall
// inside the step (dense output, heyoka's propagate_grid:
// heyoka/expose_batch_integrators.cpp:315-392).
//
// With events (HYX_EVENTS, hyx_kevent.hip), the step reports which events need a
// root scan (and then leaves the coefficients in a.tc); event processing may
// shorten the step to a terminal trigger, and a propagation ends at a terminal
// event with outcome -(index + 1).
//
// Built with HYX_CLOG defined (a third code object, for continuous output), every
// step of a propagation writes its coefficients and its start time and size to
// the next slot of the instance's log (a.clog, a.clog_t); an instance whose log
// is full parks (as for events) until the host has drained it.
#if defined(HYX_GRID) || defined(HYX_CLOG)
#define HYX_PROP_WTC true
#else
#define HYX_PROP_WTC false
#endif
#ifdef HYX_EVENTS
#define HYX_EV_ARGS , evm
#define HYX_TC_ON true
#else
#define HYX_EV_ARGS
#define HYX_TC_ON HYX_PROP_WTC
#endif
//
// The compact prelude may set HYX_SUB > 1: HYX_SUB adjacent lanes then hold the
// same instance (identical registers, identical stores), each summing every
// HYX_SUB-th convolution term inside hyx_taylor_step; claims, counts and the tape
// column are per instance.
#ifndef HYX_SUB
#define HYX_SUB 1
#endif

// ---- one instance's state held in registers ------------------------------------------
struct hyx_inst {
    double x[HYX_N];
    double pv[HYX_NPAR];
    double thi, tlo, tfhi, tflo, mdt, minh, maxh;
    hyx_u64 iter, nst;
    double hlast; // the last step's size (kept for the host when the instance parks)
    int tdir;
    unsigned gc; // propagate_grid: next grid point
    unsigned cn; // continuous output: log slots in use
#ifdef HYX_EVENTS
    unsigned long long cdact; // terminal events with a running cooldown
#endif
};

__device__ __forceinline__ void hyx_load(hyx_inst &s, const hyx_kargs &a, hyx_u64 i)
{
#pragma unroll
    for (int q = 0; q < HYX_N; ++q) s.x[q] = a.state[(hyx_u64)q * a.ld + i];
#pragma unroll
    for (int q = 0; q < HYX_NPAR; ++q) s.pv[q] = a.pars ? a.pars[(hyx_u64)q * a.ld + i] : 0.0;
    s.thi = a.t_hi[i];
    s.tlo = a.t_lo[i];
}

__device__ __forceinline__ void hyx_store_state(const hyx_inst &s, const hyx_kargs &a, hyx_u64 i)
{
#pragma unroll
    for (int q = 0; q < HYX_N; ++q) a.state[(hyx_u64)q * a.ld + i] = s.x[q];
    a.t_hi[i] = s.thi;
    a.t_lo[i] = s.tlo;
}

#ifdef HYX_GRID
// Dense output of instance i at the grid points inside the step of size h just
// taken (the time is already advanced), from the coefficients the step wrote
// to a.tc. fin: the step reached the final time, so every remaining point
// (the last grid time) is evaluated.
__device__ __forceinline__ void hyx_grid_lane(hyx_inst &s, const hyx_kargs &a, hyx_u64 i, double h, bool fin)
{
    unsigned c = s.gc;
    while (c < a.npts) {
        bool inside;
        const double dt = hyx_grid_offset(a.grid[(hyx_u64)c * a.ld + i], s.thi, s.tlo, h, s.tdir, inside);
        if (!inside && !fin) break;
#pragma unroll 1
        for (int q = 0; q < HYX_N; ++q) {
            const double *cq = a.tc + (hyx_u64)q * (HYX_P + 1) * a.ld + i;
            double r = cq[(hyx_u64)HYX_P * a.ld];
            for (int k = HYX_P - 1; k >= 0; --k) r = __builtin_fma(r, dt, cq[(hyx_u64)k * a.ld]);
            a.gout[((hyx_u64)c * HYX_N + q) * a.ld + i] = r;
        }
        ++c;
    }
    s.gc = c;
}
#endif

// One iteration of propagate_until for one instance i (lane_tape: its column
// of the compact-form tape). Returns true when the instance is done; *res then
// holds its outcome.
__device__ __forceinline__ bool hyx_prop_iter(hyx_inst &s, const hyx_kargs &a, hyx_u64 lane_tape, hyx_u64 i,
                                              hyx_i64 *res)
{
    (void)i;
    // Clamp: the remaining time, or max_delta_t if smaller.
    const double rem = hyx_dfl_sub_hi(s.tfhi, s.tflo, s.thi, s.tlo);
    const double lim = s.tdir ? ((s.mdt < rem) ? s.mdt : rem) : ((-s.mdt > rem) ? -s.mdt : rem);
    double h;
    bool nf;
    unsigned long long evm = 0;
    (void)evm;
#ifdef HYX_CLOG
    const unsigned slot = s.cn; // < a.clog_cap: an instance with a full log parks
    double *tcp = a.clog + (hyx_u64)slot * (hyx_u64)(HYX_N * (HYX_P + 1)) * a.ld + i;
    const double th0 = s.thi, tl0 = s.tlo;
#else
    double *tcp = HYX_TC_ON ? a.tc + i : nullptr;
#endif
    hyx_taylor_step<HYX_PROP_WTC>(s.x, s.pv, s.thi, lim, h, nf, tcp, a.ld, a.tape ? a.tape + lane_tape : nullptr,
                                  a.tape_stride HYX_EV_ARGS);
    int te = -1;
    bool park = false;
#ifdef HYX_EVENTS
    int te_dsgn;
    unsigned ntc;
    h = hyx_ev_step(a, i, s.pv, s.thi, s.tlo, h, s.x, nf, te, te_dsgn, ntc, evm, s.cdact, false);
    park = a.ev_pause && ((te >= 0 && !hyx_ev_stop[te]) || ntc > 0u);
#endif
    hyx_dfl_add(s.thi, s.tlo, h, 0.0);
    s.hlast = h;
#ifdef HYX_CLOG
    a.clog_t[((hyx_u64)slot * 3) * a.ld + i] = th0;
    a.clog_t[((hyx_u64)slot * 3 + 1) * a.ld + i] = tl0;
    a.clog_t[((hyx_u64)slot * 3 + 2) * a.ld + i] = h;
    s.cn = slot + 1;
#endif
    nf = nf || !isfinite(s.thi);
    s.iter += 1;
    s.nst += (h != 0.0) ? 1 : 0;
    const hyx_i64 oc = nf ? HYX_ERR_NF_STATE : ((te < 0 && h == lim) ? HYX_TIME_LIMIT : HYX_SUCCESS);
    // min_h / max_h over steps the rule chose alone: not those clamped by the time limit
    // or shortened by a terminal event.
    if (oc == HYX_SUCCESS && te < 0) {
        const double ah = fabs(h);
        s.minh = (ah < s.minh) ? ah : s.minh;
        s.maxh = (ah > s.maxh) ? ah : s.maxh;
    }
    if (oc == HYX_ERR_NF_STATE) {
        *res = HYX_ERR_NF_STATE;
        return true;
    }
    const bool fin = oc == HYX_TIME_LIMIT && h == rem;
#ifdef HYX_GRID
    hyx_grid_lane(s, a, i, h, fin);
#endif
    if (park) {
        // The host runs this step's callbacks, then resumes (or stops) the instance.
        *res = HYX_PARKED;
        return true;
    }
    if (te >= 0) {
        *res = -(hyx_i64)te - 1;
        return true;
    }
    if (fin) {
        *res = HYX_TIME_LIMIT;
        return true;
    }
    if (a.max_steps != 0 && s.iter >= a.max_steps) {
        *res = HYX_STEP_LIMIT;
        return true;
    }
#ifdef HYX_CLOG
    if (s.cn >= a.clog_cap) {
        *res = HYX_PARKED; // the host drains the log, then resumes the instance
        return true;
    }
#endif
    return false;
}

// Resume an instance of a propagation set up by hyx_iter_init_k.
__device__ __forceinline__ void hyx_resume(hyx_inst &s, const hyx_kargs &a, hyx_u64 i, unsigned st)
{
    hyx_load(s, a, i);
    s.tfhi = a.tf_hi[i];
    s.tflo = a.tf_lo[i];
    s.mdt = a.max_dt[i];
    s.tdir = (st == 0u);
    s.minh = a.min_h[i];
    s.maxh = a.max_h[i];
    s.iter = a.iters[i];
    s.nst = a.nsteps[i];
#ifdef HYX_GRID
    s.gc = a.gcur[i];
#endif
#ifdef HYX_CLOG
    s.cn = a.clog_n[i];
#endif
#ifdef HYX_EVENTS
    s.cdact = hyx_ev_cdmask(a, i);
#endif
}

// Write an instance back; done instances also record their outcome.
__device__ __forceinline__ void hyx_suspend(const hyx_inst &s, const hyx_kargs &a, hyx_u64 i, bool done, hyx_i64 res)
{
    hyx_store_state(s, a, i);
    a.min_h[i] = s.minh;
    a.max_h[i] = s.maxh;
    a.iters[i] = s.iter;
    a.nsteps[i] = s.nst;
#ifdef HYX_GRID
    a.gcur[i] = s.gc;
#endif
#ifdef HYX_CLOG
    a.clog_n[i] = s.cn;
#endif
    if (done && res == HYX_PARKED) {
        a.status[i] = HYX_ST_PARKED | (s.tdir ? 0u : 2u);
        a.h_out[i] = s.hlast;
    } else if (done) {
        a.status[i] = HYX_ST_DONE;
        a.outcome[i] = res;
    }
}

// ---- step(): one adaptive step per instance ----------------------------------------------
template <bool WTC>
__device__ __forceinline__ void hyx_step_body(const hyx_kargs &a)
{
    const hyx_u64 i = ((hyx_u64)blockIdx.x * blockDim.x + threadIdx.x) / HYX_SUB;
    if (i >= a.N) return;
    hyx_inst s;
    hyx_load(s, a, i);
    const double mdt = a.max_dt[i];
    double h;
    bool nf;
#ifdef HYX_EVENTS
    unsigned long long evm = 0;
    hyx_taylor_step<WTC>(s.x, s.pv, s.thi, mdt, h, nf, a.tc + i, a.ld, a.tape ? a.tape + i : nullptr, a.tape_stride,
                         evm);
    int te, te_dsgn;
    unsigned ntc;
    s.cdact = hyx_ev_cdmask(a, i);
    h = hyx_ev_step(a, i, s.pv, s.thi, s.tlo, h, s.x, nf, te, te_dsgn, ntc, evm, s.cdact, true);
#else
    hyx_taylor_step<WTC>(s.x, s.pv, s.thi, mdt, h, nf, WTC ? a.tc + i : nullptr, a.ld, a.tape ? a.tape + i : nullptr,
                         a.tape_stride);
    const int te = -1;
#endif
    hyx_dfl_add(s.thi, s.tlo, h, 0.0);
    nf = nf || !isfinite(s.thi);
    hyx_store_state(s, a, i);
    a.h_out[i] = h;
    // A terminal event reports -(index + 1); the host turns it into the index when
    // the event's callback lets the integration continue.
    a.outcome[i] = nf ? HYX_ERR_NF_STATE
                      : (te >= 0 ? -(hyx_i64)te - 1 : ((h == mdt) ? HYX_TIME_LIMIT : HYX_SUCCESS));
}

extern "C" __global__ void __launch_bounds__(64) hyx_step_k(hyx_kargs a) { hyx_step_body<false>(a); }

extern "C" __global__ void __launch_bounds__(64) hyx_step_tc_k(hyx_kargs a) { hyx_step_body<true>(a); }

// ---- plain form: one propagate iteration per launch ----------------------------------------
extern "C" __global__ void __launch_bounds__(64) hyx_iter_init_k(hyx_kargs a)
{
    const hyx_u64 i = (hyx_u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.N) return;
    const double rem = hyx_dfl_sub_hi(a.tf_hi[i], a.tf_lo[i], a.t_hi[i], a.t_lo[i]);
    a.status[i] = (rem >= 0.0) ? 0u : 2u;
    a.iters[i] = 0;
    a.nsteps[i] = 0;
    a.min_h[i] = __builtin_inf();
    a.max_h[i] = 0.0;
    a.outcome[i] = HYX_TIME_LIMIT;
}

extern "C" __global__ void __launch_bounds__(64) hyx_iter_k(hyx_kargs a)
{
    const hyx_u64 i = ((hyx_u64)blockIdx.x * blockDim.x + threadIdx.x) / HYX_SUB;
    const bool lead = (threadIdx.x % HYX_SUB) == 0;
    bool still = false;
    if (i < a.N) {
        const unsigned st = a.status[i];
        if (hyx_runnable(st)) {
            hyx_inst s;
            hyx_resume(s, a, i, st);
            hyx_i64 res = HYX_TIME_LIMIT;
            const bool done = hyx_prop_iter(s, a, i, i, &res);
            hyx_suspend(s, a, i, done, res);
            still = !done;
        }
    }
    const hyx_u64 m = __ballot(still && lead);
    if (threadIdx.x == 0 && m != 0) atomicAdd(a.active_count, (unsigned)__popcll(m));
}

// Clears the plain form's per-launch active counters (a kernel rather than a
// memset so that it can sit inside a captured graph).
extern "C" __global__ void hyx_clear_u32_k(unsigned *p, unsigned n)
{
    const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

extern "C" __global__ void hyx_clear_u64_k(hyx_u64 *p, unsigned n)
{
    const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

// ---- persistent form: per-wavefront work stealing -----------------------------------------
// Each wave keeps 64 instances in registers. Lanes whose instance is done
// claim new queue positions with one atomic per wave (ballot + prefix count),
// so a ragged ensemble keeps every lane busy until the queue drains. An
// instance that reaches iter_budget iterations in this launch is written back
// unfinished (counted in *unfinished, listed in tailq when given) and the host
// launches again; so every lane leaves the loop after a bounded number of iterations.
//
// Tail hand-off (a.inflight given, tail_min > 0: the system has a tail form): the
// kernel counts the instances in flight; once the queue is drained and at most
// tail_min remain, every wave writes its instances back (listed in tailq) and leaves,
// and the host continues them in the tail form (this same kernel compiled with the
// quad prelude, four lanes per instance, or the twin prelude, two), with fewer
// instructions per step per lane and bit for bit the same arithmetic. The last
// instances of a ragged ensemble, which run alone on their SIMDs, then take steps
// of lower latency.
extern "C" __global__ void __launch_bounds__(64) hyx_persist_k(hyx_kargs a)
{
    const int lane = (int)__lane_id();
    const int sub = lane % HYX_SUB;
    const int lead = lane - sub; // the lane that claims for this instance slot
    const hyx_u64 gslot = ((hyx_u64)blockIdx.x * 64 + lane) / HYX_SUB; // tape column of the slot
    hyx_inst s;
    hyx_u64 inst = 0, budget = 0;
    bool active = false;
    bool exhausted = false;
    for (;;) {
        bool claimed = false;
        const hyx_u64 idle = __ballot(!active && sub == 0);
        const unsigned nidle = (unsigned)__popcll(idle);
        if (!exhausted && nidle != 0 && (nidle * HYX_SUB >= a.refill_min || nidle == 64u / HYX_SUB)) {
            const int leader = __ffsll((long long)idle) - 1;
            hyx_u64 base = 0;
            if (lane == leader) base = atomicAdd(a.work_counter, (hyx_u64)nidle);
            base = __shfl(base, leader);
            if (base + nidle >= a.N) exhausted = true;
            if (!active) {
                // Rank of this slot among the idle slots: idle leaders below its leader lane.
                const hyx_u64 below = (lead == 0) ? 0ull : ((~0ull) >> (64 - lead));
                const hyx_u64 pos = base + (hyx_u64)__popcll(idle & below);
                if (pos < a.N) {
                    inst = a.queue ? (hyx_u64)a.queue[pos] : pos;
                    const unsigned st = a.status[inst];
                    if (hyx_runnable(st)) {
                        hyx_resume(s, a, inst, st);
                        budget = 0;
                        active = true;
                        claimed = true;
                    }
                }
            }
        }
        if (a.inflight) {
            const hyx_u64 got = __ballot(claimed && sub == 0);
            if (lane == 0 && got != 0) atomicAdd(a.inflight, (unsigned)__popcll(got));
        }
        // Queue drained and no hand-off to watch: nothing is claimed or counted any more, so
        // the remaining instances step in a loop of their own, free of the claim code and of
        // what it keeps live (the straggler of a ragged ensemble runs this loop alone).
        if (exhausted && !a.inflight) {
            while (active) {
                hyx_i64 res = HYX_TIME_LIMIT;
                const bool done = hyx_prop_iter(s, a, gslot, inst, &res);
                budget += 1;
                if (done || budget >= a.iter_budget) {
                    hyx_suspend(s, a, inst, done, res);
                    if (!done && sub == 0) {
                        const unsigned q = atomicAdd(a.unfinished, 1u);
                        if (a.tailq) a.tailq[q] = (unsigned)inst;
                    }
                    active = false;
                }
            }
            break;
        }
        // Leave only once the queue is exhausted: a wave whose claimed positions all hold
        // finished (or unarmed) instances claims again; every claim advances the shared
        // counter, so this ends after at most N claims.
        if (!__any(active) && exhausted) break;
        // Step until an instance of the wave ends, or (while the queue lasts) once when idle
        // slots, claims that found no runnable instance, are enough to claim again, or (once
        // it is drained) until the next hand-off check, every 16 iterations: the claim and
        // count code above, and the registers it needs, stay out of the loop that the steps
        // run in, and claims happen after the same steps as with a claim check on every
        // iteration (idle slots only change when an instance ends, which leaves the loop).
        // A wave with no instance claimed none this time: it claims again.
        bool ended = false;
        const unsigned nid = (unsigned)__popcll(__ballot(!active && sub == 0));
        const bool reclaim = !exhausted && nid != 0 && (nid * HYX_SUB >= a.refill_min || nid == 64u / HYX_SUB);
        if (__any(active)) {
            for (unsigned it = 1;; ++it) {
                if (active) {
                    hyx_i64 res = HYX_TIME_LIMIT;
                    const bool done = hyx_prop_iter(s, a, gslot, inst, &res);
                    budget += 1;
                    if (done || budget >= a.iter_budget) {
                        hyx_suspend(s, a, inst, done, res);
                        if (!done && sub == 0) {
                            const unsigned q = atomicAdd(a.unfinished, 1u);
                            if (a.tailq) a.tailq[q] = (unsigned)inst;
                        }
                        active = false;
                        ended = true;
                    }
                }
                if (__any(ended) || reclaim) break;
                if (exhausted && (it & 15u) == 0u) break;
            }
        }
        if (a.inflight) {
            const hyx_u64 fin = __ballot(ended && sub == 0);
            if (lane == 0 && fin != 0) atomicSub(a.inflight, (unsigned)__popcll(fin));
            if (a.tail_min != 0u && exhausted) {
                unsigned infl = 0;
                if (lane == 0) infl = __hip_atomic_load(a.inflight, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                infl = (unsigned)__shfl((int)infl, 0);
                if (infl <= a.tail_min && active) {
                    // Hand-off: written back unfinished, continued by the tail form.
                    hyx_suspend(s, a, inst, false, HYX_TIME_LIMIT);
                    if (sub == 0) {
                        const unsigned q = atomicAdd(a.unfinished, 1u);
                        a.tailq[q] = (unsigned)inst;
                    }
                    active = false;
                }
            }
        }
    }
}
