// Group form: HYX_G lanes cooperate on one instance, HYX_I = 64 / HYX_G
// instances per wavefront (hand-written CDNA4 kernels, compiled with hipRTC
// after hyx_kcommon.hip).
//
// The generated prelude defines the group plan, the per-lane table struct
// hyx_gtab (cluster input / constant bindings, owned variables, right-hand-side
// terms) with its loader hyx_gtab_load, and
//   hyx_gstep<WTC>(xo, pv, t0, max_dt, h, nf, g, fl, tab, tc, tcs)
// which advances one instance by one adaptive step. Lane g of the group
// evaluates cluster g of the decomposition (one of the isomorphic subgraphs
// the code generator found, e.g. one body pair of an N-body system) with its
// coefficient histories in registers, and owns the state variables
// tab.vo[o] (o < HYX_OWN). Per order, coefficients are exchanged through this
// instance's LDS block `fl` of HYX_FLAT doubles: the cluster outputs of the
// current order, a zero slot, and every order of every state variable
// (rows of HYX_NX = n + 1 columns, one of them zero; variable tab.vo[o] sits in
// column tab.vc[o]), which the state update reads.
//
// Every lane of the wave executes every step, also lanes whose group has no
// instance (their results are not stored): hyx_gstep exchanges data between the
// lanes of the wavefront through LDS at every order, and a block is exactly one
// wavefront.
//
// Built with HYX_GRID defined (the propagate_grid code object), the propagate
// kernels also evaluate the step's Taylor polynomials at the grid times inside
// the step; each lane evaluates its owned state variables from the
// coefficient rows in LDS, which stay intact until the next hyx_gstep.
//
// With events (HYX_EVENTS), hyx_gstep also forms the event functions' jet in every
// lane of the group from the LDS rows and reports which events need a root scan;
// the group then runs the detection of hyx_kevent.hip (every lane the same) and,
// at a terminal trigger, re-evaluates its owned variables from the LDS rows.

struct hyx_ginst {
    double xo[HYX_OWN];
    double pv[HYX_NPAR];
    double thi, tlo, tfhi, tflo, mdt, minh, maxh;
    hyx_u64 iter, nst;
    double hlast; // the last step's size (kept for the host when the instance parks)
    int tdir;
    unsigned gc; // propagate_grid: next grid point
#ifdef HYX_EVENTS
    unsigned long long cdact; // terminal events with a running cooldown
#endif
};

#ifdef HYX_EVENTS
// Event detection for the group's instance i after the step of size h from
// (thi, tlo): shortens h to a terminal trigger and re-evaluates the owned
// variables there from the LDS rows. Every lane of the group computes the same.
__device__ __forceinline__ double hyx_gevents(hyx_ginst &s, const hyx_kargs &a, hyx_u64 i, const double *fl,
                                              const hyx_gtab &t, double thi, double tlo, double h,
                                              unsigned long long evm, double mx, bool &nf, int &te, unsigned &ntc,
                                              bool always_record)
{
    int te_dsgn;
    const double hend = hyx_ev_detect(a, i, thi, tlo, h, evm, mx, te, te_dsgn, ntc, s.cdact, always_record);
    if (te < 0) return h;
    int bad = 0;
#pragma unroll
    for (int o = 0; o < HYX_OWN; ++o) {
        if (t.vo[o] < 0) continue;
        const double *row = fl + HYX_XB + t.vc[o];
#ifdef HYX_HIGH_ACC
        double r = row[0], cp = 0.0, hk = 1.0;
        for (int k = 1; k <= HYX_P; ++k) {
            hk *= hend;
            const double y = row[k * HYX_NX] * hk - cp;
            const double tt = r + y;
            cp = (tt - r) - y;
            r = tt;
        }
#else
        double r = row[HYX_P * HYX_NX];
        for (int k = HYX_P - 1; k >= 0; --k) r = __builtin_fma(r, hend, row[k * HYX_NX]);
#endif
        s.xo[o] = r;
        bad |= isfinite(r) ? 0 : 1;
    }
#pragma unroll
    for (int off = 1; off < HYX_G; off <<= 1) bad |= __shfl_xor(bad, off);
    nf = bad != 0;
    return hend;
}
#endif

__device__ __forceinline__ void hyx_gload(hyx_ginst &s, const hyx_kargs &a, hyx_u64 i, int g, const hyx_gtab &t)
{
#pragma unroll
    for (int o = 0; o < HYX_OWN; ++o)
        s.xo[o] = (t.vo[o] >= 0) ? a.state[(hyx_u64)t.vo[o] * a.ld + i] : 0.0;
#pragma unroll
    for (int q = 0; q < HYX_NPAR; ++q) s.pv[q] = a.pars ? a.pars[(hyx_u64)q * a.ld + i] : 0.0;
    s.thi = a.t_hi[i];
    s.tlo = a.t_lo[i];
}

__device__ __forceinline__ void hyx_gstore_state(const hyx_ginst &s, const hyx_kargs &a, hyx_u64 i, int g,
                                                 const hyx_gtab &t)
{
#pragma unroll
    for (int o = 0; o < HYX_OWN; ++o)
        if (t.vo[o] >= 0) a.state[(hyx_u64)t.vo[o] * a.ld + i] = s.xo[o];
    if (g == 0) {
        a.t_hi[i] = s.thi;
        a.t_lo[i] = s.tlo;
    }
}

__device__ __forceinline__ void hyx_gresume(hyx_ginst &s, const hyx_kargs &a, hyx_u64 i, unsigned st, int g,
                                            const hyx_gtab &t)
{
    hyx_gload(s, a, i, g, t);
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
#ifdef HYX_EVENTS
    s.cdact = hyx_ev_cdmask(a, i);
#endif
}

__device__ __forceinline__ void hyx_gsuspend(const hyx_ginst &s, const hyx_kargs &a, hyx_u64 i, int g,
                                             const hyx_gtab &t, bool done, hyx_i64 res)
{
    hyx_gstore_state(s, a, i, g, t);
    if (g == 0) {
        a.min_h[i] = s.minh;
        a.max_h[i] = s.maxh;
        a.iters[i] = s.iter;
        a.nsteps[i] = s.nst;
#ifdef HYX_GRID
        a.gcur[i] = s.gc;
#endif
        if (done && res == HYX_PARKED) {
            a.status[i] = HYX_ST_PARKED | (s.tdir ? 0u : 2u);
            a.h_out[i] = s.hlast;
        } else if (done) {
            a.status[i] = HYX_ST_DONE;
            a.outcome[i] = res;
        }
    }
}

#ifdef HYX_GRID
// Dense output of the group's instance i at the grid points inside the step of
// size h just taken (time already advanced): every lane of the group walks the
// same points and writes its owned variables. fin: the step reached the final
// time, so every remaining point is evaluated.
__device__ __forceinline__ void hyx_ggrid(hyx_ginst &s, const hyx_kargs &a, hyx_u64 i, const double *fl,
                                          const hyx_gtab &t, double h, bool fin)
{
    unsigned c = s.gc;
    while (c < a.npts) {
        bool inside;
        const double dt = hyx_grid_offset(a.grid[(hyx_u64)c * a.ld + i], s.thi, s.tlo, h, s.tdir, inside);
        if (!inside && !fin) break;
#pragma unroll
        for (int o = 0; o < HYX_OWN; ++o) {
            if (t.vo[o] < 0) continue;
            const double *row = fl + HYX_XB + t.vc[o];
            double r = row[HYX_P * HYX_NX];
            for (int k = HYX_P - 1; k >= 0; --k) r = __builtin_fma(r, dt, row[k * HYX_NX]);
            a.gout[((hyx_u64)c * HYX_N + t.vo[o]) * a.ld + i] = r;
        }
        ++c;
    }
    s.gc = c;
}
#endif

// One propagate_until iteration of the group's instance i; every lane of the
// group computes the same bookkeeping (h and the non-finite flag are reduced
// across the group inside hyx_gstep). wr: the group holds a live instance
// (grid output is written only then).
__device__ __forceinline__ bool hyx_gprop_iter(hyx_ginst &s, const hyx_kargs &a, int g, double *fl,
                                               const hyx_gtab &t, hyx_u64 i, bool wr, hyx_i64 *res)
{
    (void)i;
    (void)wr;
    const double rem = hyx_dfl_sub_hi(s.tfhi, s.tflo, s.thi, s.tlo);
    const double lim = s.tdir ? ((s.mdt < rem) ? s.mdt : rem) : ((-s.mdt > rem) ? -s.mdt : rem);
    double h;
    bool nf;
    int te = -1;
    bool park = false;
#ifdef HYX_EVENTS
    unsigned long long evm = 0;
    double gmx = 0.0;
    hyx_gstep<false>(s.xo, s.pv, s.thi, lim, h, nf, g, fl, t, wr ? a.tc + i : nullptr, a.ld, evm, gmx);
    if (wr) {
        unsigned ntc = 0;
        h = hyx_gevents(s, a, i, fl, t, s.thi, s.tlo, h, evm, gmx, nf, te, ntc, false);
        park = a.ev_pause && ((te >= 0 && !hyx_ev_stop[te]) || ntc > 0u);
    }
#else
    hyx_gstep<false>(s.xo, s.pv, s.thi, lim, h, nf, g, fl, t, nullptr, 0);
#endif
    hyx_dfl_add(s.thi, s.tlo, h, 0.0);
    s.hlast = h;
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
    if (wr) hyx_ggrid(s, a, i, fl, t, h, fin);
#endif
    if (park) {
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
    return false;
}

#define HYX_GROUP_SETUP                                                                                               \
    __shared__ double hyx_lds[HYX_I][HYX_FLAT];                                                                      \
    const int lane = (int)__lane_id();                                                                               \
    const int grp = lane / HYX_G;                                                                                    \
    const int g = lane % HYX_G;                                                                                      \
    double *fl = &hyx_lds[grp][0];                                                                                   \
    for (int q = g; q < HYX_FLAT; q += HYX_G) fl[q] = 0.0;                                                           \
    hyx_gtab tab;                                                                                                    \
    hyx_gtab_load(tab, g);                                                                                           \
    hyx_wsync();

// ---- step(): one adaptive step per instance ----------------------------------------------
template <bool WTC>
__device__ __forceinline__ void hyx_gstep_body(const hyx_kargs &a)
{
    HYX_GROUP_SETUP
    const hyx_u64 i = (hyx_u64)blockIdx.x * HYX_I + grp;
    const bool valid = i < a.N;
    const hyx_u64 ii = valid ? i : 0; // groups past the end step instance 0 and store nothing
    hyx_ginst s;
    hyx_gload(s, a, ii, g, tab);
    const double mdt = a.max_dt[ii];
    double h;
    bool nf;
    int te = -1;
#ifdef HYX_EVENTS
    unsigned long long evm = 0;
    double gmx = 0.0;
    hyx_gstep<WTC>(s.xo, s.pv, s.thi, mdt, h, nf, g, fl, tab, valid ? a.tc + i : nullptr, a.ld, evm, gmx);
    if (valid) {
        unsigned ntc = 0;
        s.cdact = hyx_ev_cdmask(a, i);
        h = hyx_gevents(s, a, i, fl, tab, s.thi, s.tlo, h, evm, gmx, nf, te, ntc, true);
    }
#else
    hyx_gstep<WTC>(s.xo, s.pv, s.thi, mdt, h, nf, g, fl, tab, (WTC && valid) ? a.tc + i : nullptr, a.ld);
#endif
    hyx_dfl_add(s.thi, s.tlo, h, 0.0);
    nf = nf || !isfinite(s.thi);
    if (valid) {
        hyx_gstore_state(s, a, i, g, tab);
        if (g == 0) {
            a.h_out[i] = h;
            a.outcome[i] = nf ? HYX_ERR_NF_STATE
                              : (te >= 0 ? -(hyx_i64)te - 1 : ((h == mdt) ? HYX_TIME_LIMIT : HYX_SUCCESS));
        }
    }
}

extern "C" __global__ void __launch_bounds__(64) hyx_step_k(hyx_kargs a) { hyx_gstep_body<false>(a); }

extern "C" __global__ void __launch_bounds__(64) hyx_step_tc_k(hyx_kargs a) { hyx_gstep_body<true>(a); }

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
    HYX_GROUP_SETUP
    const hyx_u64 i = (hyx_u64)blockIdx.x * HYX_I + grp;
    const unsigned st = (i < a.N) ? a.status[i] : 1u;
    const bool valid = hyx_runnable(st);
    const hyx_u64 ii = (i < a.N) ? i : 0;
    hyx_ginst s;
    hyx_gresume(s, a, ii, st, g, tab);
    hyx_i64 res = HYX_TIME_LIMIT;
    const bool done = hyx_gprop_iter(s, a, g, fl, tab, ii, valid, &res);
    if (valid) hyx_gsuspend(s, a, i, g, tab, done, res);
    const hyx_u64 m = __ballot(valid && !done && g == 0);
    if (lane == 0 && m != 0) atomicAdd(a.active_count, (unsigned)__popcll(m));
}

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

// ---- persistent form: per-group work stealing ----------------------------------------------
// Each wave holds HYX_I instances. Groups whose instance is done claim new
// queue positions together, with one atomic per wave (ballot over the group
// leaders + prefix count). An instance that reaches iter_budget iterations in
// this launch is written back unfinished (counted in *unfinished) and the host
// launches again, so every wave leaves the loop after a bounded number of
// iterations once the queue is drained.
extern "C" __global__ void __launch_bounds__(64) hyx_persist_k(hyx_kargs a)
{
    HYX_GROUP_SETUP
    hyx_ginst s;
    hyx_gload(s, a, 0, g, tab);
    s.tfhi = s.thi;
    s.tflo = s.tlo;
    s.mdt = 0.0;
    s.tdir = 1;
    s.minh = 0.0;
    s.maxh = 0.0;
    s.iter = 0;
    s.nst = 0;
    s.hlast = 0.0;
    s.gc = 0;
#ifdef HYX_EVENTS
    s.cdact = 0ull;
#endif
    hyx_u64 inst = 0, budget = 0;
    bool active = false;
    bool exhausted = false;
    for (;;) {
        const hyx_u64 idle = __ballot(g == 0 && !active);
        const unsigned nidle = (unsigned)__popcll(idle);
        if (!exhausted && nidle != 0 && (nidle * HYX_G >= a.refill_min || nidle == HYX_I)) {
            const int leader = __ffsll((long long)idle) - 1;
            hyx_u64 base = 0;
            if (lane == leader) base = atomicAdd(a.work_counter, (hyx_u64)nidle);
            base = __shfl(base, leader);
            if (base + nidle >= a.N) exhausted = true;
            if (!active) {
                // Rank of this group among the idle groups: idle leaders below its leader lane.
                const int lead_lane = grp * HYX_G;
                const hyx_u64 below_lead = (lead_lane == 0) ? 0ull : ((~0ull) >> (64 - lead_lane));
                const hyx_u64 gpos = base + (hyx_u64)__popcll(idle & below_lead);
                if (gpos < a.N) {
                    inst = a.queue ? (hyx_u64)a.queue[gpos] : gpos;
                    const unsigned st = a.status[inst];
                    if (hyx_runnable(st)) {
                        hyx_gresume(s, a, inst, st, g, tab);
                        budget = 0;
                        active = true;
                    }
                }
            }
        }
        // Leave only once the queue is exhausted: a wave whose claimed positions all hold
        // finished (or unarmed) instances claims again; every claim advances the shared
        // counter, so this ends after at most N claims.
        if (!__any(active) && exhausted) break;
        // Step until an instance of the wave ends, or (while the queue lasts) once when idle
        // groups, claims that found no runnable instance, are enough to claim again: the claim
        // code above, and the registers it needs, stay out of the loop the steps run in, and
        // claims happen after the same steps as with a claim check on every iteration (idle
        // groups only change when an instance ends, which leaves the loop). A wave with no
        // instance claimed none this time: it claims again.
        const unsigned nid = (unsigned)__popcll(__ballot(g == 0 && !active));
        const bool reclaim = !exhausted && nid != 0 && (nid * HYX_G >= a.refill_min || nid == HYX_I);
        if (__any(active)) {
            for (;;) {
                bool ended = false;
                hyx_i64 res = HYX_TIME_LIMIT;
                const bool done = hyx_gprop_iter(s, a, g, fl, tab, inst, active, &res);
                if (active) {
                    budget += 1;
                    if (done || budget >= a.iter_budget) {
                        hyx_gsuspend(s, a, inst, g, tab, done, res);
                        if (!done && g == 0) atomicAdd(a.unfinished, 1u);
                        active = false;
                        ended = true;
                    }
                }
                if (__any(ended) || reclaim) break;
            }
        }
    }
}
