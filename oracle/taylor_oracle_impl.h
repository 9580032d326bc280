/* CPU restatement of heyoka's adaptive Taylor step and propagate loop.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the parity oracle for the GPU kernels
 * and the CPU baseline that bench.py times. Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it, and only as the checker.
 *
 * This header is included twice by taylor_oracle.c, once with REAL = double and
 * once with REAL = __float128 (the real128 oracle). It interprets the op-list of
 * include/hyx_oplist.h.
 *
 * What it restates (the arithmetic lives in heyoka C++ 7.11.0, which heyoka.py
 * binds but does not vendor; CMakeLists.txt:127):
 *   - normalised-derivative recurrences per elementary op
 *     (doc/tut_taylor_method.rst:22-81; Jorba & Zou 2005, Exp. Math. 14:99);
 *   - order rule p = ceil(-ln tol / 2 + 1) (applied by the caller);
 *   - step size rho_m * e^-2 * e^(-0.7/(p-1)) from the infinity norms of the
 *     last two orders, absolute/relative switch on max|x| vs 1
 *     (doc/notebooks/Customising the adaptive integrator.ipynb:23-31);
 *   - Horner state update, or compensated summation in high-accuracy mode
 *     (Customising the adaptive integrator.ipynb:255-277);
 *   - double-length time (heyoka/expose_batch_integrators.cpp:407-449);
 *   - the batch propagate_until loop: finished lanes take zero-length steps,
 *     per-lane (outcome, min_h, max_h, nsteps) (expose_batch_integrators.cpp:393,
 *     doc/notebooks/Batch mode overview.ipynb:340-386).
 * Pinned by: tests/test_oracle_goldens.py (notebook outputs) and
 * tests/test_oracle_series.py (closed-form Taylor series, invariants).
 */

#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)
#define FN(name) CAT(CAT(ora_, SFX), CAT(_, name))

typedef struct {
    REAL hi, lo;
} FN(dfl);

/* Knuth's error-free sum (TAOCP vol. 2, 4.2.2): a + b = s + e exactly. */
static inline void FN(two_sum)(REAL a, REAL b, REAL *s, REAL *e)
{
    REAL x = a + b;
    REAL z = x - a;
    *e = (a - (x - z)) + (b - z);
    *s = x;
}

/* Dekker's fast error-free sum, valid for |a| >= |b|. */
static inline void FN(fast_two_sum)(REAL a, REAL b, REAL *s, REAL *e)
{
    REAL x = a + b;
    *e = b - (x - a);
    *s = x;
}

/* Double-length addition (Dekker 1971 / QD "ieee_add"): both parts summed with
   two_sum, then renormalised. */
static inline FN(dfl) FN(dfl_add)(FN(dfl) x, FN(dfl) y)
{
    REAL s, e, t, f;
    FN(two_sum)(x.hi, y.hi, &s, &e);
    FN(two_sum)(x.lo, y.lo, &t, &f);
    e += t;
    FN(fast_two_sum)(s, e, &s, &e);
    e += f;
    FN(fast_two_sum)(s, e, &s, &e);
    FN(dfl) r = {s, e};
    return r;
}

static inline FN(dfl) FN(dfl_neg)(FN(dfl) x)
{
    FN(dfl) r = {-x.hi, -x.lo};
    return r;
}

static inline FN(dfl) FN(dfl_norm)(REAL hi, REAL lo)
{
    FN(dfl) r;
    FN(fast_two_sum)(hi, lo, &r.hi, &r.lo);
    return r;
}

/* Parsed op-list. */
typedef struct {
    uint32_t n_state, n_par, n_nodes, n_consts;
    const uint32_t *words;
    const uint32_t **rec; /* rec[j] -> {op, nargs, aux, args...} for j >= n_state */
    const uint32_t *rhs;
    const double *consts;
} FN(sys);

static int FN(parse)(FN(sys) * s, const uint32_t *w, const double *consts)
{
    if (w[0] != HYX_OPLIST_MAGIC) return -1;
    s->n_state = w[1];
    s->n_par = w[2];
    s->n_nodes = w[3];
    s->n_consts = w[4];
    s->words = w;
    s->consts = consts;
    s->rec = (const uint32_t **)calloc(s->n_nodes ? s->n_nodes : 1, sizeof(uint32_t *));
    uint32_t off = 6;
    for (uint32_t j = s->n_state; j < s->n_nodes; ++j) {
        s->rec[j] = w + off;
        off += 3 + w[off + 1];
    }
    if (off != w[5]) {
        free(s->rec);
        return -2;
    }
    s->rhs = w + off;
    return 0;
}

/* Working set for a chunk of L lanes: tape[node][order][lane]. */
typedef struct {
    uint32_t L, P;
    REAL *tape;
    REAL *pars; /* [n_par][L] */
} FN(work);

#define TP(wk, node, k, l) ((wk)->tape[((size_t)(node) * ((wk)->P + 1) + (k)) * (wk)->L + (l)])

/* Value of an operand at order k for lane l. */
static inline REAL FN(opnd)(const FN(sys) * s, const FN(work) * wk, uint32_t w, uint32_t k, uint32_t l)
{
    uint32_t kind = HYX_OPERAND_KIND(w), idx = HYX_OPERAND_INDEX(w);
    if (kind == HYX_KIND_NODE) return TP(wk, idx, k, l);
    if (k != 0) return (REAL)0;
    if (kind == HYX_KIND_CONST) return (REAL)s->consts[idx];
    return wk->pars[(size_t)idx * wk->L + l];
}

/* Order-0 value of node j. */
static void FN(eval0)(const FN(sys) * s, FN(work) * wk, uint32_t j, REAL t0, uint32_t l)
{
    const uint32_t *r = s->rec[j];
    uint32_t op = r[0], na = r[1];
    const uint32_t *a = r + 3;
    REAL v = 0;
    switch (op) {
    case HYX_OP_ADD: v = FN(opnd)(s, wk, a[0], 0, l) + FN(opnd)(s, wk, a[1], 0, l); break;
    case HYX_OP_SUB: v = FN(opnd)(s, wk, a[0], 0, l) - FN(opnd)(s, wk, a[1], 0, l); break;
    case HYX_OP_NEG: v = -FN(opnd)(s, wk, a[0], 0, l); break;
    case HYX_OP_SUM:
        v = FN(opnd)(s, wk, a[0], 0, l);
        for (uint32_t q = 1; q < na; ++q) v += FN(opnd)(s, wk, a[q], 0, l);
        break;
    case HYX_OP_MUL: v = FN(opnd)(s, wk, a[0], 0, l) * FN(opnd)(s, wk, a[1], 0, l); break;
    case HYX_OP_DIV: v = FN(opnd)(s, wk, a[0], 0, l) / FN(opnd)(s, wk, a[1], 0, l); break;
    case HYX_OP_SQUARE: {
        REAL x = FN(opnd)(s, wk, a[0], 0, l);
        v = x * x;
        break;
    }
    case HYX_OP_SQRT: v = R_SQRT(FN(opnd)(s, wk, a[0], 0, l)); break;
    case HYX_OP_POW: v = R_POW(FN(opnd)(s, wk, a[0], 0, l), (REAL)s->consts[r[2]]); break;
    case HYX_OP_EXP: v = R_EXP(FN(opnd)(s, wk, a[0], 0, l)); break;
    case HYX_OP_LOG: v = R_LOG(FN(opnd)(s, wk, a[0], 0, l)); break;
    case HYX_OP_SIN: v = R_SIN(FN(opnd)(s, wk, a[0], 0, l)); break;
    case HYX_OP_COS: v = R_COS(FN(opnd)(s, wk, a[0], 0, l)); break;
    case HYX_OP_TIME: v = t0; break;
    default: v = (REAL)NAN;
    }
    TP(wk, j, 0, l) = v;
}

/* Order-k (k >= 1) normalised derivative of node j. */
static void FN(evalk)(const FN(sys) * s, FN(work) * wk, uint32_t j, uint32_t k, uint32_t l)
{
    const uint32_t *r = s->rec[j];
    uint32_t op = r[0], na = r[1];
    const uint32_t *a = r + 3;
    REAL v = 0;
    switch (op) {
    case HYX_OP_ADD: v = FN(opnd)(s, wk, a[0], k, l) + FN(opnd)(s, wk, a[1], k, l); break;
    case HYX_OP_SUB: v = FN(opnd)(s, wk, a[0], k, l) - FN(opnd)(s, wk, a[1], k, l); break;
    case HYX_OP_NEG: v = -FN(opnd)(s, wk, a[0], k, l); break;
    case HYX_OP_SUM:
        v = FN(opnd)(s, wk, a[0], k, l);
        for (uint32_t q = 1; q < na; ++q) v += FN(opnd)(s, wk, a[q], k, l);
        break;
    case HYX_OP_MUL: {
        /* c^[k] = sum_{i=0}^{k} a^[i] b^[k-i] */
        REAL acc = 0;
        for (uint32_t i = 0; i <= k; ++i) acc += FN(opnd)(s, wk, a[0], i, l) * FN(opnd)(s, wk, a[1], k - i, l);
        v = acc;
        break;
    }
    case HYX_OP_DIV: {
        /* c^[k] = (a^[k] - sum_{i=1}^{k} b^[i] c^[k-i]) / b^[0] */
        REAL acc = 0;
        for (uint32_t i = 1; i <= k; ++i) acc += FN(opnd)(s, wk, a[1], i, l) * TP(wk, j, k - i, l);
        v = (FN(opnd)(s, wk, a[0], k, l) - acc) / FN(opnd)(s, wk, a[1], 0, l);
        break;
    }
    case HYX_OP_SQUARE: {
        /* c^[k] = sum_{i=0}^{k} a^[i] a^[k-i], folded symmetrically */
        REAL acc = 0;
        for (uint32_t i = 0; 2 * i < k; ++i) acc += FN(opnd)(s, wk, a[0], i, l) * FN(opnd)(s, wk, a[0], k - i, l);
        acc = acc + acc;
        if (k % 2 == 0) {
            REAL m = FN(opnd)(s, wk, a[0], k / 2, l);
            acc += m * m;
        }
        v = acc;
        break;
    }
    case HYX_OP_SQRT: {
        /* c^[k] = (a^[k] - sum_{i=1}^{k-1} c^[i] c^[k-i]) / (2 c^[0]) */
        REAL acc = 0;
        for (uint32_t i = 1; 2 * i < k; ++i) acc += TP(wk, j, i, l) * TP(wk, j, k - i, l);
        acc = acc + acc;
        if (k % 2 == 0) {
            REAL m = TP(wk, j, k / 2, l);
            acc += m * m;
        }
        v = (FN(opnd)(s, wk, a[0], k, l) - acc) / (2 * TP(wk, j, 0, l));
        break;
    }
    case HYX_OP_POW: {
        /* c^[k] = 1/(k a^[0]) sum_{i=0}^{k-1} (k alpha - i (alpha + 1)) a^[k-i] c^[i] */
        REAL al = (REAL)s->consts[r[2]];
        REAL acc = 0;
        for (uint32_t i = 0; i < k; ++i)
            acc += ((REAL)k * al - (REAL)i * (al + 1)) * FN(opnd)(s, wk, a[0], k - i, l) * TP(wk, j, i, l);
        v = acc / ((REAL)k * FN(opnd)(s, wk, a[0], 0, l));
        break;
    }
    case HYX_OP_EXP: {
        /* c^[k] = (1/k) sum_{i=1}^{k} i a^[i] c^[k-i] */
        REAL acc = 0;
        for (uint32_t i = 1; i <= k; ++i) acc += (REAL)i * FN(opnd)(s, wk, a[0], i, l) * TP(wk, j, k - i, l);
        v = acc / (REAL)k;
        break;
    }
    case HYX_OP_LOG: {
        /* c^[k] = (a^[k] - (1/k) sum_{i=1}^{k-1} i c^[i] a^[k-i]) / a^[0] */
        REAL acc = 0;
        for (uint32_t i = 1; i < k; ++i) acc += (REAL)i * TP(wk, j, i, l) * FN(opnd)(s, wk, a[0], k - i, l);
        v = (FN(opnd)(s, wk, a[0], k, l) - acc / (REAL)k) / FN(opnd)(s, wk, a[0], 0, l);
        break;
    }
    case HYX_OP_SIN:
    case HYX_OP_COS: {
        /* s^[k] = (1/k) sum i a^[i] c^[k-i];  c^[k] = -(1/k) sum i a^[i] s^[k-i] */
        uint32_t partner = r[2];
        REAL acc = 0;
        for (uint32_t i = 1; i <= k; ++i) acc += (REAL)i * FN(opnd)(s, wk, a[0], i, l) * TP(wk, partner, k - i, l);
        v = acc / (REAL)k;
        if (op == HYX_OP_COS) v = -v;
        break;
    }
    case HYX_OP_TIME: v = (k == 1) ? (REAL)1 : (REAL)0; break;
    default: v = (REAL)NAN;
    }
    TP(wk, j, k, l) = v;
}

/* Fill the Taylor tape for lanes [0, L): orders 0..p of the state variables and
   orders 0..p-1 of every node. x holds the order-0 state [n][L]. */
static void FN(taylor_coeffs)(const FN(sys) * s, FN(work) * wk, const REAL *x, const REAL *t0)
{
    const uint32_t n = s->n_state, P = wk->P, L = wk->L;
    for (uint32_t i = 0; i < n; ++i)
        for (uint32_t l = 0; l < L; ++l) TP(wk, i, 0, l) = x[(size_t)i * L + l];
    for (uint32_t j = n; j < s->n_nodes; ++j)
        for (uint32_t l = 0; l < L; ++l) FN(eval0)(s, wk, j, t0[l], l);
    for (uint32_t k = 1; k <= P; ++k) {
        for (uint32_t i = 0; i < n; ++i)
            for (uint32_t l = 0; l < L; ++l) TP(wk, i, k, l) = FN(opnd)(s, wk, s->rhs[i], k - 1, l) / (REAL)k;
        if (k < P)
            for (uint32_t j = n; j < s->n_nodes; ++j)
                for (uint32_t l = 0; l < L; ++l) FN(evalk)(s, wk, j, k, l);
    }
}

static inline int FN(finite)(REAL x) { return R_ISFINITE(x); }

/* One adaptive step for lanes [0, L). max_dt[l] is the signed clamp (+-inf for
   none). Updates x and the time in place, writes h and the outcome. */
static void FN(step)(const FN(sys) * s, FN(work) * wk, REAL rhofac, int high_acc, REAL *x, FN(dfl) * t,
                     const REAL *max_dt, REAL *h_out, int64_t *oc_out)
{
    const uint32_t n = s->n_state, P = wk->P, L = wk->L;
    REAL t0[64];
    for (uint32_t l = 0; l < L; ++l) t0[l] = t[l].hi;
#ifdef ORA_GENERATED_COEFFS
    /* Compiled-port build (cport.py): the system's own generated recurrences. */
    ORA_GENERATED_COEFFS(wk->tape, wk->pars, x, t0, wk->L);
#else
    FN(taylor_coeffs)(s, wk, x, t0);
#endif
    for (uint32_t l = 0; l < L; ++l) {
        REAL mx = 0, mo = 0, mom1 = 0;
        for (uint32_t i = 0; i < n; ++i) {
            REAL a = R_FABS(TP(wk, i, 0, l)), b = R_FABS(TP(wk, i, P, l)), c = R_FABS(TP(wk, i, P - 1, l));
            mx = (a > mx) ? a : mx;
            mo = (b > mo) ? b : mo;
            mom1 = (c > mom1) ? c : mom1;
        }
        REAL num = (mx <= 1) ? (REAL)1 : mx;
        REAL rho_o = R_POW(num / mo, (REAL)1 / (REAL)P);
        REAL rho_om1 = R_POW(num / mom1, (REAL)1 / (REAL)(P - 1));
        REAL rho_m = (rho_om1 < rho_o) ? rho_om1 : rho_o;
        REAL h = rho_m * rhofac;
        REAL lim = R_FABS(max_dt[l]);
        h = (h < lim) ? h : lim;
        if (max_dt[l] < 0) h = -h;
        int fin = 1;
        for (uint32_t i = 0; i < n; ++i) {
            REAL res;
            if (!high_acc) {
                res = TP(wk, i, P, l);
                for (int k = (int)P - 1; k >= 0; --k) res = res * h + TP(wk, i, k, l);
            } else {
                /* Kahan-compensated sum of x^[k] h^k, k = 0..p. */
                REAL sum = TP(wk, i, 0, l), comp = 0, hk = 1;
                for (uint32_t k = 1; k <= P; ++k) {
                    hk *= h;
                    REAL y = TP(wk, i, k, l) * hk - comp;
                    REAL tt = sum + y;
                    comp = (tt - sum) - y;
                    sum = tt;
                }
                res = sum;
            }
            x[(size_t)i * L + l] = res;
            fin = fin && FN(finite)(res);
        }
        FN(dfl) hd = {h, 0};
        t[l] = FN(dfl_add)(t[l], hd);
        fin = fin && FN(finite)(t[l].hi);
        h_out[l] = h;
        oc_out[l] = !fin ? ORA_ERR_NF_STATE : (h == max_dt[l] ? ORA_TIME_LIMIT : ORA_SUCCESS);
    }
}

static void FN(work_init)(FN(work) * wk, const FN(sys) * s, uint32_t P, uint32_t L)
{
    wk->L = L;
    wk->P = P;
    wk->tape = (REAL *)calloc((size_t)s->n_nodes * (P + 1) * L, sizeof(REAL));
    wk->pars = (REAL *)calloc((size_t)(s->n_par ? s->n_par : 1) * L, sizeof(REAL));
}

static void FN(work_free)(FN(work) * wk)
{
    free(wk->tape);
    free(wk->pars);
}

static REAL FN(rhofac)(uint32_t P)
{
    /* e^-2 * e^(-0.7/(p-1)) */
    REAL e = R_EXP((REAL)1);
    return R_EXP((REAL)(-7) / (REAL)10 / (REAL)(P - 1)) / (e * e);
}

/* ---------------------------------------------------------------------------------------
 * Exported entry points. State and parameters are [rows][B] doubles, batch index
 * fastest (heyoka/expose_batch_integrators.cpp:101-119). For the __float128 build the
 * inputs are converted exactly and the outputs rounded; state_lo, when non-NULL,
 * receives (double)(x - (double)x).
 * ------------------------------------------------------------------------------------- */

/* One step for B lanes. tc_out (optional): [n][p+1][B] Taylor coefficients. */
int FN(step_batch)(const uint32_t *words, const double *consts, uint32_t order, int high_acc, uint32_t B,
                   double *state, double *state_lo, const double *pars, double *t_hi, double *t_lo,
                   const double *max_dt, double *h_out, int64_t *oc_out, double *tc_out)
{
    FN(sys) s;
    if (FN(parse)(&s, words, consts) != 0) return -1;
    if (order < 2 || B == 0) {
        free(s.rec);
        return -2;
    }
    const uint32_t n = s.n_state, L = 8;
    const REAL rf = FN(rhofac)(order);
    FN(work) wk;
    FN(work_init)(&wk, &s, order, L);
    REAL *x = (REAL *)malloc(sizeof(REAL) * (size_t)n * L);
    for (uint32_t b0 = 0; b0 < B; b0 += L) {
        uint32_t nl = (B - b0 < L) ? B - b0 : L;
        FN(dfl) t[8];
        REAL mdt[8], h[8];
        int64_t oc[8];
        for (uint32_t l = 0; l < L; ++l) {
            uint32_t b = b0 + (l < nl ? l : 0);
            for (uint32_t i = 0; i < n; ++i) x[(size_t)i * L + l] = (REAL)state[(size_t)i * B + b];
            for (uint32_t q = 0; q < s.n_par; ++q) wk.pars[(size_t)q * L + l] = (REAL)pars[(size_t)q * B + b];
            t[l] = FN(dfl_norm)((REAL)t_hi[b], (REAL)t_lo[b]);
            mdt[l] = (REAL)max_dt[b];
        }
        FN(step)(&s, &wk, rf, high_acc, x, t, mdt, h, oc);
        for (uint32_t l = 0; l < nl; ++l) {
            uint32_t b = b0 + l;
            for (uint32_t i = 0; i < n; ++i) {
                REAL v = x[(size_t)i * L + l];
                state[(size_t)i * B + b] = (double)v;
                if (state_lo) state_lo[(size_t)i * B + b] = (double)(v - (REAL)(double)v);
            }
            t_hi[b] = (double)t[l].hi;
            t_lo[b] = (double)(t[l].lo + (t[l].hi - (REAL)(double)t[l].hi));
            h_out[b] = (double)h[l];
            oc_out[b] = oc[l];
            if (tc_out)
                for (uint32_t i = 0; i < n; ++i)
                    for (uint32_t k = 0; k <= order; ++k)
                        tc_out[((size_t)i * (order + 1) + k) * B + b] = (double)TP(&wk, i, k, l);
        }
    }
    free(x);
    FN(work_free)(&wk);
    free(s.rec);
    return 0;
}

typedef struct {
    const FN(sys) * s;
    uint32_t order, B, L;
    int high_acc;
    double *state, *state_lo;
    const double *pars;
    double *t_hi, *t_lo;
    const double *tf_hi, *tf_lo, *max_delta_t;
    uint64_t max_steps;
    int64_t *oc;
    double *min_h, *max_h;
    uint64_t *nsteps;
    const double *grid; /* propagate_grid: [npts][B] times (NULL for propagate_until) */
    uint64_t npts;
    double *gout; /* propagate_grid: [npts][n][B] */
    volatile uint32_t next_chunk;
    int err;
} FN(prop_job);

/* Dense output of lane l (instance b) at the grid points inside the step of
   size h that ended at time t: the step's Taylor polynomials evaluated at the
   offset (g - t) + h from the start of the step, as heyoka's update_d_output
   forms it (heyoka/expose_batch_integrators.cpp:505-541). fin: the step reached
   the final time, so every remaining point is evaluated. */
static void FN(grid_eval)(FN(prop_job) * jb, FN(work) * wk, uint32_t l, uint32_t b, uint64_t *gc, FN(dfl) t, REAL h,
                          int tdir, int fin)
{
    const uint32_t n = jb->s->n_state, P = wk->P, B = jb->B;
    while (*gc < jb->npts) {
        REAL d = ((REAL)jb->grid[*gc * B + b] - t.hi) - t.lo;
        int inside = tdir ? (d <= 0) : (d >= 0);
        if (!inside && !fin) break;
        REAL dt = d + h;
        for (uint32_t i = 0; i < n; ++i) {
            REAL r = TP(wk, i, P, l);
            for (int k = (int)P - 1; k >= 0; --k) r = r * dt + TP(wk, i, k, l);
            jb->gout[(*gc * n + i) * B + b] = (double)r;
        }
        ++*gc;
    }
}

/* propagate_until for the lanes of one chunk, run in lock-step like one batch
   integrator of the reference: lanes that are done take zero-length steps. */
static void FN(prop_chunk)(FN(prop_job) * jb, FN(work) * wk, uint32_t b0, uint32_t nl)
{
    const FN(sys) *s = jb->s;
    const uint32_t n = s->n_state, L = jb->L, B = jb->B;
    const REAL rf = FN(rhofac)(jb->order);
    REAL *x = (REAL *)malloc(sizeof(REAL) * (size_t)n * L);
    FN(dfl) t[64], tf[64];
    REAL mdt[64], h[64], min_h[64], max_h[64], lim[64], remv[64];
    int64_t oc[64], res_oc[64];
    uint64_t nst[64], gc[64];
    int done[64], tdir[64];
    for (uint32_t l = 0; l < L; ++l) {
        uint32_t b = b0 + (l < nl ? l : 0);
        for (uint32_t i = 0; i < n; ++i) x[(size_t)i * L + l] = (REAL)jb->state[(size_t)i * B + b];
        for (uint32_t q = 0; q < s->n_par; ++q) wk->pars[(size_t)q * L + l] = (REAL)jb->pars[(size_t)q * B + b];
        t[l] = FN(dfl_norm)((REAL)jb->t_hi[b], (REAL)jb->t_lo[b]);
        tf[l] = FN(dfl_norm)((REAL)jb->tf_hi[b], (REAL)jb->tf_lo[b]);
        mdt[l] = (REAL)jb->max_delta_t[b];
        FN(dfl) rem = FN(dfl_add)(tf[l], FN(dfl_neg)(t[l]));
        tdir[l] = rem.hi >= 0;
        min_h[l] = (REAL)INFINITY;
        max_h[l] = 0;
        nst[l] = 0;
        done[l] = (l >= nl);
        res_oc[l] = ORA_TIME_LIMIT;
        gc[l] = 1;
        if (jb->grid && l < nl)
            for (uint32_t i = 0; i < n; ++i) jb->gout[(size_t)i * B + b] = jb->state[(size_t)i * B + b];
    }
    uint64_t iter = 0;
    for (;;) {
        for (uint32_t l = 0; l < L; ++l) {
            if (done[l]) {
                lim[l] = 0;
                continue;
            }
            /* the clamp is the remaining time, or max_delta_t if smaller */
            REAL rem = FN(dfl_add)(tf[l], FN(dfl_neg)(t[l])).hi;
            remv[l] = rem;
            lim[l] = tdir[l] ? ((mdt[l] < rem) ? mdt[l] : rem) : ((-mdt[l] > rem) ? -mdt[l] : rem);
        }
        FN(step)(s, wk, rf, jb->high_acc, x, t, lim, h, oc);
        ++iter;
        int all_done = 1;
        for (uint32_t l = 0; l < L; ++l) {
            if (done[l]) continue;
            if (h[l] != 0) ++nst[l];
            if (oc[l] == ORA_SUCCESS) {
                REAL ah = R_FABS(h[l]);
                min_h[l] = (ah < min_h[l]) ? ah : min_h[l];
                max_h[l] = (ah > max_h[l]) ? ah : max_h[l];
            }
            if (jb->grid && oc[l] != ORA_ERR_NF_STATE)
                FN(grid_eval)(jb, wk, l, b0 + l, &gc[l], t[l], h[l], tdir[l],
                              oc[l] == ORA_TIME_LIMIT && h[l] == remv[l]);
            if (oc[l] == ORA_ERR_NF_STATE) {
                res_oc[l] = ORA_ERR_NF_STATE;
                done[l] = 1;
            } else if (oc[l] == ORA_TIME_LIMIT && h[l] == remv[l]) {
                res_oc[l] = ORA_TIME_LIMIT;
                done[l] = 1;
            } else if (jb->max_steps != 0 && iter >= jb->max_steps) {
                res_oc[l] = ORA_STEP_LIMIT;
                done[l] = 1;
            }
            all_done = all_done && done[l];
        }
        if (all_done) break;
    }
    for (uint32_t l = 0; l < nl; ++l) {
        uint32_t b = b0 + l;
        for (uint32_t i = 0; i < n; ++i) {
            REAL v = x[(size_t)i * L + l];
            jb->state[(size_t)i * B + b] = (double)v;
            if (jb->state_lo) jb->state_lo[(size_t)i * B + b] = (double)(v - (REAL)(double)v);
        }
        jb->t_hi[b] = (double)t[l].hi;
        jb->t_lo[b] = (double)(t[l].lo + (t[l].hi - (REAL)(double)t[l].hi));
        jb->oc[b] = res_oc[l];
        jb->min_h[b] = (double)min_h[l];
        jb->max_h[b] = (double)max_h[l];
        jb->nsteps[b] = nst[l];
    }
    free(x);
}

static void *FN(prop_worker)(void *arg)
{
    FN(prop_job) *jb = (FN(prop_job) *)arg;
    FN(work) wk;
    FN(work_init)(&wk, jb->s, jb->order, jb->L);
    for (;;) {
        uint32_t c = __atomic_fetch_add(&jb->next_chunk, 1u, __ATOMIC_RELAXED);
        uint64_t b0 = (uint64_t)c * jb->L;
        if (b0 >= jb->B) break;
        uint32_t nl = (uint32_t)((jb->B - b0 < jb->L) ? jb->B - b0 : jb->L);
        FN(prop_chunk)(jb, &wk, (uint32_t)b0, nl);
    }
    FN(work_free)(&wk);
    return NULL;
}

static void FN(prop_run)(FN(prop_job) * jb, int nthreads)
{
    jb->next_chunk = 0;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    if (nthreads == 1) {
        FN(prop_worker)(jb);
    } else {
        pthread_t th[256];
        for (int i = 0; i < nthreads; ++i) pthread_create(&th[i], NULL, FN(prop_worker), jb);
        for (int i = 0; i < nthreads; ++i) pthread_join(th[i], NULL);
    }
}

/* propagate_until for B lanes, processed in lock-step chunks of lane_batch lanes
   (1..64) by nthreads threads. tf = (tf_hi, tf_lo) is the final time per lane,
   max_delta_t[b] > 0 (+inf for none). */
int FN(propagate_until)(const uint32_t *words, const double *consts, uint32_t order, int high_acc, uint32_t B,
                        uint32_t lane_batch, double *state, double *state_lo, const double *pars, double *t_hi,
                        double *t_lo, const double *tf_hi, const double *tf_lo, const double *max_delta_t,
                        uint64_t max_steps, int64_t *oc, double *min_h, double *max_h, uint64_t *nsteps,
                        int nthreads)
{
    FN(sys) s;
    if (FN(parse)(&s, words, consts) != 0) return -1;
    if (order < 2 || lane_batch == 0 || lane_batch > 64) {
        free(s.rec);
        return -2;
    }
    FN(prop_job) jb;
    memset(&jb, 0, sizeof(jb));
    jb.s = &s;
    jb.order = order;
    jb.B = B;
    jb.L = lane_batch;
    jb.high_acc = high_acc;
    jb.state = state;
    jb.state_lo = state_lo;
    jb.pars = pars;
    jb.t_hi = t_hi;
    jb.t_lo = t_lo;
    jb.tf_hi = tf_hi;
    jb.tf_lo = tf_lo;
    jb.max_delta_t = max_delta_t;
    jb.max_steps = max_steps;
    jb.oc = oc;
    jb.min_h = min_h;
    jb.max_h = max_h;
    jb.nsteps = nsteps;
    FN(prop_run)(&jb, nthreads);
    free(s.rec);
    return 0;
}

/* propagate_grid (heyoka/expose_batch_integrators.cpp:315-392): propagate_until
   to the last grid time of each lane, writing the state at every grid time
   ([npts][B], first row = current time) into gout [npts][n][B] by dense output
   of the step containing it. gout must be prefilled (NaN) by the caller: points
   a lane does not reach keep that value. */
int FN(propagate_grid)(const uint32_t *words, const double *consts, uint32_t order, int high_acc, uint32_t B,
                       uint32_t lane_batch, double *state, const double *pars, double *t_hi, double *t_lo,
                       const double *grid, uint64_t npts, const double *max_delta_t, uint64_t max_steps, int64_t *oc,
                       double *min_h, double *max_h, uint64_t *nsteps, double *gout, int nthreads)
{
    if (npts == 0) return -3;
    FN(sys) s;
    if (FN(parse)(&s, words, consts) != 0) return -1;
    if (order < 2 || lane_batch == 0 || lane_batch > 64) {
        free(s.rec);
        return -2;
    }
    double *tf_lo = (double *)calloc(B ? B : 1, sizeof(double));
    FN(prop_job) jb;
    memset(&jb, 0, sizeof(jb));
    jb.s = &s;
    jb.order = order;
    jb.B = B;
    jb.L = lane_batch;
    jb.high_acc = high_acc;
    jb.state = state;
    jb.pars = pars;
    jb.t_hi = t_hi;
    jb.t_lo = t_lo;
    jb.tf_hi = grid + (size_t)(npts - 1) * B;
    jb.tf_lo = tf_lo;
    jb.max_delta_t = max_delta_t;
    jb.max_steps = max_steps;
    jb.oc = oc;
    jb.min_h = min_h;
    jb.max_h = max_h;
    jb.nsteps = nsteps;
    jb.grid = grid;
    jb.npts = npts;
    jb.gout = gout;
    FN(prop_run)(&jb, nthreads);
    free(tf_lo);
    free(s.rec);
    return 0;
}

/* rho factor actually used, for tests. */
double FN(rhofac_value)(uint32_t order) { return (double)FN(rhofac)(order); }

#undef TP
#undef FN
#undef CAT
#undef CAT_
