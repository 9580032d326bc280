// Hand-written CDNA4 kernels of the batch/ensemble Taylor integrator: the part
// shared by the lane form (hyx_klane.hip) and the group form (hyx_kgroup.hip).
//
// These files are compiled at run time with hipRTC, after a per-system prelude
// that the runtime generates from the op-list (hyx_codegen.cpp).
//
// Layout in HBM: every per-instance array is structure-of-arrays with the
// instance index fastest (state[i * ld + inst]), so the 64 lanes of a wave
// touch 512 contiguous bytes per access.
//
// Semantics follow heyoka's batch integrator (heyoka/expose_batch_integrators.cpp):
//   - step(max_dt): |h| clamped to |max_dt|, sign of max_dt; outcome
//     time_limit when h == max_dt, err_nf_state on a non-finite state/time
//     (:233-242, doc/notebooks/Batch mode overview.ipynb:242-245,320);
//   - propagate_until(t): steps clamped to the remaining time and to
//     max_delta_t, double-length time, per-instance (outcome, min_h, max_h,
//     nsteps) with min_h/max_h over unclamped steps only (:256-314, :393).
// Each instance is independent: one instance's error never stops another.

typedef unsigned long long hyx_u64;
typedef long long hyx_i64;

#define HYX_SUCCESS (-4294967297LL)
#define HYX_STEP_LIMIT (-4294967298LL)
#define HYX_TIME_LIMIT (-4294967299LL)
#define HYX_ERR_NF_STATE (-4294967300LL)
#define HYX_PARKED (1LL << 40) // internal: the instance parks for the host (never an outcome)

// Kernel argument block. The host fills the identical struct (hyx_runtime.cpp).
struct hyx_kargs {
    double *state;          // [n][ld]
    const double *pars;     // [npar][ld]
    double *t_hi;           // [N]
    double *t_lo;           // [N]
    const double *tf_hi;    // [N] final time (propagate)
    const double *tf_lo;    // [N]
    const double *max_dt;   // [N] step: signed clamp; propagate: max_delta_t > 0
    double *h_out;          // [N] step: h taken
    hyx_i64 *outcome;       // [N]
    double *min_h;          // [N]
    double *max_h;          // [N]
    hyx_u64 *nsteps;        // [N]
    hyx_u64 *iters;         // [N] plain form: iterations taken so far
    unsigned *status;       // [N] plain form: 0 active forward, 2 active backward, 1 done
    double *tc;             // [n][P+1][ld] Taylor coefficients (step with write_tc)
    double *tape;           // per-lane scratch tape for compact code (may be null)
    hyx_u64 *work_counter;  // persistent form: next queue position
    unsigned *active_count; // plain form: active instances after this launch
    const unsigned *queue;  // persistent form: instance order (null = identity)
    hyx_u64 N;
    hyx_u64 ld;
    hyx_u64 max_steps;      // 0 = unlimited
    hyx_u64 tape_stride;    // lanes per tape row
    hyx_u64 iter_budget;    // persistent form: iterations per instance per launch
    unsigned *unfinished;   // persistent form: instances suspended by the budget
    unsigned refill_min;    // persistent form: refill when this many lanes idle
    unsigned pad_;
    const double *grid;     // propagate_grid: [npts][ld] output times per instance
    double *gout;           // propagate_grid: [npts][n][ld] dense output at the grid times
    unsigned *gcur;         // propagate_grid: [N] next grid point of each instance
    hyx_u64 npts;           // propagate_grid: grid points (0 outside propagate_grid)
    double *et;             // events: [event nodes * (P+1)][ld] event-function coefficients
    double *evcd;           // events: [n_te][3][ld] cooldowns: trigger time hi, lo, length (< 0: none)
    hyx_i64 *te;            // events: [N] last step's terminal event: 4 * index + d_sgn + 1, or -1
    unsigned *ntc;          // events: [N] non-terminal events found in the last step
    double *nt_time;        // events: [nt_cap][ld] their times
    hyx_i64 *nt_code;       // events: [nt_cap][ld] 4 * (non-terminal index) + d_sgn + 1
    unsigned nt_cap;        // events: records kept per instance (0: none)
    unsigned ev_pause;      // events: propagations park a lane after a step that needs the host
    double *clog;           // continuous output: [clog_cap][n * (P+1)][ld] coefficients of each step
    double *clog_t;         // continuous output: [clog_cap][3][ld] start time hi, lo and size of each step
    unsigned *clog_n;       // continuous output: [N] steps logged since the host last drained
    unsigned clog_cap;      // continuous output: steps per instance before it parks
    unsigned pad3_;
    unsigned *inflight;     // persistent lane form with a twin form: instances in flight (null: off)
    unsigned *tailq;        // persistent form: [N] instances written back unfinished, by slot of *unfinished
    unsigned tail_min;      // hand-off: once drained and at most this many in flight, write back (0: never)
    unsigned pad4_;
};

// Propagation status of an instance: 0 / 2 running forward / backward, 1 done,
// 4 / 6 parked (forward / backward) after a step whose events need the host.
#define HYX_ST_DONE 1u
#define HYX_ST_PARKED 4u
__device__ __forceinline__ bool hyx_runnable(unsigned st) { return (st & ~2u) == 0u; }

// ---- double-length time -----------------------------------------------------------
// Knuth's error-free sum (TAOCP vol. 2, 4.2.2).
__device__ __forceinline__ void hyx_two_sum(double a, double b, double &s, double &e)
{
    double x = a + b;
    double z = x - a;
    e = (a - (x - z)) + (b - z);
    s = x;
}

// Dekker's fast error-free sum, |a| >= |b|.
__device__ __forceinline__ void hyx_fast_two_sum(double a, double b, double &s, double &e)
{
    double x = a + b;
    e = b - (x - a);
    s = x;
}

// (hi, lo) += (bh, bl), renormalised.
__device__ __forceinline__ void hyx_dfl_add(double &hi, double &lo, double bh, double bl)
{
    double s, e, t, f;
    hyx_two_sum(hi, bh, s, e);
    hyx_two_sum(lo, bl, t, f);
    e += t;
    hyx_fast_two_sum(s, e, s, e);
    e += f;
    hyx_fast_two_sum(s, e, hi, lo);
}

// hi part of (ah, al) - (bh, bl).
__device__ __forceinline__ double hyx_dfl_sub_hi(double ah, double al, double bh, double bl)
{
    double hi = ah, lo = al;
    hyx_dfl_add(hi, lo, -bh, -bl);
    return hi;
}


// ---- slot operations (serving: continuous batching over the instances of a batch) ----------
// 64-bit moves, so outcome / nsteps (integers) and doubles travel bit-exactly.
extern "C" __global__ void hyx_scatter_k(hyx_u64 *dst, hyx_u64 ld, unsigned rows, const hyx_u64 *src,
                                         const unsigned *idx, unsigned count)
{
    const hyx_u64 t = (hyx_u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (hyx_u64)rows * count) return;
    const hyx_u64 r = t / count, c = t % count;
    dst[r * ld + idx[c]] = src[r * count + c];
}

extern "C" __global__ void hyx_gather_k(const hyx_u64 *srcm, hyx_u64 ld, unsigned rows, hyx_u64 *dst,
                                        const unsigned *idx, unsigned count)
{
    const hyx_u64 t = (hyx_u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (hyx_u64)rows * count) return;
    const hyx_u64 r = t / count, c = t % count;
    dst[r * count + c] = srcm[r * ld + idx[c]];
}

extern "C" __global__ void hyx_gather32_k(const unsigned *srcm, unsigned *dst, const unsigned *idx, unsigned count)
{
    const unsigned c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < count) dst[c] = srcm[idx[c]];
}

extern "C" __global__ void hyx_scatter32_k(unsigned *dst, const unsigned *src, const unsigned *idx, unsigned count)
{
    const unsigned c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < count) dst[idx[c]] = src[c];
}

extern "C" __global__ void hyx_fill_u32_k(unsigned *p, unsigned v, hyx_u64 n)
{
    const hyx_u64 i = (hyx_u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Grid-stride fill of n 64-bit words (the NaN prefill of propagate_grid's output).
extern "C" __global__ void hyx_fill_u64_k(hyx_u64 *p, hyx_u64 v, hyx_u64 n)
{
    const hyx_u64 stride = (hyx_u64)gridDim.x * blockDim.x;
    for (hyx_u64 i = (hyx_u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

// Dense-output time of grid point g for a step of size h that ended at
// (t_hi, t_lo): the offset from the start of the step, as heyoka's
// update_d_output forms it ((g - t) + h; heyoka/expose_batch_integrators.cpp:505-541,
// doc/notebooks/Batch mode overview.ipynb:702).
// *inside is false when g lies beyond the end of the step in the direction of
// integration.
__device__ __forceinline__ double hyx_grid_offset(double g, double t_hi, double t_lo, double h, int tdir, bool &inside)
{
    const double d = (g - t_hi) - t_lo;
    inside = tdir ? (d <= 0.0) : (d >= 0.0);
    return d + h;
}

// Arm slots for a propagation: final time, max_delta_t, direction and zeroed
// bookkeeping, exactly as hyx_iter_init_k does for a whole batch.
// stage = [tf_hi(count), tf_lo(count), max_dt(count)].
extern "C" __global__ void hyx_slots_arm_k(hyx_kargs a, const unsigned *idx, unsigned count, const double *stage)
{
    const unsigned c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= count) return;
    const hyx_u64 i = idx[c];
    const double th = stage[c], tl = stage[count + c];
    ((double *)a.tf_hi)[i] = th;
    ((double *)a.tf_lo)[i] = tl;
    ((double *)a.max_dt)[i] = stage[2 * (hyx_u64)count + c];
    const double rem = hyx_dfl_sub_hi(th, tl, a.t_hi[i], a.t_lo[i]);
    a.status[i] = (rem >= 0.0) ? 0u : 2u;
    a.iters[i] = 0;
    a.nsteps[i] = 0;
    a.min_h[i] = __builtin_inf();
    a.max_h[i] = 0.0;
    a.outcome[i] = HYX_TIME_LIMIT;
}
