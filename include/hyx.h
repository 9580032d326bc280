/* C-ABI of the MI355X batch/ensemble Taylor integrator (libhyx.so).
 *
 * This is the drop-in boundary for heyoka.py's batch integrator path. Each entry
 * point replaces one piece of what the pybind11 class
 * heyoka.core.taylor_adaptive_batch_dbl binds from heyoka C++
 * (heyoka/expose_batch_integrators.cpp:78-673); the comment on each cites it.
 * Signatures use plain pointers and sizes only. Every function returns 0 on
 * success and a negative HYX_E* code on failure; hyx_last_error() then gives a
 * message (thread-local). One batch object must not be used by two threads at
 * once; distinct objects may be used concurrently.
 *
 * Layouts (identical to the reference's, :101-119): state is [n][N] and
 * parameters [n_par][N], row-major doubles with the instance index fastest;
 * per-instance vectors are [N].
 */
#ifndef HYX_H
#define HYX_H

#include <stddef.h>
#include <stdint.h>

#include "hyx_oplist.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hyx_sys hyx_sys;
typedef struct hyx_batch hyx_batch;

/* Per-instance result of a propagation, as in propagate_res
 * (heyoka/expose_batch_integrators.cpp:393): outcome is a taylor_outcome value
 * (heyoka/core.cpp:324-329). */
typedef struct {
    int64_t outcome;
    double min_h;
    double max_h;
    uint64_t nsteps;
} hyx_prop_res;

/* taylor_outcome values. */
#define HYX_OUTCOME_SUCCESS (-4294967297LL)
#define HYX_OUTCOME_STEP_LIMIT (-4294967298LL)
#define HYX_OUTCOME_TIME_LIMIT (-4294967299LL)
#define HYX_OUTCOME_ERR_NF_STATE (-4294967300LL)

/* Error codes. */
#define HYX_OK 0
#define HYX_EINVAL (-1)
#define HYX_EHIP (-2)
#define HYX_ERTC (-3)
#define HYX_ENOMEM (-4)
#define HYX_ENODEV (-5)

/* hyx_sys_create flags. */
#define HYX_FLAG_HIGH_ACCURACY 1u /* compensated summation in the state update */
#define HYX_FLAG_COMPILE_ONLY 2u  /* compile into the code-object cache; no device needed */
#define HYX_FLAG_FORCE_COMPACT 4u /* loop over orders, Taylor tape in device memory */
#define HYX_FLAG_FORCE_UNROLLED 8u /* fully unrolled, histories in registers */
#define HYX_FLAG_SOURCE_ONLY 16u   /* generate the HIP source only (hyx_sys_source); no compile, no device */
#define HYX_FLAG_FORCE_GROUP 32u   /* group form: G lanes per instance over isomorphic clusters */
#define HYX_FLAG_GRID 64u          /* hyx_sys_precompile: the propagate_grid code object */
#define HYX_FLAG_SPLIT_CHAINS 128u /* two accumulators per convolution sum of >= 8 terms (default: one) */
#define HYX_FLAG_NO_PIPELINE 256u  /* group form: two LDS exchanges per order even where one suffices */
/* Compact form: lanes per instance (default 2); each lane sums every n-th convolution term. */
#define HYX_FLAG_SUB1 512u
#define HYX_FLAG_SUB4 1024u
#define HYX_FLAG_SUB8 2048u
/* Wave form: the nodes spread over the lanes of a group (64 lanes per instance by default), every
 * coefficient row in LDS. HYX_FLAG_WAVE32 / HYX_FLAG_WAVE16: 32 or 16 lanes per instance. */
#define HYX_FLAG_FORCE_WAVE 4096u
#define HYX_FLAG_WAVE32 8192u
#define HYX_FLAG_WAVE16 16384u
/* Unrolled form: keep every order of the state variables in LDS rows (HYX_FLAG_LDS_ROWS) or
 * in registers (HYX_FLAG_NO_LDS_ROWS); neither: chosen from the register budget. */
#define HYX_FLAG_LDS_ROWS 32768u
#define HYX_FLAG_NO_LDS_ROWS 65536u
/* Unrolled form: no twin form (see hyx_batch_set_twin_tail). */
#define HYX_FLAG_NO_TWIN 131072u
/* Unrolled form: the twin form (two lanes per instance) continues the last instances even
 * where the quad form (four lanes) applies. */
#define HYX_FLAG_NO_QUAD 262144u

/* Code forms (hyx_sys_code_form). */
#define HYX_FORM_UNROLLED 0 /* one instance per lane, fully unrolled, histories in registers */
#define HYX_FORM_COMPACT 1  /* one instance per lane, loop over orders, Taylor tape in device memory */
#define HYX_FORM_GROUP 2    /* G lanes per instance, one cluster per lane, histories in registers, LDS exchange */
#define HYX_FORM_WAVE 3     /* G lanes per instance, nodes spread over the lanes, coefficient rows in LDS */

/* Fields for hyx_batch_write / hyx_batch_read / hyx_batch_device_ptr. */
#define HYX_FIELD_STATE 0   /* [n][N] */
#define HYX_FIELD_PARS 1    /* [n_par][N] */
#define HYX_FIELD_TIME_HI 2 /* [N] */
#define HYX_FIELD_TIME_LO 3 /* [N] */

/* Propagation forms for hyx_propagate_until. */
#define HYX_MODE_PERSISTENT 0 /* one persistent work-stealing launch (resumed if the budget runs out) */
#define HYX_MODE_PLAIN 1      /* one kernel launch per integration step */
#define HYX_MODE_GRAPH 2      /* the plain form's launches replayed from a captured hipGraph */

/* Build the integrator for one ODE system: parse the op-list (hyx_oplist.h),
 * generate its Taylor-step kernels and compile them for the current device's
 * architecture (gfx950) with hipRTC, through an in-memory and on-disk
 * code-object cache. `order` is the Taylor order p = ceil(-ln(tol)/2 + 1).
 * Replaces the construction of heyoka::taylor_adaptive_batch<double>
 * (heyoka/expose_batch_integrators.cpp:91-212): taylor_decompose + LLVM JIT,
 * whose caches are enabled at heyoka/core.cpp:148-149. */
int hyx_sys_create(const uint32_t *ops, uint32_t n_words, const double *consts, uint32_t n_consts, uint32_t order,
                   uint32_t flags, hyx_sys **out);
/* As hyx_sys_create with HYX_FLAG_COMPILE_ONLY, for an explicit target such as "gfx950". */
int hyx_sys_precompile(const uint32_t *ops, uint32_t n_words, const double *consts, uint32_t n_consts,
                       uint32_t order, uint32_t flags, const char *arch);
/* order / dim properties (expose_batch_integrators.cpp:542-544). */
int hyx_sys_order(const hyx_sys *s);
int hyx_sys_dim(const hyx_sys *s);
int hyx_sys_npar(const hyx_sys *s);
/* 1 if the compact code form (tape in device memory) was generated, 0 otherwise. */
int hyx_sys_is_compact(const hyx_sys *s);
/* The generated code form: HYX_FORM_UNROLLED, HYX_FORM_COMPACT, HYX_FORM_GROUP or HYX_FORM_WAVE. */
int hyx_sys_code_form(const hyx_sys *s);
/* Lanes of a wavefront that one instance occupies (64 / instances per wavefront). No
 * counterpart in the reference, whose instances are SIMD lanes of a CPU batch. */
int hyx_sys_lanes_per_instance(const hyx_sys *s);
/* The generated HIP source (for inspection); returns its length, copies up to cap-1 bytes. */
size_t hyx_sys_source(const hyx_sys *s, char *buf, size_t cap);
/* Register usage of the compiled Taylor-step kernels: VGPRs, scratch bytes per lane. */
int hyx_sys_kernel_info(const hyx_sys *s, int *vgprs, int *scratch_bytes);
void hyx_sys_destroy(hyx_sys *s);

/* N device-resident instances of one system (the batch_size of
 * taylor_adaptive_batch, :116). State, parameters and time start at zero. */
int hyx_batch_create(hyx_sys *s, uint64_t n_inst, hyx_batch **out);
void hyx_batch_destroy(hyx_batch *b);
uint64_t hyx_batch_size(const hyx_batch *b);

/* Copy a field host -> device / device -> host; the state / pars / time
 * accessors of the reference (:394-472) map onto these. */
int hyx_batch_write(hyx_batch *b, int field, const double *src);
int hyx_batch_read(hyx_batch *b, int field, double *dst);
/* Device pointer of a field (device-resident use, no copies). */
int hyx_batch_device_ptr(hyx_batch *b, int field, void **out);

/* One adaptive step of every instance: step() / step(max_delta_t)
 * (:233-242). max_dt: [N] signed clamps, or NULL for +inf (step_backward passes
 * -inf). outcome / h: [N] host outputs (step_res), may be NULL. write_tc keeps the
 * Taylor coefficients for hyx_read_tc (tc property, :473-489). */
int hyx_step(hyx_batch *b, const double *max_dt, int write_tc, int64_t *outcome, double *h);
/* Taylor coefficients of the last step with write_tc: [n][p+1][N]. */
int hyx_read_tc(hyx_batch *b, double *dst);

/* propagate_until (:279-314) for every instance: tf_hi / tf_lo ([N], tf_lo may be
 * NULL) is the double-length final time; max_delta_t: [N] positive values or NULL
 * (no limit); max_steps: 0 = unlimited; res: [N] host output, may be NULL. */
int hyx_propagate_until(hyx_batch *b, const double *tf_hi, const double *tf_lo, const double *max_delta_t,
                        uint64_t max_steps, int mode, hyx_prop_res *res);

/* No counterpart in the reference: propagate_until for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). It runs the kernels of hyx_propagate_until's persistent form over the
 * caller's device arrays: state [n][N], pars [n_par][N] (NULL when the system has no
 * parameters), t_hi / t_lo [N] (read, then written with the final times), tf_hi [N],
 * tf_lo [N] or NULL (0), max_delta_t [N] or NULL (no limit); outputs outcome, min_h,
 * max_h, nsteps [N] (propagate_res). It allocates nothing: its scratch (per-instance
 * bookkeeping, work-queue counters, default arrays, the compact form's Taylor tape) is a
 * fixed device array of HYX_NOALLOC_SCRATCH_BYTES per GPU, held from the first call on a
 * device until the process ends; a call whose scratch would not fit fails with HYX_EINVAL
 * and a message giving the bytes needed (about 20-44 bytes per instance, plus the tape).
 * The kernels are enqueued on `stream` (a hipStream_t; NULL = the null stream) and the call
 * returns without waiting: results are valid once the stream has completed. Every call on
 * one GPU shares that GPU's scratch, so calls that overlap on one GPU race. Systems with
 * events are not supported. The current device must be the system's. */
#define HYX_NOALLOC_SCRATCH_BYTES (64u << 20)
int hyx_propagate_until_noalloc(hyx_sys *s, uint64_t n_inst, double *state, const double *pars, double *t_hi,
                                double *t_lo, const double *tf_hi, const double *tf_lo, const double *max_delta_t,
                                uint64_t max_steps, int64_t *outcome, double *min_h, double *max_h,
                                uint64_t *nsteps, void *stream);

/* propagate_grid (heyoka/expose_batch_integrators.cpp:315-392): propagates every
 * instance to its last grid time and writes the state at each grid time by dense
 * output of the step that contains it. grid: [n_pts][N] host times; per instance
 * finite, strictly monotonic, the first equal to the current time. out:
 * [n_pts][n][N] host output (the reference's (n_pts, n, B) array); points an
 * instance does not reach (step limit, non-finite state) are NaN. Steps are those
 * of propagate_until to the last grid time. HYX_MODE_GRAPH runs as HYX_MODE_PLAIN.
 * The first call on a system compiles its grid code object (HYX_GRID). */
int hyx_propagate_grid(hyx_batch *b, const double *grid, uint64_t n_pts, const double *max_delta_t,
                       uint64_t max_steps, int mode, hyx_prop_res *res, double *out);

/* Event detection: the t_events / nt_events of taylor_adaptive_batch
 * (heyoka/expose_batch_integrators.cpp:166-208, events bound at
 * heyoka/taylor_expose_events.cpp:195-300). The event op-list has a row per state
 * variable of the system, in the same order (any further rows are zero-valued
 * padding), and the event functions as the right-hand sides of its first
 * n_te + n_nt rows, terminal events first. direction: [n_te + n_nt] values -1
 * (negative), 0 (any), 1 (positive); cooldown: [n_te], negative = deduced from tol
 * and the event's time derivative at the trigger. Systems with events use the lane
 * code forms (unrolled or compact); every step detects the events of every instance:
 * a terminal trigger shortens the step to it (outcome -(index + 1)), and propagations
 * end there; non-terminal triggers are recorded for hyx_events_read. */
typedef struct {
    const uint32_t *ops;
    uint32_t n_words;
    const double *consts;
    uint32_t n_consts;
    uint32_t n_te, n_nt;
    const int32_t *direction;
    const double *cooldown;
    const int32_t *stopping; /* [n_te]: 1 = no callback, the event always stops; NULL = all 1 */
    double tol;
} hyx_events;
int hyx_sys_create_events(const uint32_t *ops, uint32_t n_words, const double *consts, uint32_t n_consts,
                          uint32_t order, uint32_t flags, const hyx_events *ev, hyx_sys **out);
/* with_events / len(t_events) / len(nt_events) (expose_batch_integrators.cpp:651-656). */
int hyx_sys_n_events(const hyx_sys *s, uint32_t *n_te, uint32_t *n_nt);
/* Events of the last hyx_step (what heyoka's batch step passes to the callbacks):
 * te [N] 4 * (terminal index) + d_sgn + 1, or -1; nt_count [N] non-terminal
 * triggers; nt_time / nt_code [cap][N] the first min(count, cap) of them: trigger
 * time and 4 * (non-terminal index) + d_sgn + 1. A count above cap means records
 * were lost. Any pointer may be NULL. */
int hyx_events_read(hyx_batch *b, int64_t *te, uint32_t *nt_count, double *nt_time, int64_t *nt_code,
                    uint32_t *cap);
/* Parking (no counterpart in the reference, whose propagate loop calls the callbacks
 * between lockstep steps): with pause on, hyx_propagate_until / _grid park an instance
 * at the end of each step whose events need the host (a non-terminal trigger, or a
 * terminal one of an event with a callback) and return when every instance is done
 * or parked. hyx_events_parked lists the parked instances (lanes: [N] capacity);
 * hyx_events_gather fetches, per listed instance, what that step's callbacks see: te,
 * nt_count, nt_time / nt_code [cap][count], the step size h [count] and, if tc is not
 * NULL, the step's Taylor coefficients [n][p+1][count]; hyx_events_resume sets each
 * listed instance running again (stop[c] == 0) or done with outcome stop[c];
 * hyx_propagate_resume continues the propagation (grid_out: the output of a resumed
 * propagate_grid, else NULL). */
int hyx_batch_set_event_pause(hyx_batch *b, int on);
int hyx_events_parked(hyx_batch *b, uint32_t *lanes, uint32_t *count);
int hyx_events_gather(hyx_batch *b, const uint32_t *lanes, uint32_t count, int64_t *te, uint32_t *nt_count,
                      double *nt_time, int64_t *nt_code, double *h, double *tc);
int hyx_events_resume(hyx_batch *b, const uint32_t *lanes, uint32_t count, const int64_t *stop);
int hyx_propagate_resume(hyx_batch *b, uint64_t max_steps, int mode, hyx_prop_res *res, double *grid_out);

/* Continuous output on the device (c_output=True of propagate_until / propagate_for,
 * heyoka/taylor_expose_c_output.cpp:269-520): with a step log of cap slots per
 * instance, propagations write every step's Taylor coefficients and its start time
 * (hi, lo) and size to the instance's log, and an instance whose log is full parks
 * (hyx_events_parked / hyx_events_resume with stop = NULL). hyx_clog_drain copies the
 * logs of the listed instances out and empties them: nrec [count] steps logged, tc
 * [cap][n][p+1][count] and tt [cap][3][count] (any may be NULL). cap = 0 removes the
 * log. Lane code forms, systems without events. */
int hyx_batch_set_clog(hyx_batch *b, uint32_t cap);
int hyx_clog_drain(hyx_batch *b, const uint32_t *lanes, uint32_t count, uint32_t *nrec, double *tc, double *tt);
/* te_cooldowns / reset_cooldowns (expose_batch_integrators.cpp:652-654): [n_te][3][N]
 * trigger time (hi, lo) and cooldown length; a negative length = no cooldown. */
int hyx_cooldowns_read(hyx_batch *b, double *dst);
int hyx_cooldowns_write(hyx_batch *b, const double *src);

/* Kernel time (ms, HIP events on the batch's stream) and launch count of the
 * last hyx_propagate_until / hyx_step call. */
int hyx_batch_last_timing(const hyx_batch *b, double *kernel_ms, uint64_t *launches);
/* (No counterpart in the reference: diagnostics of the tail hand-off.) Of the last
 * persistent hyx_propagate_until, the time (ms) from the first tail-form launch to the
 * end; 0 when no instance was handed off. */
int hyx_batch_last_tail_ms(const hyx_batch *b, double *tail_ms);
/* Persistent-form tuning: waves per CU (0 = occupancy limit) and the idle-lane
 * count that triggers a refill (1..64). */
int hyx_batch_set_persistent_config(hyx_batch *b, uint32_t waves_per_cu, uint32_t refill_min);
/* Persistent form: iterations one instance may take per launch before it is
 * written back and the kernel relaunched (default 65536). Bounds the duration of
 * every launch; results do not depend on it. */
int hyx_batch_set_iter_budget(hyx_batch *b, uint64_t iters);
/* Autotuner (no counterpart in the reference, whose LLVM JIT has no launch shape): times the persistent propagation of the batch's current contents
 * (state, pars, time as written) to tf_hi [N] for every candidate (waves per CU in
 * {occupancy, 1, 2, 4, 8} x refill threshold in {1, 8, 32}), restoring the contents
 * before each run and at the end; keeps the fastest as the batch's persistent config.
 * With cache_path, a result cached for the same code object, device and batch size is
 * reused (from_cache = 1), and a new one is appended. Launch parameters never change
 * results, so tuning is invisible to the arithmetic. */
typedef struct {
    uint32_t waves_per_cu;
    uint32_t refill_min;
    double rate; /* instance-steps/s of the winning run (kernel time) */
    int from_cache;
} hyx_tune_res;
int hyx_autotune(hyx_batch *b, const double *tf_hi, const char *cache_path, hyx_tune_res *best);

/* Twin form (no counterpart in the reference, whose instances are CPU SIMD lanes): an
 * unrolled system whose right-hand sides add isomorphic subexpressions (the planar CR3BP:
 * the pull of each primary) also gets a form with two lanes per instance that splits the
 * isomorphic nodes between the lanes with the unrolled form's arithmetic, bit for bit;
 * where the twin's nodes split again under a pairing of the state variables (the planar
 * CR3BP: x <-> y, vx <-> vy), the quad form takes its place, four lanes per instance, the
 * same arithmetic (HYX_FLAG_NO_QUAD: keep the twin). A persistent propagation hands its
 * last instances to it once its queue is drained and at most tail_min instances are in
 * flight (default: eight per CU; 0: never), so the longest-running instances of a ragged
 * ensemble take lower-latency steps. Results do not depend on the hand-off.
 * hyx_sys_twin_source: that form's HIP source (length 0: the system has none);
 * hyx_batch_twin_launches: its launches in the batch's propagations so far. */
int hyx_batch_set_twin_tail(hyx_batch *b, uint32_t tail_min);
size_t hyx_sys_twin_source(const hyx_sys *s, char *buf, size_t cap);
uint64_t hyx_batch_twin_launches(const hyx_batch *b);

/* Persistent-form instance order (a permutation of 0..N-1), or NULL for identity. */
int hyx_batch_set_queue(hyx_batch *b, const uint32_t *order);
int hyx_batch_sync(hyx_batch *b);

/* Continuous batching over the instance slots of one batch (serving). No
 * counterpart in the reference, which has no serving loop: these split
 * hyx_propagate_until into arm / bounded epochs / harvest so that slots can be
 * armed and harvested while others are still propagating. Every slot's
 * arithmetic is that of hyx_propagate_until, so results do not depend on what
 * else shares the batch. One epoch at a time per batch; distinct batches
 * (streams) may run epochs concurrently. */
int hyx_slots_reset(hyx_batch *b); /* mark every slot idle */
/* Scatter / gather one field for a list of slots: src / dst are [rows][count]. */
int hyx_slots_write(hyx_batch *b, int field, const uint32_t *slots, uint32_t count, const double *src);
int hyx_slots_read(hyx_batch *b, int field, const uint32_t *slots, uint32_t count, double *dst);
/* Arm slots for propagate_until(tf) (tf_lo, max_delta_t may be NULL). */
int hyx_slots_arm(hyx_batch *b, const uint32_t *slots, uint32_t count, const double *tf_hi, const double *tf_lo,
                  const double *max_delta_t);
/* Start one epoch (asynchronous): every armed slot advances by at most iter_budget steps. */
int hyx_slots_launch(hyx_batch *b, uint64_t iter_budget);
/* Wait for the epoch; n_active = armed slots still unfinished. */
int hyx_slots_wait(hyx_batch *b, uint32_t *n_active);
/* Non-blocking: *finished = 1 when the epoch's kernels are complete (or none runs),
 * so a server can harvest each batch as soon as its own epoch ends. */
int hyx_slots_poll(hyx_batch *b, int *finished);
/* Snapshot / restore of slots (no counterpart in the reference: serving eviction under
 * HBM pressure; the continuation contract is that of the reference's pickling,
 * heyoka/pickle_wrappers.hpp:28-74): every per-instance word a propagation depends on
 * (state, pars, time, final time, max_delta_t, min_h / max_h, nsteps, iterations,
 * outcome, status), hyx_slots_snapshot_words() words per slot, column-major per field.
 * Restoring into any free slots of a batch of the same system continues bit for bit. */
uint64_t hyx_slots_snapshot_words(const hyx_batch *b);
int hyx_slots_snapshot(hyx_batch *b, const uint32_t *slots, uint32_t count, uint64_t *dst);
int hyx_slots_restore(hyx_batch *b, const uint32_t *slots, uint32_t count, const uint64_t *src);
/* done[c] = 1 once slot c finished (then res[c] holds its propagate result). */
int hyx_slots_results(hyx_batch *b, const uint32_t *slots, uint32_t count, uint32_t *done, hyx_prop_res *res);

/* Compiled functions: heyoka's cfunc (heyoka/cfunc.cpp:90-121, the call operator
 * :121-440) evaluated on the GPU, one lane per point. The op-list holds the function's
 * variables as its rows 0..n_in-1 (any further rows are zero-valued padding) and the
 * outputs as the right-hand sides of rows 0..n_out-1; numbers, par[] and time as in a
 * system. */
typedef struct hyx_cfunc hyx_cfunc;
int hyx_cfunc_create(const uint32_t *ops, uint32_t n_words, const double *consts, uint32_t n_consts, uint32_t n_in,
                     uint32_t n_out, uint32_t flags /* HYX_FLAG_COMPILE_ONLY / SOURCE_ONLY */, hyx_cfunc **out);
size_t hyx_cfunc_source(const hyx_cfunc *c, char *buf, size_t cap);
/* Host arrays: in [n_in][n_points], pars [npar][n_points] or NULL, time [n_points] or
 * NULL (0), out [n_out][n_points]. */
int hyx_cfunc_eval(hyx_cfunc *c, const double *in, const double *pars, uint32_t npar, const double *time,
                   uint64_t n_points, double *out);
/* Device diagnostics: evaluate at every instance of a batch from its device-resident
 * state, parameters and time (the function's variables must be the batch's state
 * variables, in order), ordered after the batch's pending work; out [n_out][N] host. */
int hyx_cfunc_eval_batch(hyx_cfunc *c, hyx_batch *b, double *out);
void hyx_cfunc_destroy(hyx_cfunc *c);

const char *hyx_last_error(void);
/* Number of HIP devices visible (0 without a GPU). */
int hyx_device_count(void);

#ifdef __cplusplus
}
#endif

#endif
