// Host runtime of the MI355X Taylor integrator: the C-ABI of include/hyx.h.
//
// hyx_sys:   a parsed op-list, its generated HIP source (hyx_codegen.cpp +
//            hyx_kernels.hip) and the hipRTC-compiled module. Code objects are
//            cached in memory and on disk, keyed by a hash of source, options
//            and target, as heyoka caches its LLVM JIT output
//            (heyoka/core.cpp:148-149, 191-297).
// hyx_batch: N device-resident instances (SoA, instance index fastest) plus the
//            per-instance bookkeeping of propagate_until, on one HIP stream.

#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <dlfcn.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <limits>
#include <mutex>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/hyx.h"
#include "hyx_internal.h"

#include "build/hyx_kernels_embed.inc" // k_hyx_kcommon, k_hyx_klane, k_hyx_kgroup, k_hyx_kcfunc, k_hyx_kevent: kernel templates

namespace
{

thread_local std::string g_err;

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

#define HIPCHK(x)                                                                                                      \
    do {                                                                                                               \
        hipError_t e_ = (x);                                                                                           \
        if (e_ != hipSuccess) return fail(HYX_EHIP, std::string(#x) + ": " + hipGetErrorString(e_));                \
    } while (0)

// Kernel argument block; must match struct hyx_kargs in hyx_kernels.hip.
struct KArgs {
    double *state;
    const double *pars;
    double *t_hi;
    double *t_lo;
    const double *tf_hi;
    const double *tf_lo;
    const double *max_dt;
    double *h_out;
    int64_t *outcome;
    double *min_h;
    double *max_h;
    uint64_t *nsteps;
    uint64_t *iters;
    unsigned *status;
    double *tc;
    double *tape;
    uint64_t *work_counter;
    unsigned *active_count;
    const unsigned *queue;
    uint64_t N;
    uint64_t ld;
    uint64_t max_steps;
    uint64_t tape_stride;
    uint64_t iter_budget;
    unsigned *unfinished;
    unsigned refill_min;
    unsigned pad_;
    const double *grid;
    double *gout;
    unsigned *gcur;
    uint64_t npts;
    double *et;
    double *evcd;
    int64_t *te;
    unsigned *ntc;
    double *nt_time;
    int64_t *nt_code;
    unsigned nt_cap;
    unsigned ev_pause;
    double *clog;
    double *clog_t;
    unsigned *clog_n;
    unsigned clog_cap;
    unsigned pad3_;
    unsigned *inflight;
    unsigned *tailq;
    unsigned tail_min;
    unsigned pad4_;
};

// ---------------------------------------------------------------------------------------------
// Code-object cache.

std::mutex g_cache_mu;
std::unordered_map<std::string, std::vector<char>> g_mem_cache;

uint64_t fnv1a(const std::string &s, uint64_t h = 1469598103934665603ull)
{
    for (unsigned char c : s) {
        h ^= c;
        h *= 1099511628211ull;
    }
    return h;
}

std::string lib_dir()
{
    Dl_info info;
    if (dladdr((void *)&fnv1a, &info) && info.dli_fname) {
        std::string p = info.dli_fname;
        auto pos = p.rfind('/');
        if (pos != std::string::npos) return p.substr(0, pos);
    }
    return ".";
}

std::string cache_dir()
{
    const char *e = std::getenv("HYX_CACHE_DIR");
    if (e && *e) return e;
    return lib_dir() + "/code_cache";
}

std::vector<std::string> rtc_options(const std::string &arch)
{
    return {"--offload-arch=" + arch, "-O3", "-std=c++17", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
}

int get_code(const std::string &src, const std::string &arch, std::vector<char> &code)
{
    const auto opts = rtc_options(arch);
    std::string keysrc = src;
    for (auto &o : opts) keysrc += "\n" + o;
    char kbuf[32];
    std::snprintf(kbuf, sizeof(kbuf), "%016llx", (unsigned long long)fnv1a(keysrc));
    const std::string key = arch + "-" + kbuf;
    {
        std::lock_guard<std::mutex> lk(g_cache_mu);
        auto it = g_mem_cache.find(key);
        if (it != g_mem_cache.end()) {
            code = it->second;
            return 0;
        }
    }
    const std::string dir = cache_dir();
    const std::string path = dir + "/" + key + ".co";
    {
        std::ifstream f(path, std::ios::binary);
        if (f) {
            code.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
            if (!code.empty()) {
                std::lock_guard<std::mutex> lk(g_cache_mu);
                g_mem_cache[key] = code;
                return 0;
            }
        }
    }
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, src.c_str(), "hyx_system.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS)
        return fail(HYX_ERTC, "hiprtcCreateProgram failed");
    std::vector<const char *> copts;
    for (auto &o : opts) copts.push_back(o.c_str());
    const hiprtcResult rc = hiprtcCompileProgram(prog, (int)copts.size(), copts.data());
    if (rc != HIPRTC_SUCCESS) {
        size_t ls = 0;
        hiprtcGetProgramLogSize(prog, &ls);
        std::string log(ls + 1, '\0');
        hiprtcGetProgramLog(prog, &log[0]);
        hiprtcDestroyProgram(&prog);
        return fail(HYX_ERTC, "hipRTC compilation failed: " + log.substr(0, 4000));
    }
    size_t cs = 0;
    hiprtcGetCodeSize(prog, &cs);
    code.resize(cs);
    hiprtcGetCode(prog, code.data());
    hiprtcDestroyProgram(&prog);
    {
        std::lock_guard<std::mutex> lk(g_cache_mu);
        g_mem_cache[key] = code;
    }
    // Best-effort disk cache: write to a temporary name, then rename.
    ::mkdir(dir.c_str(), 0755);
    const std::string tmp = path + ".tmp" + std::to_string((long)::getpid());
    {
        std::ofstream f(tmp, std::ios::binary);
        if (f) f.write(code.data(), (std::streamsize)code.size());
    }
    std::rename(tmp.c_str(), path.c_str());
    return 0;
}

std::string device_arch(int dev)
{
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, dev) != hipSuccess) return "gfx950";
    std::string a = p.gcnArchName;
    auto pos = a.find(':');
    return pos == std::string::npos ? a : a.substr(0, pos);
}

} // namespace

// ---------------------------------------------------------------------------------------------

struct hyx_sys {
    hyx::Sys S;
    uint32_t order = 0;
    uint32_t flags = 0;
    bool compact = false;
    int form = 0;                 // HYX_FORM_UNROLLED / COMPACT / GROUP / WAVE
    uint32_t inst_per_block = 64; // instances handled by one 64-lane block
    int sub = 1;                  // lanes per instance (lane forms)
    std::string src;
    std::vector<char> code;
    int device = -1;
    hipModule_t mod = nullptr;
    hipFunction_t k_step = nullptr, k_step_tc = nullptr, k_init = nullptr, k_iter = nullptr, k_persist = nullptr,
                  k_clr32 = nullptr, k_clr64 = nullptr, k_scatter = nullptr, k_gather = nullptr, k_gather32 = nullptr,
                  k_scatter32 = nullptr,
                  k_fill32 = nullptr, k_fill64 = nullptr, k_arm = nullptr;
    int vgprs = -1, scratch = -1;
    // propagate_grid: the same source built with HYX_GRID, loaded on first use.
    std::mutex gmu;
    hipModule_t gmod = nullptr;
    hipFunction_t kg_init = nullptr, kg_iter = nullptr, kg_persist = nullptr;
    // Continuous output: the same source built with HYX_CLOG, loaded on first use.
    hipModule_t cmod = nullptr;
    hipFunction_t kc_init = nullptr, kc_iter = nullptr, kc_persist = nullptr;
    // Twin form (unrolled systems whose right-hand sides add isomorphic terms): its own
    // source and module, whose persistent kernel continues the last instances of a
    // persistent propagation two lanes per instance (empty source: no twin form), or four
    // lanes per instance in the quad form where that applies (tsub: lanes per instance).
    std::string tsrc;
    int tsub = 2;
    hipModule_t tmod = nullptr;
    hipFunction_t kt_persist = nullptr;
    // Events: the event op-list (rows = state variables, right-hand sides = event
    // functions, terminal ones first) and its specification.
    bool has_events = false;
    hyx::Sys E;
    hyx::EventSpec ev;
    uint64_t ev_rows = 0; // event-node coefficient rows per instance
};

struct hyx_batch {
    hyx_sys *sys = nullptr;
    uint64_t N = 0;
    int device = 0;
    int n_cu = 0;
    hipStream_t stream = nullptr;
    double *state = nullptr, *pars = nullptr, *thi = nullptr, *tlo = nullptr, *tfhi = nullptr, *tflo = nullptr,
           *mdt = nullptr, *hout = nullptr, *minh = nullptr, *maxh = nullptr, *tc = nullptr;
    int64_t *oc = nullptr;
    uint64_t *nsteps = nullptr, *iters = nullptr, *work = nullptr;
    unsigned *status = nullptr, *active = nullptr, *unfinished = nullptr, *queue = nullptr;
    double *tape = nullptr;     // compact code: [nodes * (p + 1)][tape_stride]
    uint64_t tape_stride = 0;
    uint32_t waves_per_cu = 0, refill_min = 1;
    uint64_t iter_budget = 1ull << 16;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipEvent_t evt = nullptr; // the first tail-form launch of a persistent propagation
    double last_ms = 0.0;
    double last_tail_ms = 0.0; // from that launch to the end (0: no hand-off)
    uint64_t last_launches = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    void *stage = nullptr;       // serving: device staging for slot data
    size_t stage_bytes = 0;
    unsigned *stage_idx = nullptr;
    size_t stage_idx_n = 0;
    bool slots_running = false;
    double *grid = nullptr, *gout = nullptr; // propagate_grid: [npts][N] times, [npts][n][N] output
    unsigned *gcur = nullptr;                // propagate_grid: [N] next grid point
    size_t grid_cap = 0, gout_cap = 0;       // allocated doubles
    uint64_t npts = 0;
    // Events (systems with events only): et = the event-node coefficient rows, which
    // follow the state rows inside tc; [n_te][3][N] cooldowns, the last step's terminal event and its direction,
    // and its non-terminal records ([nt_cap][N] offsets and codes).
    double *et = nullptr, *evcd = nullptr, *nt_time = nullptr;
    int64_t *te = nullptr, *nt_code = nullptr;
    unsigned *ntc = nullptr;
    unsigned nt_cap = 0;
    bool ev_pause = false;       // propagations park instances whose events need the host
    bool last_grid = false;      // the last propagation was a propagate_grid (for resumes)
    // Continuous output: per-instance logs of step coefficients / start times and sizes.
    double *clog = nullptr, *clog_t = nullptr;
    unsigned *clog_n = nullptr;
    unsigned clog_cap = 0;
    // Twin hand-off of persistent propagations: the in-flight counter, two lists of
    // instances written back unfinished ([N] each) and the hand-off threshold (0: off).
    unsigned *inflight = nullptr, *tailq = nullptr;
    uint32_t tail_min = 0;
    uint64_t twin_handoffs = 0; // tail-form launches of the last propagations (diagnostics)
};

namespace
{

const int kPlainGroup = 16; // plain form: launches between host checks

int build_sys(const uint32_t *ops, uint32_t n_words, const double *consts, uint32_t n_consts, uint32_t order,
              uint32_t flags, hyx_sys &s, const hyx_events *ev = nullptr)
{
    if (!ops) return fail(HYX_EINVAL, "null op-list");
    std::string err;
    if (hyx::parse_oplist(ops, n_words, consts, n_consts, s.S, err) != 0) return fail(HYX_EINVAL, err);
    if (order < 2 || order > 200) return fail(HYX_EINVAL, "the Taylor order must be in [2, 200]");
    s.order = order;
    s.flags = flags;
    if (ev) {
        const uint32_t nev = ev->n_te + ev->n_nt;
        if (nev == 0) return fail(HYX_EINVAL, "an event specification needs at least one event");
        if (!ev->ops || (nev && !ev->direction) || (ev->n_te && !ev->cooldown))
            return fail(HYX_EINVAL, "null pointer in the event specification");
        if (hyx::parse_oplist(ev->ops, ev->n_words, ev->consts, ev->n_consts, s.E, err) != 0)
            return fail(HYX_EINVAL, "event op-list: " + err);
        if (s.E.n < s.S.n || s.E.n < nev)
            return fail(HYX_EINVAL, "the event op-list must have a row per state variable and per event");
        if (s.E.npar > s.S.npar)
            return fail(HYX_EINVAL, "the event functions use more parameters than the system");
        s.ev.n_te = ev->n_te;
        s.ev.tol = ev->tol;
        for (uint32_t e = 0; e < nev; ++e) {
            if (ev->direction[e] < -1 || ev->direction[e] > 1) return fail(HYX_EINVAL, "invalid event direction");
            s.ev.dir.push_back(ev->direction[e]);
        }
        for (uint32_t e = 0; e < ev->n_te; ++e) {
            if (!(ev->cooldown[e] == ev->cooldown[e])) return fail(HYX_EINVAL, "a NaN event cooldown was given");
            s.ev.cooldown.push_back(ev->cooldown[e]);
            s.ev.stopping.push_back(ev->stopping ? (ev->stopping[e] != 0) : 1);
        }
        if (!(ev->tol > 0.0)) return fail(HYX_EINVAL, "the event specification needs the tolerance");
        s.has_events = true;
        s.ev_rows = (uint64_t)(s.E.nnodes - s.E.n) * (order + 1);
    }
    try {
        const bool ha = (flags & HYX_FLAG_HIGH_ACCURACY) != 0;
        hyx::GroupPlan gp;
        std::string why;
        const bool lane_compact = hyx::choose_compact(s.S, (int)order);
        if (flags & HYX_FLAG_FORCE_COMPACT)
            s.form = HYX_FORM_COMPACT;
        else if (flags & HYX_FLAG_FORCE_UNROLLED)
            s.form = HYX_FORM_UNROLLED;
        else if (flags & HYX_FLAG_FORCE_GROUP) {
            if (!hyx::plan_groups(s.S, gp, why)) return fail(HYX_EINVAL, "the group form does not apply: " + why);
            s.form = HYX_FORM_GROUP;
        } else if (flags & HYX_FLAG_FORCE_WAVE)
            s.form = HYX_FORM_WAVE;
        else if (lane_compact && hyx::plan_groups(s.S, gp, why))
            s.form = HYX_FORM_GROUP;
        else
            s.form = lane_compact ? HYX_FORM_COMPACT : HYX_FORM_UNROLLED;
        s.compact = s.form == HYX_FORM_COMPACT;
        const int split = (flags & HYX_FLAG_SPLIT_CHAINS) ? 8 : 0;
        if (s.has_events && s.ev.dir.size() > 64) return fail(HYX_EINVAL, "at most 64 events are supported");
        std::string wave_src;
        if (s.form == HYX_FORM_WAVE) {
            const int G = (flags & HYX_FLAG_WAVE16) ? 16 : (flags & HYX_FLAG_WAVE32) ? 32 : 64;
            wave_src = hyx::generate_wave_prelude(s.S, (int)order, ha, G, s.has_events ? &s.E : nullptr,
                                                  (uint32_t)s.ev.dir.size(), split, why);
            if (wave_src.empty()) return fail(HYX_EINVAL, "the wave form does not apply: " + why);
            s.inst_per_block = 64u / (uint32_t)G;
        } else if (s.form == HYX_FORM_COMPACT && !(flags & HYX_FLAG_FORCE_COMPACT)) {
            // Systems too large for the register forms: the wave form when its LDS rows fit,
            // 32 lanes per instance (measured on the J2 + variational ensemble: 5.9x the
            // compact form's 2-lane tape in device memory; 16 lanes: 5.1x; 64 lanes: 3.2x),
            // else 64.
            for (int G : {32, 64}) {
                wave_src = hyx::generate_wave_prelude(s.S, (int)order, ha, G, s.has_events ? &s.E : nullptr,
                                                      (uint32_t)s.ev.dir.size(), split, why);
                if (!wave_src.empty()) {
                    s.form = HYX_FORM_WAVE;
                    s.compact = false;
                    s.inst_per_block = 64u / (uint32_t)G;
                    break;
                }
            }
        }
        if (s.form == HYX_FORM_GROUP || s.form == HYX_FORM_WAVE) {
            if (s.form == HYX_FORM_GROUP) s.inst_per_block = 64u / (uint32_t)gp.G;
            s.src = (s.has_events ? hyx::events_forward_decl() : std::string()) +
                    (s.form == HYX_FORM_WAVE
                         ? wave_src
                         : hyx::generate_group_prelude(s.S, gp, (int)order, ha, split, !(flags & HYX_FLAG_NO_PIPELINE),
                                                       s.has_events ? &s.E : nullptr, (uint32_t)s.ev.dir.size())) +
                    "\n";
            if (s.has_events)
                s.src += hyx::generate_events_prelude(s.E, s.S.n, (int)order, s.ev, ha) + "\n" + k_hyx_kcommon + "\n" +
                         k_hyx_kevent + "\n";
            else
                s.src += std::string(k_hyx_kcommon) + "\n";
            s.src += k_hyx_kgroup;
        } else {
            // Compact form: lanes per instance (default 2: measured best on the J2 +
            // variational ensemble, 1.5x over one lane), each summing every sub-th
            // convolution term; the unrolled form keeps one lane per instance.
            int sub = 1;
            if (s.compact)
                sub = (flags & HYX_FLAG_SUB1) ? 1 : (flags & HYX_FLAG_SUB4) ? 4 : (flags & HYX_FLAG_SUB8) ? 8 : 2;
            s.inst_per_block = 64u / (uint32_t)sub;
            s.sub = sub;
            if (s.has_events && s.ev.dir.size() > 64) return fail(HYX_EINVAL, "at most 64 events are supported");
            if (s.form == HYX_FORM_UNROLLED && !s.has_events && !(flags & HYX_FLAG_NO_TWIN)) {
                std::string twhy;
                std::string tp;
                if (!(flags & HYX_FLAG_NO_QUAD)) tp = hyx::generate_quad_prelude(s.S, (int)order, ha, split, twhy);
                s.tsub = tp.empty() ? 2 : 4;
                if (tp.empty()) tp = hyx::generate_twin_prelude(s.S, (int)order, ha, split, twhy);
                if (!tp.empty()) s.tsrc = tp + "\n" + k_hyx_kcommon + "\n" + k_hyx_klane;
            }
            s.src = (s.has_events ? hyx::events_forward_decl() : std::string()) +
                    hyx::generate_prelude(s.S, (int)order, ha, s.compact, split, sub, s.has_events ? &s.E : nullptr,
                                          (uint32_t)s.ev.dir.size(),
                                          (flags & HYX_FLAG_LDS_ROWS) ? 1 : (flags & HYX_FLAG_NO_LDS_ROWS) ? 0 : -1) +
                    "\n";
            if (s.has_events)
                s.src += hyx::generate_events_prelude(s.E, s.S.n, (int)order, s.ev, ha) + "\n" + k_hyx_kcommon + "\n" +
                         k_hyx_kevent + "\n";
            else
                s.src += std::string(k_hyx_kcommon) + "\n";
            s.src += k_hyx_klane;
        }
    } catch (const std::exception &e) {
        return fail(HYX_EINVAL, e.what());
    }
    return 0;
}

KArgs make_args(hyx_batch *b)
{
    KArgs a;
    std::memset(&a, 0, sizeof(a));
    a.state = b->state;
    a.pars = b->pars;
    a.t_hi = b->thi;
    a.t_lo = b->tlo;
    a.tf_hi = b->tfhi;
    a.tf_lo = b->tflo;
    a.max_dt = b->mdt;
    a.h_out = b->hout;
    a.outcome = b->oc;
    a.min_h = b->minh;
    a.max_h = b->maxh;
    a.nsteps = b->nsteps;
    a.iters = b->iters;
    a.status = b->status;
    a.tc = b->tc;
    a.tape = b->tape;
    a.work_counter = b->work;
    a.active_count = b->active;
    a.queue = b->queue;
    a.N = b->N;
    a.ld = b->N;
    a.max_steps = 0;
    a.tape_stride = b->tape_stride;
    a.iter_budget = b->iter_budget;
    a.unfinished = b->unfinished;
    a.refill_min = b->refill_min;
    a.et = b->et;
    a.evcd = b->evcd;
    a.te = b->te;
    a.ntc = b->ntc;
    a.nt_time = b->nt_time;
    a.nt_code = b->nt_code;
    // Records are kept by hyx_step and by propagations that park.
    a.nt_cap = b->ev_pause ? b->nt_cap : 0;
    a.ev_pause = b->ev_pause ? 1u : 0u;
    a.clog = b->clog;
    a.clog_t = b->clog_t;
    a.clog_n = b->clog_n;
    a.clog_cap = b->clog_cap;
    return a;
}

hipError_t launch(hipFunction_t f, uint64_t grid, hipStream_t st, KArgs *a)
{
    void *params[] = {a};
    return hipModuleLaunchKernel(f, (unsigned)grid, 1, 1, 64, 1, 1, 0, st, params, nullptr);
}

hipError_t clear32(hyx_batch *b, unsigned *p, unsigned n)
{
    void *params[] = {&p, &n};
    return hipModuleLaunchKernel(b->sys->k_clr32, (n + 63) / 64, 1, 1, 64, 1, 1, 0, b->stream, params, nullptr);
}

hipError_t clear64(hyx_batch *b, uint64_t *p, unsigned n)
{
    void *params[] = {&p, &n};
    return hipModuleLaunchKernel(b->sys->k_clr64, (n + 63) / 64, 1, 1, 64, 1, 1, 0, b->stream, params, nullptr);
}

uint64_t blocks_for(uint64_t N) { return (N + 63) / 64; }

// The group and wave forms run the group kernels (G lanes per instance, LDS rows).
bool group_kernels(const hyx_sys *s) { return s->form == HYX_FORM_GROUP || s->form == HYX_FORM_WAVE; }

// Blocks of the step / plain / persistent kernels: one block holds 64 instances
// in the lane form and 64 / G in the group form.
uint64_t blocks_of(const hyx_batch *b) { return (b->N + b->sys->inst_per_block - 1) / b->sys->inst_per_block; }

// Compact code keeps its Taylor tape in device memory, one column per lane:
// lanes are instances for the step / plain kernels and persistent-grid lanes
// for the persistent kernel.
int ensure_tape(hyx_batch *b, uint64_t lanes)
{
    if (!b->sys->compact) return 0;
    const uint64_t want = std::max<uint64_t>(lanes, b->N);
    if (b->tape && b->tape_stride >= want) return 0;
    if (b->tape) HIPCHK(hipFree(b->tape));
    b->tape = nullptr;
    const uint64_t rows = (uint64_t)b->sys->S.nnodes * (b->sys->order + 1);
    if (hipMalloc((void **)&b->tape, rows * want * 8) != hipSuccess) {
        b->tape = nullptr;
        b->tape_stride = 0;
        return fail(HYX_ENOMEM, "hipMalloc of the Taylor tape failed (" + std::to_string(rows * want * 8) + " bytes)");
    }
    b->tape_stride = want;
    if (b->graph_exec) {
        hipGraphExecDestroy(b->graph_exec);
        hipGraphDestroy(b->graph);
        b->graph_exec = nullptr;
        b->graph = nullptr;
    }
    return 0;
}

int upload(hyx_batch *b, double *dst, const double *src, uint64_t count, double fill)
{
    if (src) {
        HIPCHK(hipMemcpyAsync(dst, src, count * sizeof(double), hipMemcpyHostToDevice, b->stream));
    } else {
        std::vector<double> tmp(count, fill);
        HIPCHK(hipMemcpyAsync(dst, tmp.data(), count * sizeof(double), hipMemcpyHostToDevice, b->stream));
        HIPCHK(hipStreamSynchronize(b->stream));
    }
    return 0;
}

} // namespace

extern "C" {

// ---- compiled functions ------------------------------------------------------------------

struct hyx_cfunc {
    hyx::Sys S;
    uint32_t n_in = 0, n_out = 0;
    std::string src;
    int device = -1;
    hipModule_t mod = nullptr;
    hipFunction_t k = nullptr;
    hipStream_t stream = nullptr;
    double *dbuf = nullptr; // device staging for host-array evaluations
    size_t dcap = 0;        // doubles
};

int hyx_cfunc_create(const uint32_t *ops, uint32_t n_words, const double *consts, uint32_t n_consts, uint32_t n_in,
                     uint32_t n_out, uint32_t flags, hyx_cfunc **out)
{
    if (!out || !ops) return fail(HYX_EINVAL, "null argument");
    *out = nullptr;
    auto *c = new hyx_cfunc;
    auto bad = [&](int code, const std::string &m) {
        hyx_cfunc_destroy(c);
        return fail(code, m);
    };
    std::string err;
    if (hyx::parse_oplist(ops, n_words, consts, n_consts, c->S, err) != 0) return bad(HYX_EINVAL, err);
    if (n_in > c->S.n || n_out > c->S.n || n_out == 0)
        return bad(HYX_EINVAL, "invalid input / output counts for the op-list of a compiled function");
    c->n_in = n_in;
    c->n_out = n_out;
    try {
        c->src = hyx::generate_cfunc_prelude(c->S, n_out) + "\n" + k_hyx_kcfunc;
    } catch (const std::exception &e) {
        return bad(HYX_EINVAL, e.what());
    }
    if (flags & HYX_FLAG_SOURCE_ONLY) {
        *out = c;
        return 0;
    }
    std::vector<char> code;
    if (flags & HYX_FLAG_COMPILE_ONLY) {
        if (int rc = get_code(c->src, "gfx950", code)) return bad(rc, hyx_last_error());
        *out = c;
        return 0;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return bad(HYX_ENODEV, "no HIP device is available; compiled functions run on an MI355X (gfx950)");
    if (hipGetDevice(&c->device) != hipSuccess) return bad(HYX_EHIP, "hipGetDevice failed");
    if (int rc = get_code(c->src, device_arch(c->device), code)) return bad(rc, hyx_last_error());
    if (hipModuleLoadData(&c->mod, code.data()) != hipSuccess) return bad(HYX_EHIP, "hipModuleLoadData failed");
    if (hipModuleGetFunction(&c->k, c->mod, "hyx_cfunc_k") != hipSuccess)
        return bad(HYX_EHIP, "hipModuleGetFunction(hyx_cfunc_k) failed");
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
        return bad(HYX_EHIP, "hipStreamCreate failed");
    *out = c;
    return 0;
}

size_t hyx_cfunc_source(const hyx_cfunc *c, char *buf, size_t cap)
{
    if (!c) return 0;
    if (buf && cap) {
        const size_t n = std::min(cap - 1, c->src.size());
        std::memcpy(buf, c->src.data(), n);
        buf[n] = '\0';
    }
    return c->src.size();
}

static int cfunc_launch(hyx_cfunc *c, hipStream_t st, const double *in, uint64_t ld_in, const double *pars,
                        uint64_t ld_par, uint32_t npar, const double *time, double *outp, uint64_t ld_out, uint64_t n)
{
    if (n == 0) return 0;
    void *params[] = {&in, &ld_in, &pars, &ld_par, &npar, &time, &outp, &ld_out, &n};
    HIPCHK(hipModuleLaunchKernel(c->k, (unsigned)((n + 63) / 64), 1, 1, 64, 1, 1, 0, st, params, nullptr));
    return 0;
}

int hyx_cfunc_eval(hyx_cfunc *c, const double *in, const double *pars, uint32_t npar, const double *time,
                   uint64_t n_points, double *out)
{
    if (!c || !in || !out) return fail(HYX_EINVAL, "null argument");
    if (!c->mod) return fail(HYX_EINVAL, "the compiled function was created without a device");
    if (npar > c->S.npar && c->S.npar > 0) npar = c->S.npar;
    const uint64_t N = n_points, nin = c->S.n, nout = c->n_out;
    const size_t need = (size_t)(nin + npar + 1 + nout) * N;
    if (need > c->dcap) {
        if (c->dbuf) HIPCHK(hipFree(c->dbuf));
        c->dbuf = nullptr;
        c->dcap = 0;
        HIPCHK(hipMalloc((void **)&c->dbuf, std::max<size_t>(need, 1) * 8));
        c->dcap = need;
    }
    double *din = c->dbuf, *dpar = din + nin * N, *dt = dpar + (size_t)npar * N, *dout = dt + N;
    // Inputs beyond the function's variables (padding rows of the op-list) are zero.
    HIPCHK(hipMemsetAsync(din, 0, nin * N * 8, c->stream));
    HIPCHK(hipMemcpyAsync(din, in, (size_t)c->n_in * N * 8, hipMemcpyHostToDevice, c->stream));
    if (npar && pars) HIPCHK(hipMemcpyAsync(dpar, pars, (size_t)npar * N * 8, hipMemcpyHostToDevice, c->stream));
    if (time) HIPCHK(hipMemcpyAsync(dt, time, N * 8, hipMemcpyHostToDevice, c->stream));
    if (int rc = cfunc_launch(c, c->stream, din, N, (npar && pars) ? dpar : nullptr, N, npar, time ? dt : nullptr, dout,
                              N, N))
        return rc;
    HIPCHK(hipMemcpyAsync(out, dout, nout * N * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

int hyx_cfunc_eval_batch(hyx_cfunc *c, hyx_batch *b, double *out)
{
    if (!c || !b || !out) return fail(HYX_EINVAL, "null argument");
    if (!c->mod) return fail(HYX_EINVAL, "the compiled function was created without a device");
    if (c->n_in != b->sys->S.n || c->S.n != b->sys->S.n)
        return fail(HYX_EINVAL, "the compiled function's variables must be the batch's state variables");
    const uint64_t N = b->N, nout = c->n_out;
    const uint32_t npar = std::min<uint32_t>(c->S.npar, b->sys->S.npar);
    if (c->dcap < nout * N) {
        if (c->dbuf) HIPCHK(hipFree(c->dbuf));
        c->dbuf = nullptr;
        c->dcap = 0;
        HIPCHK(hipMalloc((void **)&c->dbuf, std::max<size_t>(nout * N, 1) * 8));
        c->dcap = nout * N;
    }
    // On the batch's stream: ordered after its propagations, reading its device-resident state.
    if (int rc = cfunc_launch(c, b->stream, b->state, N, npar ? b->pars : nullptr, N, npar, b->thi, c->dbuf, N, N))
        return rc;
    HIPCHK(hipMemcpyAsync(out, c->dbuf, nout * N * 8, hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    return 0;
}

void hyx_cfunc_destroy(hyx_cfunc *c)
{
    if (!c) return;
    if (c->stream) hipStreamSynchronize(c->stream);
    if (c->dbuf) hipFree(c->dbuf);
    if (c->stream) hipStreamDestroy(c->stream);
    if (c->mod) hipModuleUnload(c->mod);
    delete c;
}

const char *hyx_last_error(void) { return g_err.c_str(); }

int hyx_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int hyx_sys_precompile(const uint32_t *ops, uint32_t n_words, const double *consts, uint32_t n_consts,
                       uint32_t order, uint32_t flags, const char *arch)
{
    hyx_sys s;
    int rc = build_sys(ops, n_words, consts, n_consts, order, flags, s);
    if (rc) return rc;
    std::vector<char> code;
    if (!(flags & HYX_FLAG_GRID) && !s.tsrc.empty())
        if (int rc2 = get_code(s.tsrc, arch ? arch : "gfx950", code)) return rc2;
    return get_code((flags & HYX_FLAG_GRID) ? "#define HYX_GRID 1\n" + s.src : s.src, arch ? arch : "gfx950", code);
}

static int create_sys(const uint32_t *ops, uint32_t n_words, const double *consts, uint32_t n_consts, uint32_t order,
                      uint32_t flags, const hyx_events *ev, hyx_sys **out)
{
    if (!out) return fail(HYX_EINVAL, "null output pointer");
    *out = nullptr;
    auto *s = new hyx_sys;
    int rc = build_sys(ops, n_words, consts, n_consts, order, flags, *s, ev);
    if (rc) {
        delete s;
        return rc;
    }
    if (flags & HYX_FLAG_SOURCE_ONLY) {
        *out = s;
        return 0;
    }
    if (flags & HYX_FLAG_COMPILE_ONLY) {
        rc = get_code(s->src, "gfx950", s->code);
        if (!rc && !s->tsrc.empty()) {
            std::vector<char> tcode;
            rc = get_code(s->tsrc, "gfx950", tcode);
        }
        if (rc) {
            delete s;
            return rc;
        }
        *out = s;
        return 0;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        delete s;
        return fail(HYX_ENODEV, "no HIP device is available; the hyx GPU integrator needs an MI355X (gfx950)");
    }
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        delete s;
        return fail(HYX_EHIP, "hipGetDevice failed");
    }
    s->device = dev;
    rc = get_code(s->src, device_arch(dev), s->code);
    if (rc) {
        delete s;
        return rc;
    }
    hipError_t e = hipModuleLoadData(&s->mod, s->code.data());
    if (e != hipSuccess) {
        delete s;
        return fail(HYX_EHIP, std::string("hipModuleLoadData: ") + hipGetErrorString(e));
    }
    struct {
        hipFunction_t *f;
        const char *name;
    } fns[] = {{&s->k_step, "hyx_step_k"},       {&s->k_step_tc, "hyx_step_tc_k"}, {&s->k_init, "hyx_iter_init_k"},
               {&s->k_iter, "hyx_iter_k"},       {&s->k_persist, "hyx_persist_k"}, {&s->k_clr32, "hyx_clear_u32_k"},
               {&s->k_clr64, "hyx_clear_u64_k"}, {&s->k_scatter, "hyx_scatter_k"}, {&s->k_gather, "hyx_gather_k"},
               {&s->k_gather32, "hyx_gather32_k"}, {&s->k_scatter32, "hyx_scatter32_k"}, {&s->k_fill32, "hyx_fill_u32_k"}, {&s->k_fill64, "hyx_fill_u64_k"},
               {&s->k_arm, "hyx_slots_arm_k"}};
    for (auto &f : fns) {
        e = hipModuleGetFunction(f.f, s->mod, f.name);
        if (e != hipSuccess) {
            hipModuleUnload(s->mod);
            delete s;
            return fail(HYX_EHIP, std::string("hipModuleGetFunction(") + f.name + "): " + hipGetErrorString(e));
        }
    }
    if (!s->tsrc.empty()) {
        std::vector<char> tcode;
        rc = get_code(s->tsrc, device_arch(dev), tcode);
        if (!rc && hipModuleLoadData(&s->tmod, tcode.data()) == hipSuccess &&
            hipModuleGetFunction(&s->kt_persist, s->tmod, "hyx_persist_k") == hipSuccess) {
            // loaded
        } else {
            if (s->tmod) hipModuleUnload(s->tmod);
            s->tmod = nullptr;
            s->kt_persist = nullptr;
            s->tsrc.clear();
            if (rc) {
                hipModuleUnload(s->mod);
                delete s;
                return rc;
            }
        }
    }
    int v = -1, sc = -1;
    if (hipFuncGetAttribute(&v, HIP_FUNC_ATTRIBUTE_NUM_REGS, s->k_persist) == hipSuccess) s->vgprs = v;
    if (hipFuncGetAttribute(&sc, HIP_FUNC_ATTRIBUTE_LOCAL_SIZE_BYTES, s->k_persist) == hipSuccess) s->scratch = sc;
    *out = s;
    return 0;
}

int hyx_sys_create(const uint32_t *ops, uint32_t n_words, const double *consts, uint32_t n_consts, uint32_t order,
                   uint32_t flags, hyx_sys **out)
{
    return create_sys(ops, n_words, consts, n_consts, order, flags, nullptr, out);
}

int hyx_sys_create_events(const uint32_t *ops, uint32_t n_words, const double *consts, uint32_t n_consts,
                          uint32_t order, uint32_t flags, const hyx_events *ev, hyx_sys **out)
{
    if (!ev) return fail(HYX_EINVAL, "null event specification");
    return create_sys(ops, n_words, consts, n_consts, order, flags, ev, out);
}

int hyx_sys_n_events(const hyx_sys *s, uint32_t *n_te, uint32_t *n_nt)
{
    if (!s) return fail(HYX_EINVAL, "null system");
    const uint32_t nev = (uint32_t)s->ev.dir.size();
    if (n_te) *n_te = s->has_events ? s->ev.n_te : 0;
    if (n_nt) *n_nt = s->has_events ? nev - s->ev.n_te : 0;
    return 0;
}

int hyx_sys_order(const hyx_sys *s) { return s ? (int)s->order : HYX_EINVAL; }
int hyx_sys_dim(const hyx_sys *s) { return s ? (int)s->S.n : HYX_EINVAL; }
int hyx_sys_npar(const hyx_sys *s) { return s ? (int)s->S.npar : HYX_EINVAL; }
int hyx_sys_is_compact(const hyx_sys *s) { return s ? (int)s->compact : HYX_EINVAL; }
int hyx_sys_code_form(const hyx_sys *s) { return s ? s->form : HYX_EINVAL; }
int hyx_sys_lanes_per_instance(const hyx_sys *s) { return s ? (int)(64u / s->inst_per_block) : HYX_EINVAL; }

size_t hyx_sys_twin_source(const hyx_sys *s, char *buf, size_t cap)
{
    if (!s) return 0;
    if (buf && cap) {
        const size_t n = std::min(cap - 1, s->tsrc.size());
        std::memcpy(buf, s->tsrc.data(), n);
        buf[n] = '\0';
    }
    return s->tsrc.size();
}

int hyx_batch_set_twin_tail(hyx_batch *b, uint32_t tail_min)
{
    if (!b) return fail(HYX_EINVAL, "null batch");
    if (tail_min > 0 && !b->sys->kt_persist) return fail(HYX_EINVAL, "the system has no twin form");
    b->tail_min = tail_min;
    return 0;
}

uint64_t hyx_batch_twin_launches(const hyx_batch *b) { return b ? b->twin_handoffs : 0; }

size_t hyx_sys_source(const hyx_sys *s, char *buf, size_t cap)
{
    if (!s) return 0;
    if (buf && cap) {
        const size_t n = std::min(cap - 1, s->src.size());
        std::memcpy(buf, s->src.data(), n);
        buf[n] = '\0';
    }
    return s->src.size();
}

int hyx_sys_kernel_info(const hyx_sys *s, int *vgprs, int *scratch_bytes)
{
    if (!s) return fail(HYX_EINVAL, "null system");
    if (vgprs) *vgprs = s->vgprs;
    if (scratch_bytes) *scratch_bytes = s->scratch;
    return 0;
}

void hyx_sys_destroy(hyx_sys *s)
{
    if (!s) return;
    if (s->cmod) hipModuleUnload(s->cmod);
    if (s->gmod) hipModuleUnload(s->gmod);
    if (s->tmod) hipModuleUnload(s->tmod);
    if (s->mod) hipModuleUnload(s->mod);
    delete s;
}

int hyx_batch_create(hyx_sys *s, uint64_t n_inst, hyx_batch **out)
{
    if (!s || !out) return fail(HYX_EINVAL, "null argument");
    if (!s->mod) return fail(HYX_EINVAL, "the system was created compile-only");
    if (n_inst == 0 || n_inst > (1ull << 32)) return fail(HYX_EINVAL, "the batch size must be in [1, 2^32]");
    *out = nullptr;
    auto *b = new hyx_batch;
    b->sys = s;
    b->N = n_inst;
    b->device = s->device;
    auto cleanup = [&](int code, const std::string &m) {
        hyx_batch_destroy(b);
        return fail(code, m);
    };
    if (hipSetDevice(b->device) != hipSuccess) return cleanup(HYX_EHIP, "hipSetDevice failed");
    hipDeviceGetAttribute(&b->n_cu, hipDeviceAttributeMultiprocessorCount, b->device);
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
        return cleanup(HYX_EHIP, "hipStreamCreate failed");
    hipEventCreate(&b->ev0);
    hipEventCreate(&b->ev1);
    hipEventCreate(&b->evt);
    const uint64_t N = n_inst, n = s->S.n, np = std::max<uint32_t>(s->S.npar, 1);
    struct {
        void **p;
        size_t bytes;
    } allocs[] = {
        {(void **)&b->state, n * N * 8},  {(void **)&b->pars, np * N * 8}, {(void **)&b->thi, N * 8},
        {(void **)&b->tlo, N * 8},        {(void **)&b->tfhi, N * 8},      {(void **)&b->tflo, N * 8},
        {(void **)&b->mdt, N * 8},        {(void **)&b->hout, N * 8},      {(void **)&b->minh, N * 8},
        {(void **)&b->maxh, N * 8},       {(void **)&b->oc, N * 8},        {(void **)&b->nsteps, N * 8},
        {(void **)&b->iters, N * 8},      {(void **)&b->status, N * 4},    {(void **)&b->work, 64},
        {(void **)&b->active, kPlainGroup * 4}, {(void **)&b->unfinished, 64},
    };
    for (auto &al : allocs) {
        if (hipMalloc(al.p, al.bytes) != hipSuccess) return cleanup(HYX_ENOMEM, "hipMalloc failed (device memory)");
        if (hipMemsetAsync(*al.p, 0, al.bytes, b->stream) != hipSuccess) return cleanup(HYX_EHIP, "hipMemset failed");
    }
    if (s->has_events) {
        // Every step of a system with events writes its coefficients (tc) and the
        // event functions' (et); non-terminal records: 16 per instance and step.
        const uint64_t P1 = s->order + 1, nte = std::max<uint32_t>(s->ev.n_te, 1);
        b->nt_cap = 16;
        struct {
            void **p;
            size_t bytes;
        } ev_allocs[] = {
            {(void **)&b->tc, (n * P1 + s->ev_rows + (uint64_t)s->ev.dir.size() * P1) * N * 8},
            {(void **)&b->evcd, 3 * nte * N * 8},
            {(void **)&b->te, N * 8},
            {(void **)&b->ntc, N * 4},
            {(void **)&b->nt_time, (uint64_t)b->nt_cap * N * 8},
            {(void **)&b->nt_code, (uint64_t)b->nt_cap * N * 8},
        };
        for (auto &al : ev_allocs) {
            if (hipMalloc(al.p, al.bytes) != hipSuccess) return cleanup(HYX_ENOMEM, "hipMalloc failed (device memory)");
            if (hipMemsetAsync(*al.p, 0, al.bytes, b->stream) != hipSuccess)
                return cleanup(HYX_EHIP, "hipMemset failed");
        }
        b->et = b->tc + n * P1 * N; // event-node rows follow the state rows
        // No cooldown running: every field -1 (only the length is read as a flag).
        const std::vector<double> m1(3 * nte * N, -1.0);
        if (hipMemcpyAsync(b->evcd, m1.data(), m1.size() * 8, hipMemcpyHostToDevice, b->stream) != hipSuccess)
            return cleanup(HYX_EHIP, "hipMemcpy failed");
        if (hipStreamSynchronize(b->stream) != hipSuccess) return cleanup(HYX_EHIP, "hipStreamSynchronize failed");
    }
    if (s->kt_persist) {
        // Twin hand-off bookkeeping; hand-off once the drained queue leaves at most eight
        // instances in flight per CU (the ragged CR3BP ensemble of 131,072 instances with
        // the quad form, three runs each: 2.50 s at two per CU, 2.30 s at eight, 2.36 s at
        // sixteen).
        if (hipMalloc((void **)&b->inflight, 64) != hipSuccess || hipMalloc((void **)&b->tailq, 2 * N * 4) != hipSuccess)
            return cleanup(HYX_ENOMEM, "hipMalloc failed (device memory)");
        b->tail_min = 8u * (uint32_t)std::max(b->n_cu, 1);
    }
    if (hipStreamSynchronize(b->stream) != hipSuccess) return cleanup(HYX_EHIP, "hipStreamSynchronize failed");
    *out = b;
    return 0;
}

void hyx_batch_destroy(hyx_batch *b)
{
    if (!b) return;
    if (b->stream) hipStreamSynchronize(b->stream);
    if (b->graph_exec) hipGraphExecDestroy(b->graph_exec);
    if (b->graph) hipGraphDestroy(b->graph);
    void *ptrs[] = {b->state, b->pars, b->thi,    b->tlo,   b->tfhi,   b->tflo,  b->mdt,
                    b->hout,  b->minh, b->maxh,   b->tc,    b->oc,     b->nsteps, b->iters,
                    b->work,  b->status, b->active, b->unfinished, b->queue, b->tape, b->stage, b->stage_idx,
                    b->grid,  b->gout, b->gcur, b->evcd, b->te, b->ntc, b->nt_time, b->nt_code,
                    b->clog, b->clog_t, b->clog_n, b->inflight, b->tailq};
    for (void *p : ptrs)
        if (p) hipFree(p);
    if (b->ev0) hipEventDestroy(b->ev0);
    if (b->ev1) hipEventDestroy(b->ev1);
    if (b->evt) hipEventDestroy(b->evt);
    if (b->stream) hipStreamDestroy(b->stream);
    delete b;
}

uint64_t hyx_batch_size(const hyx_batch *b) { return b ? b->N : 0; }

static int field_info(hyx_batch *b, int field, double **p, uint64_t *count)
{
    const uint64_t N = b->N;
    switch (field) {
    case HYX_FIELD_STATE: *p = b->state; *count = (uint64_t)b->sys->S.n * N; return 0;
    case HYX_FIELD_PARS: *p = b->pars; *count = (uint64_t)b->sys->S.npar * N; return 0;
    case HYX_FIELD_TIME_HI: *p = b->thi; *count = N; return 0;
    case HYX_FIELD_TIME_LO: *p = b->tlo; *count = N; return 0;
    default: return fail(HYX_EINVAL, "unknown field");
    }
}

int hyx_batch_write(hyx_batch *b, int field, const double *src)
{
    if (!b || !src) return fail(HYX_EINVAL, "null argument");
    double *p;
    uint64_t count;
    if (int rc = field_info(b, field, &p, &count)) return rc;
    if (count == 0) return 0;
    HIPCHK(hipMemcpyAsync(p, src, count * 8, hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    return 0;
}

int hyx_batch_read(hyx_batch *b, int field, double *dst)
{
    if (!b || !dst) return fail(HYX_EINVAL, "null argument");
    double *p;
    uint64_t count;
    if (int rc = field_info(b, field, &p, &count)) return rc;
    if (count == 0) return 0;
    HIPCHK(hipMemcpyAsync(dst, p, count * 8, hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    return 0;
}

int hyx_batch_device_ptr(hyx_batch *b, int field, void **out)
{
    if (!b || !out) return fail(HYX_EINVAL, "null argument");
    double *p;
    uint64_t count;
    if (int rc = field_info(b, field, &p, &count)) return rc;
    *out = p;
    return 0;
}

int hyx_step(hyx_batch *b, const double *max_dt, int write_tc, int64_t *outcome, double *h)
{
    if (!b) return fail(HYX_EINVAL, "null batch");
    const uint64_t N = b->N;
    if (write_tc && !b->tc) {
        const size_t bytes = (size_t)b->sys->S.n * (b->sys->order + 1) * N * 8;
        HIPCHK(hipMalloc((void **)&b->tc, bytes));
    }
    if (int rc = upload(b, b->mdt, max_dt, N, std::numeric_limits<double>::infinity())) return rc;
    if (int rc = ensure_tape(b, N)) return rc;
    KArgs a = make_args(b);
    a.nt_cap = b->nt_cap;
    a.ev_pause = 0;
    HIPCHK(hipEventRecord(b->ev0, b->stream));
    HIPCHK(launch(write_tc ? b->sys->k_step_tc : b->sys->k_step, blocks_of(b), b->stream, &a));
    HIPCHK(hipEventRecord(b->ev1, b->stream));
    if (outcome) HIPCHK(hipMemcpyAsync(outcome, b->oc, N * 8, hipMemcpyDeviceToHost, b->stream));
    if (h) HIPCHK(hipMemcpyAsync(h, b->hout, N * 8, hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    float ms = 0.f;
    hipEventElapsedTime(&ms, b->ev0, b->ev1);
    b->last_ms = ms;
    b->last_launches = 1;
    return 0;
}

int hyx_events_read(hyx_batch *b, int64_t *te, uint32_t *nt_count, double *nt_time, int64_t *nt_code, uint32_t *cap)
{
    if (!b) return fail(HYX_EINVAL, "null batch");
    if (!b->sys->has_events) return fail(HYX_EINVAL, "the system has no events");
    const uint64_t N = b->N;
    if (cap) *cap = b->nt_cap;
    if (te) HIPCHK(hipMemcpyAsync(te, b->te, N * 8, hipMemcpyDeviceToHost, b->stream));
    if (nt_count) HIPCHK(hipMemcpyAsync(nt_count, b->ntc, N * 4, hipMemcpyDeviceToHost, b->stream));
    if (nt_time)
        HIPCHK(hipMemcpyAsync(nt_time, b->nt_time, (uint64_t)b->nt_cap * N * 8, hipMemcpyDeviceToHost, b->stream));
    if (nt_code)
        HIPCHK(hipMemcpyAsync(nt_code, b->nt_code, (uint64_t)b->nt_cap * N * 8, hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    return 0;
}

int hyx_cooldowns_read(hyx_batch *b, double *dst)
{
    if (!b || !dst) return fail(HYX_EINVAL, "null argument");
    if (!b->sys->has_events) return fail(HYX_EINVAL, "the system has no events");
    const uint64_t cnt = 3ull * b->sys->ev.n_te * b->N;
    if (cnt == 0) return 0;
    HIPCHK(hipMemcpyAsync(dst, b->evcd, cnt * 8, hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    return 0;
}

int hyx_cooldowns_write(hyx_batch *b, const double *src)
{
    if (!b || !src) return fail(HYX_EINVAL, "null argument");
    if (!b->sys->has_events) return fail(HYX_EINVAL, "the system has no events");
    const uint64_t cnt = 3ull * b->sys->ev.n_te * b->N;
    if (cnt == 0) return 0;
    HIPCHK(hipMemcpyAsync(b->evcd, src, cnt * 8, hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    return 0;
}

int hyx_read_tc(hyx_batch *b, double *dst)
{
    if (!b || !dst) return fail(HYX_EINVAL, "null argument");
    if (!b->tc) return fail(HYX_EINVAL, "no Taylor coefficients: call hyx_step with write_tc first");
    const size_t bytes = (size_t)b->sys->S.n * (b->sys->order + 1) * b->N * 8;
    HIPCHK(hipMemcpyAsync(dst, b->tc, bytes, hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    return 0;
}

} // extern "C"

namespace
{

// Loads the propagate_grid code object of a system (the same source with
// HYX_GRID defined) on first use.
int ensure_grid_module(hyx_sys *s)
{
    std::lock_guard<std::mutex> lk(s->gmu);
    if (s->gmod) return 0;
    std::vector<char> code;
    if (int rc = get_code("#define HYX_GRID 1\n" + s->src, device_arch(s->device), code)) return rc;
    hipModule_t m = nullptr;
    HIPCHK(hipModuleLoadData(&m, code.data()));
    struct {
        hipFunction_t *f;
        const char *name;
    } fns[] = {{&s->kg_init, "hyx_iter_init_k"}, {&s->kg_iter, "hyx_iter_k"}, {&s->kg_persist, "hyx_persist_k"}};
    for (auto &f : fns) {
        hipError_t e = hipModuleGetFunction(f.f, m, f.name);
        if (e != hipSuccess) {
            hipModuleUnload(m);
            return fail(HYX_EHIP, std::string("hipModuleGetFunction(") + f.name + "): " + hipGetErrorString(e));
        }
    }
    s->gmod = m;
    return 0;
}

// Loads the continuous-output code object of a system (HYX_CLOG) on first use.
int ensure_clog_module(hyx_sys *s)
{
    std::lock_guard<std::mutex> lk(s->gmu);
    if (s->cmod) return 0;
    std::vector<char> code;
    if (int rc = get_code("#define HYX_CLOG 1\n" + s->src, device_arch(s->device), code)) return rc;
    hipModule_t m = nullptr;
    HIPCHK(hipModuleLoadData(&m, code.data()));
    struct {
        hipFunction_t *f;
        const char *name;
    } fns[] = {{&s->kc_init, "hyx_iter_init_k"}, {&s->kc_iter, "hyx_iter_k"}, {&s->kc_persist, "hyx_persist_k"}};
    for (auto &f : fns) {
        hipError_t e = hipModuleGetFunction(f.f, m, f.name);
        if (e != hipSuccess) {
            hipModuleUnload(m);
            return fail(HYX_EHIP, std::string("hipModuleGetFunction(") + f.name + "): " + hipGetErrorString(e));
        }
    }
    s->cmod = m;
    return 0;
}

// The propagate loop shared by propagate_until and propagate_grid: final
// times, max_delta_t (and the grid buffers) are already on the device.
int run_propagation(hyx_batch *b, uint64_t max_steps, int mode, hyx_prop_res *res, bool grid_run, bool init = true)
{
    const uint64_t N = b->N;
    hyx_sys *s = b->sys;
    const bool clog = b->clog_cap > 0;
    if (clog && grid_run) return fail(HYX_EINVAL, "continuous output does not combine with propagate_grid");
    if (clog) {
        if (int rc = ensure_clog_module(s)) return rc;
    }
    hipFunction_t k_init = grid_run ? s->kg_init : clog ? s->kc_init : s->k_init;
    hipFunction_t k_iter = grid_run ? s->kg_iter : clog ? s->kc_iter : s->k_iter;
    hipFunction_t k_persist = grid_run ? s->kg_persist : clog ? s->kc_persist : s->k_persist;
    if (clog && mode == HYX_MODE_GRAPH) mode = HYX_MODE_PLAIN; // the captured graph uses the plain code object
    if (grid_run && mode == HYX_MODE_GRAPH) mode = HYX_MODE_PLAIN; // no captured graph for the grid code object
    if (b->ev_pause && mode == HYX_MODE_GRAPH) mode = HYX_MODE_PLAIN; // the graph holds arguments without parking
    uint64_t pgrid = 0;
    if (mode == HYX_MODE_PERSISTENT) {
        int occ = 0;
        HIPCHK(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_persist, 64, 0));
        const uint64_t per_cu = b->waves_per_cu ? b->waves_per_cu : (uint64_t)std::max(occ, 1);
        pgrid = std::max<uint64_t>(1, std::min<uint64_t>(per_cu * (uint64_t)b->n_cu, blocks_of(b)));
    }
    if (int rc = ensure_tape(b, pgrid * b->sys->inst_per_block)) return rc;
    KArgs a = make_args(b);
    a.max_steps = max_steps;
    if (grid_run) {
        a.grid = b->grid;
        a.gout = b->gout;
        a.gcur = b->gcur;
        a.npts = b->npts;
    }
    if (init) HIPCHK(launch(k_init, blocks_for(N), b->stream, &a));
    b->last_grid = grid_run;
    uint64_t launches = 0;
    bool tail_started = false;
    HIPCHK(hipEventRecord(b->ev0, b->stream));
    if (mode == HYX_MODE_PERSISTENT) {
        // With a tail form (quad or twin), the last instances are handed off to it: the lane
        // kernel writes them back once few remain in flight, and the tail kernel continues the
        // listed ones (and any an iteration budget writes back) until none is left.
        const bool twin = s->kt_persist && b->tail_min > 0 && !grid_run && !clog && !b->ev_pause;
        uint64_t grid = pgrid;
        hipFunction_t kp = k_persist;
        int cur = 0;
        if (twin) {
            HIPCHK(clear32(b, b->inflight, 1));
            a.inflight = b->inflight;
            a.tail_min = b->tail_min;
            a.tailq = b->tailq;
        }
        uint64_t twin_grid = 0;
        if (twin) {
            int occ = 0;
            HIPCHK(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&occ, s->kt_persist, 64, 0));
            twin_grid = (uint64_t)std::max(occ, 1) * (uint64_t)b->n_cu;
        }
        for (;;) {
            HIPCHK(clear64(b, b->work, 1));
            HIPCHK(clear32(b, b->unfinished, 1));
            HIPCHK(launch(kp, grid, b->stream, &a));
            ++launches;
            unsigned unf = 0;
            HIPCHK(hipMemcpyAsync(&unf, b->unfinished, 4, hipMemcpyDeviceToHost, b->stream));
            HIPCHK(hipStreamSynchronize(b->stream));
            if (unf == 0) break;
            if (twin) {
                // Continue the listed instances in the tail form, 16 per wave (twin: 32).
                if (kp != s->kt_persist) {
                    HIPCHK(hipEventRecord(b->evt, b->stream));
                    tail_started = true;
                }
                kp = s->kt_persist;
                a.queue = b->tailq + (uint64_t)cur * N;
                a.N = unf;
                cur ^= 1;
                a.tailq = b->tailq + (uint64_t)cur * N;
                a.inflight = nullptr;
                a.tail_min = 0;
                grid = std::max<uint64_t>(1, std::min<uint64_t>(twin_grid, ((uint64_t)unf * s->tsub + 63) / 64));
                ++b->twin_handoffs;
            }
        }
    } else {
        const uint64_t grid = blocks_of(b);
        if (mode == HYX_MODE_GRAPH && !b->graph_exec) {
            HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            hipError_t e = clear32(b, b->active, kPlainGroup);
            for (int i = 0; i < kPlainGroup && e == hipSuccess; ++i) {
                KArgs ai = a;
                ai.active_count = b->active + i;
                ai.max_steps = 0; // the graph is replayed only when max_steps == 0
                e = launch(k_iter, grid, b->stream, &ai);
            }
            hipGraph_t g = nullptr;
            hipError_t e2 = hipStreamEndCapture(b->stream, &g);
            if (e != hipSuccess || e2 != hipSuccess)
                return fail(HYX_EHIP, std::string("graph capture failed: ") + hipGetErrorString(e != hipSuccess ? e : e2));
            b->graph = g;
            HIPCHK(hipGraphInstantiate(&b->graph_exec, g, nullptr, nullptr, 0));
        }
        for (;;) {
            if (mode == HYX_MODE_GRAPH && max_steps == 0) {
                HIPCHK(hipGraphLaunch(b->graph_exec, b->stream));
            } else {
                HIPCHK(clear32(b, b->active, kPlainGroup));
                for (int i = 0; i < kPlainGroup; ++i) {
                    KArgs ai = a;
                    ai.active_count = b->active + i;
                    HIPCHK(launch(k_iter, grid, b->stream, &ai));
                }
            }
            launches += kPlainGroup;
            unsigned act = 0;
            HIPCHK(hipMemcpyAsync(&act, b->active + (kPlainGroup - 1), 4, hipMemcpyDeviceToHost, b->stream));
            HIPCHK(hipStreamSynchronize(b->stream));
            if (act == 0) break;
        }
    }
    HIPCHK(hipEventRecord(b->ev1, b->stream));
    HIPCHK(hipEventSynchronize(b->ev1));
    float ms = 0.f;
    hipEventElapsedTime(&ms, b->ev0, b->ev1);
    b->last_ms = ms;
    b->last_launches = launches;
    b->last_tail_ms = 0.0;
    if (tail_started) {
        float tms = 0.f;
        hipEventElapsedTime(&tms, b->evt, b->ev1);
        b->last_tail_ms = tms;
    }
    if (res) {
        std::vector<int64_t> oc(N);
        std::vector<double> mn(N), mx(N);
        std::vector<uint64_t> ns(N);
        HIPCHK(hipMemcpyAsync(oc.data(), b->oc, N * 8, hipMemcpyDeviceToHost, b->stream));
        HIPCHK(hipMemcpyAsync(mn.data(), b->minh, N * 8, hipMemcpyDeviceToHost, b->stream));
        HIPCHK(hipMemcpyAsync(mx.data(), b->maxh, N * 8, hipMemcpyDeviceToHost, b->stream));
        HIPCHK(hipMemcpyAsync(ns.data(), b->nsteps, N * 8, hipMemcpyDeviceToHost, b->stream));
        HIPCHK(hipStreamSynchronize(b->stream));
        for (uint64_t i = 0; i < N; ++i) res[i] = hyx_prop_res{oc[i], mn[i], mx[i], ns[i]};
    }
    return 0;
}

// Grows a device buffer to hold at least `count` elements of `elem` bytes.
template <class T>
int reserve(T **p, size_t *cap, size_t count)
{
    if (count <= *cap && *p) return 0;
    if (*p) HIPCHK(hipFree(*p));
    *p = nullptr;
    *cap = 0;
    HIPCHK(hipMalloc((void **)p, count * sizeof(T)));
    *cap = count;
    return 0;
}

} // namespace

extern "C" {

int hyx_propagate_until(hyx_batch *b, const double *tf_hi, const double *tf_lo, const double *max_delta_t,
                        uint64_t max_steps, int mode, hyx_prop_res *res)
{
    if (!b || !tf_hi) return fail(HYX_EINVAL, "null argument");
    if (mode != HYX_MODE_PERSISTENT && mode != HYX_MODE_PLAIN && mode != HYX_MODE_GRAPH)
        return fail(HYX_EINVAL, "unknown propagation mode");
    const uint64_t N = b->N;
    if (int rc = upload(b, b->tfhi, tf_hi, N, 0.0)) return rc;
    if (int rc = upload(b, b->tflo, tf_lo, N, 0.0)) return rc;
    if (int rc = upload(b, b->mdt, max_delta_t, N, std::numeric_limits<double>::infinity())) return rc;
    return run_propagation(b, max_steps, mode, res, false);
}

} // extern "C"

int hyx_internal_fail(int code, const std::string &msg) { return fail(code, msg); }

// The body of hyx_propagate_until_noalloc (hyx_noalloc.hip): the persistent form over
// caller-owned device arrays. Scratch layout inside [scratch, scratch + cap), 256-byte
// aligned pieces: the work counter and the unfinished count, then per instance the
// iteration count, the status word and the last step size, a zero parameter row when the
// system has no parameters, tf_lo / max_delta_t when the caller passes none, and the
// compact form's Taylor tape, whose columns are the persistent grid's slots (the grid is
// narrowed until the tape fits). One launch with an unbounded iteration budget runs every
// instance to its end, so nothing is read back: the kernels are only enqueued on the
// caller's stream, and the host allocates nothing.
int hyx_noalloc_propagate(hyx_sys *s, const hyx::NoAllocArgs &na, unsigned char *scratch, uint64_t cap)
{
    if (!s || !s->mod) return fail(HYX_EINVAL, "the system was not created for a device");
    if (s->has_events) return fail(HYX_EINVAL, "systems with events are not supported without allocation");
    const uint64_t N = na.n_inst;
    if (N == 0 || N > (1ull << 32)) return fail(HYX_EINVAL, "the number of instances must be in [1, 2^32]");
    if (!na.state || !na.t_hi || !na.t_lo || !na.tf_hi || !na.outcome || !na.min_h || !na.max_h || !na.nsteps)
        return fail(HYX_EINVAL, "null device array");
    if (s->S.npar > 0 && !na.pars) return fail(HYX_EINVAL, "the system has parameters: pars must be given");
    int cur = -1;
    HIPCHK(hipGetDevice(&cur));
    if (cur != s->device) return fail(HYX_EINVAL, "the current device is not the one the system was created on");
    hipStream_t st = (hipStream_t)na.stream;
    auto up = [](uint64_t x) { return (x + 255) & ~255ull; };
    uint64_t off = 256;
    auto take = [&](uint64_t bytes) {
        const uint64_t o = off;
        off = up(off + bytes);
        return o;
    };
    const uint64_t o_iters = take(N * 8), o_status = take(N * 4), o_hout = take(N * 8);
    const uint64_t o_pars = na.pars ? 0 : take(N * 8);
    const uint64_t o_tflo = na.tf_lo ? 0 : take(N * 8);
    const uint64_t o_mdt = na.max_dt ? 0 : take(N * 8);
    if (off > cap)
        return fail(HYX_EINVAL, "hyx_propagate_until_noalloc: " + std::to_string(N) + " instances need " +
                                    std::to_string(off) + " bytes of scratch, beyond the fixed " + std::to_string(cap) +
                                    " bytes (HYX_NOALLOC_SCRATCH_BYTES)");
    int occ = 0, n_cu = 0;
    HIPCHK(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&occ, s->k_persist, 64, 0));
    HIPCHK(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, s->device));
    const uint64_t blocks = (N + s->inst_per_block - 1) / s->inst_per_block;
    uint64_t pgrid = std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)std::max(occ, 1) * (uint64_t)n_cu, blocks));
    uint64_t o_tape = 0, stride = 0;
    if (s->compact) {
        const uint64_t rows = (uint64_t)s->S.nnodes * (s->order + 1);
        const uint64_t room = cap > off ? cap - off : 0;
        const uint64_t per_block = rows * s->inst_per_block * 8;
        pgrid = std::min<uint64_t>(pgrid, room / per_block);
        if (pgrid == 0)
            return fail(HYX_EINVAL, "hyx_propagate_until_noalloc: the Taylor tape of one block (" +
                                        std::to_string(per_block) + " bytes) does not fit the scratch left (" +
                                        std::to_string(room) + " bytes)");
        stride = pgrid * s->inst_per_block;
        o_tape = take(rows * stride * 8);
    }
    unsigned char *const sc = scratch;
    KArgs a;
    std::memset(&a, 0, sizeof(a));
    a.state = na.state;
    a.pars = na.pars ? na.pars : (const double *)(sc + o_pars);
    a.t_hi = na.t_hi;
    a.t_lo = na.t_lo;
    a.tf_hi = na.tf_hi;
    a.tf_lo = na.tf_lo ? na.tf_lo : (const double *)(sc + o_tflo);
    a.max_dt = na.max_dt ? na.max_dt : (const double *)(sc + o_mdt);
    a.h_out = (double *)(sc + o_hout);
    a.outcome = na.outcome;
    a.min_h = na.min_h;
    a.max_h = na.max_h;
    a.nsteps = na.nsteps;
    a.iters = (uint64_t *)(sc + o_iters);
    a.status = (unsigned *)(sc + o_status);
    a.tape = s->compact ? (double *)(sc + o_tape) : nullptr;
    a.tape_stride = stride;
    a.work_counter = (uint64_t *)sc;
    a.unfinished = (unsigned *)(sc + 8);
    a.N = N;
    a.ld = N;
    a.max_steps = na.max_steps;
    a.iter_budget = ~0ull;
    a.refill_min = 1;
    auto fill64 = [&](const void *p, double v) -> hipError_t {
        uint64_t bits;
        std::memcpy(&bits, &v, 8);
        void *pp = const_cast<void *>(p);
        uint64_t n = N;
        void *params[] = {&pp, &bits, &n};
        const unsigned g = (unsigned)std::min<uint64_t>((N + 255) / 256, 4096);
        return hipModuleLaunchKernel(s->k_fill64, g, 1, 1, 256, 1, 1, 0, st, params, nullptr);
    };
    // The scratch keeps the previous call's contents: everything read before it is
    // written is set here (the counters, the default arrays; the init kernel sets the
    // per-instance bookkeeping).
    if (!na.pars) HIPCHK(fill64(a.pars, 0.0));
    if (!na.tf_lo) HIPCHK(fill64(a.tf_lo, 0.0));
    if (!na.max_dt) HIPCHK(fill64(a.max_dt, std::numeric_limits<double>::infinity()));
    {
        uint64_t *w = a.work_counter;
        unsigned one = 1;
        void *p64[] = {&w, &one};
        HIPCHK(hipModuleLaunchKernel(s->k_clr64, 1, 1, 1, 64, 1, 1, 0, st, p64, nullptr));
        unsigned *u = a.unfinished;
        void *p32[] = {&u, &one};
        HIPCHK(hipModuleLaunchKernel(s->k_clr32, 1, 1, 1, 64, 1, 1, 0, st, p32, nullptr));
    }
    HIPCHK(launch(s->k_init, blocks_for(N), st, &a));
    HIPCHK(launch(s->k_persist, pgrid, st, &a));
    return 0;
}

extern "C" {

int hyx_propagate_grid(hyx_batch *b, const double *grid, uint64_t n_pts, const double *max_delta_t,
                       uint64_t max_steps, int mode, hyx_prop_res *res, double *out)
{
    if (!b || !grid || !out) return fail(HYX_EINVAL, "null argument");
    if (n_pts == 0) return fail(HYX_EINVAL, "the time grid of propagate_grid() must not be empty");
    if (mode != HYX_MODE_PERSISTENT && mode != HYX_MODE_PLAIN && mode != HYX_MODE_GRAPH)
        return fail(HYX_EINVAL, "unknown propagation mode");
    const uint64_t N = b->N, n = b->sys->S.n, P = b->sys->order;
    // Grid checks, per instance: finite, starting at the current time, strictly monotonic.
    std::vector<double> thi(N);
    HIPCHK(hipMemcpyAsync(thi.data(), b->thi, N * 8, hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    for (uint64_t i = 0; i < n_pts * N; ++i)
        if (!std::isfinite(grid[i])) return fail(HYX_EINVAL, "a non-finite time was passed to propagate_grid()");
    for (uint64_t i = 0; i < N; ++i) {
        if (grid[i] != thi[i])
            return fail(HYX_EINVAL, "the first point of the time grid of instance " + std::to_string(i) +
                                        " does not match its current time in propagate_grid()");
        if (n_pts < 2) continue;
        const bool fwd = grid[(n_pts - 1) * N + i] > grid[i];
        for (uint64_t j = 1; j < n_pts; ++j) {
            const double a = grid[(j - 1) * N + i], c = grid[j * N + i];
            if (fwd ? !(c > a) : !(c < a))
                return fail(HYX_EINVAL, "a non-monotonic time grid was passed to propagate_grid() (instance " +
                                            std::to_string(i) + ")");
        }
    }
    if (int rc = ensure_grid_module(b->sys)) return rc;
    if (!group_kernels(b->sys) && !b->tc) HIPCHK(hipMalloc((void **)&b->tc, (size_t)n * (P + 1) * N * 8));
    if (int rc = reserve(&b->grid, &b->grid_cap, n_pts * N)) return rc;
    if (int rc = reserve(&b->gout, &b->gout_cap, n_pts * n * N)) return rc;
    if (!b->gcur) HIPCHK(hipMalloc((void **)&b->gcur, N * 4));
    HIPCHK(hipMemcpyAsync(b->grid, grid, n_pts * N * 8, hipMemcpyHostToDevice, b->stream));
    {
        // Unreached points stay NaN (a propagation that ends early, as in heyoka's batch propagate_grid).
        uint64_t *p = (uint64_t *)b->gout;
        uint64_t v = 0x7ff8000000000000ull, cnt = n_pts * n * N;
        void *params[] = {&p, &v, &cnt};
        const unsigned blocks = (unsigned)std::min<uint64_t>((cnt + 255) / 256, 8192);
        HIPCHK(hipModuleLaunchKernel(b->sys->k_fill64, blocks, 1, 1, 256, 1, 1, 0, b->stream, params, nullptr));
        unsigned *q = b->gcur;
        unsigned one = 1;
        uint64_t nn = N;
        void *params2[] = {&q, &one, &nn};
        HIPCHK(hipModuleLaunchKernel(b->sys->k_fill32, (unsigned)((N + 63) / 64), 1, 1, 64, 1, 1, 0, b->stream, params2,
                                     nullptr));
    }
    // The first grid point is the current time: its output is the current state.
    HIPCHK(hipMemcpyAsync(b->gout, b->state, n * N * 8, hipMemcpyDeviceToDevice, b->stream));
    HIPCHK(hipMemcpyAsync(b->tfhi, b->grid + (n_pts - 1) * N, N * 8, hipMemcpyDeviceToDevice, b->stream));
    if (int rc = upload(b, b->tflo, nullptr, N, 0.0)) return rc;
    if (int rc = upload(b, b->mdt, max_delta_t, N, std::numeric_limits<double>::infinity())) return rc;
    b->npts = n_pts;
    if (int rc = run_propagation(b, max_steps, mode, res, true)) return rc;
    HIPCHK(hipMemcpyAsync(out, b->gout, n_pts * n * N * 8, hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    return 0;
}

int hyx_autotune(hyx_batch *b, const double *tf_hi, const char *cache_path, hyx_tune_res *best)
{
    if (!b || !tf_hi) return fail(HYX_EINVAL, "null argument");
    const uint64_t N = b->N, n = b->sys->S.n, np = std::max<uint32_t>(b->sys->S.npar, 1);
    const std::string key = std::to_string(fnv1a(b->sys->src)) + "-" + device_arch(b->device) + "-" +
                            std::to_string(N);
    // A cached result for this code object, device and batch size.
    if (cache_path) {
        std::ifstream in(cache_path);
        std::string k;
        unsigned w, r;
        double rate;
        while (in >> k >> w >> r >> rate)
            if (k == key) {
                b->waves_per_cu = w;
                b->refill_min = r;
                if (best) *best = hyx_tune_res{w, r, rate, 1};
                return 0;
            }
    }
    // Snapshot of the batch's contents: every candidate starts from it, and it is restored at the end.
    std::vector<double> snap((size_t)(n + np + 2) * N);
    HIPCHK(hipMemcpyAsync(snap.data(), b->state, n * N * 8, hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipMemcpyAsync(snap.data() + n * N, b->pars, np * N * 8, hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipMemcpyAsync(snap.data() + (n + np) * N, b->thi, N * 8, hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipMemcpyAsync(snap.data() + (n + np + 1) * N, b->tlo, N * 8, hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    auto restore = [&]() -> int {
        HIPCHK(hipMemcpyAsync(b->state, snap.data(), n * N * 8, hipMemcpyHostToDevice, b->stream));
        HIPCHK(hipMemcpyAsync(b->pars, snap.data() + n * N, np * N * 8, hipMemcpyHostToDevice, b->stream));
        HIPCHK(hipMemcpyAsync(b->thi, snap.data() + (n + np) * N, N * 8, hipMemcpyHostToDevice, b->stream));
        HIPCHK(hipMemcpyAsync(b->tlo, snap.data() + (n + np + 1) * N, N * 8, hipMemcpyHostToDevice, b->stream));
        HIPCHK(hipStreamSynchronize(b->stream));
        return 0;
    };
    const uint32_t waves[] = {0, 1, 2, 4, 8};
    const uint32_t refills[] = {1, 8, 32};
    hyx_tune_res top{0, 1, -1.0, 0};
    std::vector<hyx_prop_res> res(N);
    for (uint32_t w : waves)
        for (uint32_t r : refills) {
            if (int rc = restore()) return rc;
            b->waves_per_cu = w;
            b->refill_min = r;
            if (int rc = upload(b, b->tfhi, tf_hi, N, 0.0)) return rc;
            if (int rc = upload(b, b->tflo, nullptr, N, 0.0)) return rc;
            if (int rc = upload(b, b->mdt, nullptr, N, std::numeric_limits<double>::infinity())) return rc;
            if (int rc = run_propagation(b, 0, HYX_MODE_PERSISTENT, res.data(), false)) return rc;
            double steps = 0.0;
            for (auto &x : res) steps += (double)x.nsteps;
            const double rate = b->last_ms > 0.0 ? steps / (b->last_ms * 1e-3) : 0.0;
            if (rate > top.rate) top = hyx_tune_res{w, r, rate, 0};
        }
    if (int rc = restore()) return rc;
    b->waves_per_cu = top.waves_per_cu;
    b->refill_min = top.refill_min;
    if (cache_path) {
        std::ofstream out(cache_path, std::ios::app);
        out << key << " " << top.waves_per_cu << " " << top.refill_min << " " << top.rate << "\n";
    }
    if (best) *best = top;
    return 0;
}

int hyx_batch_last_timing(const hyx_batch *b, double *kernel_ms, uint64_t *launches)
{
    if (!b) return fail(HYX_EINVAL, "null batch");
    if (kernel_ms) *kernel_ms = b->last_ms;
    if (launches) *launches = b->last_launches;
    return 0;
}

int hyx_batch_last_tail_ms(const hyx_batch *b, double *tail_ms)
{
    if (!b || !tail_ms) return fail(HYX_EINVAL, "null argument");
    *tail_ms = b->last_tail_ms;
    return 0;
}

int hyx_batch_set_persistent_config(hyx_batch *b, uint32_t waves_per_cu, uint32_t refill_min)
{
    if (!b) return fail(HYX_EINVAL, "null batch");
    if (refill_min < 1 || refill_min > 64) return fail(HYX_EINVAL, "refill_min must be in [1, 64]");
    if (waves_per_cu > 64) return fail(HYX_EINVAL, "waves_per_cu must be in [0, 64]");
    b->waves_per_cu = waves_per_cu;
    b->refill_min = refill_min;
    return 0;
}

int hyx_batch_set_iter_budget(hyx_batch *b, uint64_t iters)
{
    if (!b) return fail(HYX_EINVAL, "null batch");
    if (iters == 0) return fail(HYX_EINVAL, "the iteration budget must be positive");
    b->iter_budget = iters;
    return 0;
}

int hyx_batch_set_queue(hyx_batch *b, const uint32_t *order)
{
    if (!b) return fail(HYX_EINVAL, "null batch");
    if (!order) {
        if (b->queue) hipFree(b->queue);
        b->queue = nullptr;
        return 0;
    }
    std::vector<char> seen(b->N, 0);
    for (uint64_t i = 0; i < b->N; ++i) {
        if (order[i] >= b->N || seen[order[i]]) return fail(HYX_EINVAL, "the queue must be a permutation of 0..N-1");
        seen[order[i]] = 1;
    }
    if (!b->queue) HIPCHK(hipMalloc((void **)&b->queue, b->N * 4));
    HIPCHK(hipMemcpyAsync(b->queue, order, b->N * 4, hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    return 0;
}


// ---- serving: slots ------------------------------------------------------------------------

static int stage_reserve(hyx_batch *b, size_t bytes, size_t nidx)
{
    if (bytes > b->stage_bytes) {
        if (b->stage) HIPCHK(hipFree(b->stage));
        b->stage = nullptr;
        b->stage_bytes = 0;
        HIPCHK(hipMalloc(&b->stage, bytes));
        b->stage_bytes = bytes;
    }
    if (nidx > b->stage_idx_n) {
        if (b->stage_idx) HIPCHK(hipFree(b->stage_idx));
        b->stage_idx = nullptr;
        b->stage_idx_n = 0;
        HIPCHK(hipMalloc((void **)&b->stage_idx, nidx * 4));
        b->stage_idx_n = nidx;
    }
    return 0;
}

static int check_slots(hyx_batch *b, const uint32_t *slots, uint32_t count)
{
    if (!slots && count) return fail(HYX_EINVAL, "null slot list");
    for (uint32_t c = 0; c < count; ++c)
        if (slots[c] >= b->N) return fail(HYX_EINVAL, "slot index out of range");
    if (b->slots_running) return fail(HYX_EINVAL, "a slot epoch is running: call hyx_slots_wait first");
    return 0;
}

static hipError_t launch_raw(hyx_batch *b, hipFunction_t f, uint64_t threads, void **params)
{
    return hipModuleLaunchKernel(f, (unsigned)((threads + 63) / 64), 1, 1, 64, 1, 1, 0, b->stream, params, nullptr);
}

int hyx_slots_reset(hyx_batch *b)
{
    if (!b) return fail(HYX_EINVAL, "null batch");
    if (b->slots_running) return fail(HYX_EINVAL, "a slot epoch is running: call hyx_slots_wait first");
    unsigned *p = b->status;
    unsigned v = 1u;
    uint64_t n = b->N;
    void *params[] = {&p, &v, &n};
    HIPCHK(launch_raw(b, b->sys->k_fill32, n, params));
    HIPCHK(hipStreamSynchronize(b->stream));
    return 0;
}

int hyx_slots_write(hyx_batch *b, int field, const uint32_t *slots, uint32_t count, const double *src)
{
    if (!b || (!src && count)) return fail(HYX_EINVAL, "null argument");
    if (int rc = check_slots(b, slots, count)) return rc;
    if (count == 0) return 0;
    double *p;
    uint64_t total;
    if (int rc = field_info(b, field, &p, &total)) return rc;
    const unsigned rows = (unsigned)(total / b->N);
    if (rows == 0) return 0;
    if (int rc = stage_reserve(b, (size_t)rows * count * 8, count)) return rc;
    HIPCHK(hipMemcpyAsync(b->stage_idx, slots, (size_t)count * 4, hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipMemcpyAsync(b->stage, src, (size_t)rows * count * 8, hipMemcpyHostToDevice, b->stream));
    uint64_t ld = b->N;
    void *dst = p;
    void *params[] = {&dst, &ld, (void *)&rows, &b->stage, &b->stage_idx, &count};
    HIPCHK(launch_raw(b, b->sys->k_scatter, (uint64_t)rows * count, params));
    HIPCHK(hipStreamSynchronize(b->stream));
    return 0;
}

static int gather64(hyx_batch *b, const void *srcm, unsigned rows, const uint32_t *slots, uint32_t count, void *dst)
{
    if (int rc = stage_reserve(b, (size_t)rows * count * 8, count)) return rc;
    HIPCHK(hipMemcpyAsync(b->stage_idx, slots, (size_t)count * 4, hipMemcpyHostToDevice, b->stream));
    uint64_t ld = b->N;
    const void *sm = srcm;
    void *params[] = {(void *)&sm, &ld, &rows, &b->stage, &b->stage_idx, &count};
    HIPCHK(launch_raw(b, b->sys->k_gather, (uint64_t)rows * count, params));
    HIPCHK(hipMemcpyAsync(dst, b->stage, (size_t)rows * count * 8, hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    return 0;
}

int hyx_slots_read(hyx_batch *b, int field, const uint32_t *slots, uint32_t count, double *dst)
{
    if (!b || (!dst && count)) return fail(HYX_EINVAL, "null argument");
    if (int rc = check_slots(b, slots, count)) return rc;
    if (count == 0) return 0;
    double *p;
    uint64_t total;
    if (int rc = field_info(b, field, &p, &total)) return rc;
    const unsigned rows = (unsigned)(total / b->N);
    if (rows == 0) return 0;
    return gather64(b, p, rows, slots, count, dst);
}

// Per-slot columns of a snapshot, in blob order: state, pars, then the propagation
// bookkeeping (time, final time, max_delta_t, min_h, max_h, nsteps, iterations,
// outcome), then the 32-bit status widened to one word.
static std::vector<std::pair<void *, unsigned>> snapshot_cols(hyx_batch *b)
{
    return {{b->state, b->sys->S.n}, {b->pars, b->sys->S.npar}, {b->thi, 1},  {b->tlo, 1},
            {b->tfhi, 1},            {b->tflo, 1},              {b->mdt, 1},  {b->minh, 1},
            {b->maxh, 1},            {b->nsteps, 1},            {b->iters, 1}, {b->oc, 1}};
}

uint64_t hyx_slots_snapshot_words(const hyx_batch *b)
{
    return b ? (uint64_t)b->sys->S.n + b->sys->S.npar + 11 : 0;
}

int hyx_slots_snapshot(hyx_batch *b, const uint32_t *slots, uint32_t count, uint64_t *dst)
{
    if (!b || (!dst && count)) return fail(HYX_EINVAL, "null argument");
    if (int rc = check_slots(b, slots, count)) return rc;
    if (count == 0) return 0;
    uint64_t *o = dst;
    for (auto &c : snapshot_cols(b)) {
        if (c.second == 0) continue;
        if (int rc = gather64(b, c.first, c.second, slots, count, o)) return rc;
        o += (size_t)c.second * count;
    }
    // status: 32-bit column, gathered then widened.
    if (int rc = stage_reserve(b, (size_t)count * 8, count)) return rc;
    HIPCHK(hipMemcpyAsync(b->stage_idx, slots, (size_t)count * 4, hipMemcpyHostToDevice, b->stream));
    unsigned *st = b->status, *sg = (unsigned *)b->stage, *idx = b->stage_idx;
    void *params[] = {&st, &sg, &idx, &count};
    HIPCHK(launch_raw(b, b->sys->k_gather32, count, params));
    std::vector<unsigned> s32(count);
    HIPCHK(hipMemcpyAsync(s32.data(), b->stage, (size_t)count * 4, hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    for (uint32_t c = 0; c < count; ++c) o[c] = s32[c];
    return 0;
}

int hyx_slots_restore(hyx_batch *b, const uint32_t *slots, uint32_t count, const uint64_t *src)
{
    if (!b || (!src && count)) return fail(HYX_EINVAL, "null argument");
    if (int rc = check_slots(b, slots, count)) return rc;
    if (count == 0) return 0;
    const uint64_t *in = src;
    const uint64_t words = hyx_slots_snapshot_words(b);
    if (int rc = stage_reserve(b, (size_t)words * count * 8, count)) return rc;
    HIPCHK(hipMemcpyAsync(b->stage_idx, slots, (size_t)count * 4, hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipMemcpyAsync(b->stage, src, (size_t)words * count * 8, hipMemcpyHostToDevice, b->stream));
    uint64_t off = 0;
    for (auto &c : snapshot_cols(b)) {
        if (c.second == 0) continue;
        void *dstp = c.first;
        uint64_t ld = b->N;
        unsigned rows = c.second;
        void *srcp = (char *)b->stage + off * 8;
        void *params[] = {&dstp, &ld, &rows, &srcp, &b->stage_idx, &count};
        HIPCHK(launch_raw(b, b->sys->k_scatter, (uint64_t)rows * count, params));
        off += (uint64_t)rows * count;
    }
    std::vector<unsigned> s32(count);
    for (uint32_t c = 0; c < count; ++c) s32[c] = (unsigned)in[off + c];
    // The 32-bit status goes through a second staging area after the 64-bit columns.
    unsigned *sg = (unsigned *)((char *)b->stage + off * 8);
    HIPCHK(hipMemcpyAsync(sg, s32.data(), (size_t)count * 4, hipMemcpyHostToDevice, b->stream));
    unsigned *st = b->status, *idx = b->stage_idx;
    void *params[] = {&st, &sg, &idx, &count};
    HIPCHK(launch_raw(b, b->sys->k_scatter32, count, params));
    HIPCHK(hipStreamSynchronize(b->stream));
    return 0;
}

int hyx_slots_arm(hyx_batch *b, const uint32_t *slots, uint32_t count, const double *tf_hi, const double *tf_lo,
                  const double *max_delta_t)
{
    if (!b || (!tf_hi && count)) return fail(HYX_EINVAL, "null argument");
    if (int rc = check_slots(b, slots, count)) return rc;
    if (count == 0) return 0;
    std::vector<double> st((size_t)3 * count);
    for (uint32_t c = 0; c < count; ++c) {
        st[c] = tf_hi[c];
        st[count + c] = tf_lo ? tf_lo[c] : 0.0;
        st[2 * (size_t)count + c] = max_delta_t ? max_delta_t[c] : std::numeric_limits<double>::infinity();
        if (!(st[2 * (size_t)count + c] > 0.0)) return fail(HYX_EINVAL, "max_delta_t must be positive");
    }
    if (int rc = stage_reserve(b, st.size() * 8, count)) return rc;
    HIPCHK(hipMemcpyAsync(b->stage_idx, slots, (size_t)count * 4, hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipMemcpyAsync(b->stage, st.data(), st.size() * 8, hipMemcpyHostToDevice, b->stream));
    KArgs a = make_args(b);
    void *params[] = {&a, &b->stage_idx, &count, &b->stage};
    HIPCHK(launch_raw(b, b->sys->k_arm, count, params));
    HIPCHK(hipStreamSynchronize(b->stream));
    return 0;
}

int hyx_slots_launch(hyx_batch *b, uint64_t iter_budget)
{
    if (!b) return fail(HYX_EINVAL, "null batch");
    if (b->slots_running) return fail(HYX_EINVAL, "a slot epoch is already running");
    if (iter_budget == 0) return fail(HYX_EINVAL, "the iteration budget must be positive");
    int occ = 0;
    HIPCHK(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&occ, b->sys->k_persist, 64, 0));
    const uint64_t per_cu = b->waves_per_cu ? b->waves_per_cu : (uint64_t)std::max(occ, 1);
    const uint64_t grid = std::max<uint64_t>(1, std::min<uint64_t>(per_cu * (uint64_t)b->n_cu, blocks_of(b)));
    if (int rc = ensure_tape(b, grid * b->sys->inst_per_block)) return rc;
    KArgs a = make_args(b);
    a.iter_budget = iter_budget;
    HIPCHK(clear64(b, b->work, 1));
    HIPCHK(clear32(b, b->unfinished, 1));
    HIPCHK(hipEventRecord(b->ev0, b->stream));
    HIPCHK(launch(b->sys->k_persist, grid, b->stream, &a));
    HIPCHK(hipEventRecord(b->ev1, b->stream));
    b->slots_running = true;
    return 0;
}

int hyx_slots_wait(hyx_batch *b, uint32_t *n_active)
{
    if (!b) return fail(HYX_EINVAL, "null batch");
    if (!b->slots_running) {
        if (n_active) *n_active = 0;
        return 0;
    }
    unsigned unf = 0;
    hipError_t e = hipMemcpyAsync(&unf, b->unfinished, 4, hipMemcpyDeviceToHost, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    b->slots_running = false;
    if (e != hipSuccess) return fail(HYX_EHIP, std::string("hyx_slots_wait: ") + hipGetErrorString(e));
    float ms = 0.f;
    hipEventElapsedTime(&ms, b->ev0, b->ev1);
    b->last_ms = ms;
    b->last_launches = 1;
    if (n_active) *n_active = unf;
    return 0;
}

int hyx_slots_poll(hyx_batch *b, int *finished)
{
    if (!b || !finished) return fail(HYX_EINVAL, "null argument");
    if (!b->slots_running) {
        *finished = 1;
        return 0;
    }
    const hipError_t e = hipEventQuery(b->ev1);
    if (e == hipSuccess) {
        *finished = 1;
        return 0;
    }
    if (e == hipErrorNotReady) {
        *finished = 0;
        return 0;
    }
    return fail(HYX_EHIP, std::string("hyx_slots_poll: ") + hipGetErrorString(e));
}

int hyx_slots_results(hyx_batch *b, const uint32_t *slots, uint32_t count, uint32_t *done, hyx_prop_res *res)
{
    if (!b) return fail(HYX_EINVAL, "null batch");
    if (int rc = check_slots(b, slots, count)) return rc;
    if (count == 0) return 0;
    if (done) {
        if (int rc = stage_reserve(b, (size_t)count * 8, count)) return rc;
        HIPCHK(hipMemcpyAsync(b->stage_idx, slots, (size_t)count * 4, hipMemcpyHostToDevice, b->stream));
        const unsigned *srcm = b->status;
        void *params[] = {(void *)&srcm, &b->stage, &b->stage_idx, &count};
        HIPCHK(launch_raw(b, b->sys->k_gather32, count, params));
        HIPCHK(hipMemcpyAsync(done, b->stage, (size_t)count * 4, hipMemcpyDeviceToHost, b->stream));
        HIPCHK(hipStreamSynchronize(b->stream));
        for (uint32_t c = 0; c < count; ++c) done[c] = (done[c] == 1u) ? 1u : 0u;
    }
    if (res) {
        std::vector<uint64_t> tmp((size_t)count);
        const void *cols[4] = {b->oc, b->minh, b->maxh, b->nsteps};
        std::vector<uint64_t> all((size_t)4 * count);
        for (int q = 0; q < 4; ++q) {
            if (int rc = gather64(b, cols[q], 1, slots, count, tmp.data())) return rc;
            std::memcpy(all.data() + (size_t)q * count, tmp.data(), (size_t)count * 8);
        }
        for (uint32_t c = 0; c < count; ++c) {
            std::memcpy(&res[c].outcome, &all[c], 8);
            std::memcpy(&res[c].min_h, &all[count + c], 8);
            std::memcpy(&res[c].max_h, &all[2 * (size_t)count + c], 8);
            std::memcpy(&res[c].nsteps, &all[3 * (size_t)count + c], 8);
        }
    }
    return 0;
}

int hyx_batch_sync(hyx_batch *b)
{
    if (!b) return fail(HYX_EINVAL, "null batch");
    HIPCHK(hipStreamSynchronize(b->stream));
    return 0;
}

// ---- events: parking, gathering what the callbacks see, resuming --------------------------

int hyx_batch_set_event_pause(hyx_batch *b, int on)
{
    if (!b) return fail(HYX_EINVAL, "null batch");
    if (on && !b->sys->has_events) return fail(HYX_EINVAL, "the system has no events");
    b->ev_pause = on != 0;
    return 0;
}

static int read_status(hyx_batch *b, std::vector<unsigned> &st)
{
    st.resize(b->N);
    HIPCHK(hipMemcpyAsync(st.data(), b->status, b->N * 4, hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    return 0;
}

int hyx_events_parked(hyx_batch *b, uint32_t *lanes, uint32_t *count)
{
    if (!b || !count) return fail(HYX_EINVAL, "null argument");
    std::vector<unsigned> st;
    if (int rc = read_status(b, st)) return rc;
    uint32_t c = 0;
    for (uint64_t i = 0; i < b->N; ++i)
        if (st[i] & 4u) {
            if (lanes) lanes[c] = (uint32_t)i;
            ++c;
        }
    *count = c;
    return 0;
}

int hyx_events_gather(hyx_batch *b, const uint32_t *lanes, uint32_t count, int64_t *te, uint32_t *nt_count,
                      double *nt_time, int64_t *nt_code, double *h, double *tc)
{
    if (!b) return fail(HYX_EINVAL, "null batch");
    if (!b->sys->has_events) return fail(HYX_EINVAL, "the system has no events");
    if (int rc = check_slots(b, lanes, count)) return rc;
    if (count == 0) return 0;
    if (te) {
        if (int rc = gather64(b, b->te, 1, lanes, count, te)) return rc;
    }
    if (nt_count) {
        if (int rc = stage_reserve(b, (size_t)count * 8, count)) return rc;
        HIPCHK(hipMemcpyAsync(b->stage_idx, lanes, (size_t)count * 4, hipMemcpyHostToDevice, b->stream));
        unsigned *src = b->ntc, *dst = (unsigned *)b->stage, *idx = b->stage_idx;
        void *params[] = {&src, &dst, &idx, &count};
        HIPCHK(launch_raw(b, b->sys->k_gather32, count, params));
        HIPCHK(hipMemcpyAsync(nt_count, b->stage, (size_t)count * 4, hipMemcpyDeviceToHost, b->stream));
        HIPCHK(hipStreamSynchronize(b->stream));
    }
    if (nt_time) {
        if (int rc = gather64(b, b->nt_time, b->nt_cap, lanes, count, nt_time)) return rc;
    }
    if (nt_code) {
        if (int rc = gather64(b, b->nt_code, b->nt_cap, lanes, count, nt_code)) return rc;
    }
    if (h) {
        if (int rc = gather64(b, b->hout, 1, lanes, count, h)) return rc;
    }
    if (tc) {
        const unsigned rows = b->sys->S.n * (b->sys->order + 1);
        if (int rc = gather64(b, b->tc, rows, lanes, count, tc)) return rc;
    }
    return 0;
}

int hyx_events_resume(hyx_batch *b, const uint32_t *lanes, uint32_t count, const int64_t *stop)
{
    if (!b) return fail(HYX_EINVAL, "null batch");
    if (int rc = check_slots(b, lanes, count)) return rc;
    if (count == 0) return 0;
    std::vector<unsigned> st;
    if (int rc = read_status(b, st)) return rc;
    std::vector<uint32_t> sl;
    std::vector<int64_t> oc;
    for (uint32_t c = 0; c < count; ++c) {
        const uint32_t i = lanes[c];
        if (!(st[i] & 4u)) return fail(HYX_EINVAL, "instance " + std::to_string(i) + " is not parked");
        if (stop && stop[c] != 0) {
            st[i] = 1u;
            sl.push_back(i);
            oc.push_back(stop[c]);
        } else {
            st[i] &= 3u;
        }
    }
    HIPCHK(hipMemcpyAsync(b->status, st.data(), b->N * 4, hipMemcpyHostToDevice, b->stream));
    if (!sl.empty()) {
        const uint32_t n = (uint32_t)sl.size();
        if (int rc = stage_reserve(b, (size_t)n * 8, n)) return rc;
        HIPCHK(hipMemcpyAsync(b->stage_idx, sl.data(), (size_t)n * 4, hipMemcpyHostToDevice, b->stream));
        HIPCHK(hipMemcpyAsync(b->stage, oc.data(), (size_t)n * 8, hipMemcpyHostToDevice, b->stream));
        uint64_t ld = b->N;
        unsigned rows = 1;
        void *dst = b->oc;
        void *params[] = {&dst, &ld, &rows, &b->stage, &b->stage_idx, (void *)&n};
        HIPCHK(launch_raw(b, b->sys->k_scatter, n, params));
    }
    HIPCHK(hipStreamSynchronize(b->stream));
    return 0;
}

int hyx_propagate_resume(hyx_batch *b, uint64_t max_steps, int mode, hyx_prop_res *res, double *grid_out)
{
    if (!b) return fail(HYX_EINVAL, "null batch");
    if (mode != HYX_MODE_PERSISTENT && mode != HYX_MODE_PLAIN && mode != HYX_MODE_GRAPH)
        return fail(HYX_EINVAL, "unknown propagation mode");
    if (grid_out && !b->last_grid) return fail(HYX_EINVAL, "the propagation being resumed is not a propagate_grid");
    if (b->slots_running) return fail(HYX_EINVAL, "a slot epoch is running: call hyx_slots_wait first");
    const bool grid = b->last_grid;
    if (int rc = run_propagation(b, max_steps, mode, res, grid, false)) return rc;
    if (grid && grid_out) {
        HIPCHK(hipMemcpyAsync(grid_out, b->gout, b->npts * b->sys->S.n * b->N * 8, hipMemcpyDeviceToHost, b->stream));
        HIPCHK(hipStreamSynchronize(b->stream));
    }
    return 0;
}

// ---- continuous output: per-instance step logs ------------------------------------------

int hyx_batch_set_clog(hyx_batch *b, uint32_t cap)
{
    if (!b) return fail(HYX_EINVAL, "null batch");
    if (cap && b->sys->has_events) return fail(HYX_EINVAL, "the step log does not combine with events");
    if (cap && group_kernels(b->sys)) return fail(HYX_EINVAL, "the group and wave forms have no step log");
    void *ptrs[] = {b->clog, b->clog_t, b->clog_n};
    for (void *p : ptrs)
        if (p) HIPCHK(hipFree(p));
    b->clog = b->clog_t = nullptr;
    b->clog_n = nullptr;
    b->clog_cap = 0;
    if (cap == 0) return 0;
    const uint64_t N = b->N, rows = (uint64_t)b->sys->S.n * (b->sys->order + 1);
    if (hipMalloc((void **)&b->clog, (size_t)cap * rows * N * 8) != hipSuccess ||
        hipMalloc((void **)&b->clog_t, (size_t)cap * 3 * N * 8) != hipSuccess ||
        hipMalloc((void **)&b->clog_n, (size_t)N * 4) != hipSuccess) {
        hyx_batch_set_clog(b, 0);
        return fail(HYX_ENOMEM, "hipMalloc of the step log failed");
    }
    HIPCHK(hipMemsetAsync(b->clog_n, 0, N * 4, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    b->clog_cap = cap;
    return 0;
}

int hyx_clog_drain(hyx_batch *b, const uint32_t *lanes, uint32_t count, uint32_t *nrec, double *tc, double *tt)
{
    if (!b) return fail(HYX_EINVAL, "null batch");
    if (!b->clog_cap) return fail(HYX_EINVAL, "no step log: call hyx_batch_set_clog first");
    if (int rc = check_slots(b, lanes, count)) return rc;
    if (count == 0) return 0;
    const unsigned rows = b->sys->S.n * (b->sys->order + 1);
    if (nrec) {
        if (int rc = stage_reserve(b, (size_t)count * 8, count)) return rc;
        HIPCHK(hipMemcpyAsync(b->stage_idx, lanes, (size_t)count * 4, hipMemcpyHostToDevice, b->stream));
        unsigned *src = b->clog_n, *dst = (unsigned *)b->stage, *idx = b->stage_idx;
        void *params[] = {&src, &dst, &idx, &count};
        HIPCHK(launch_raw(b, b->sys->k_gather32, count, params));
        HIPCHK(hipMemcpyAsync(nrec, b->stage, (size_t)count * 4, hipMemcpyDeviceToHost, b->stream));
        HIPCHK(hipStreamSynchronize(b->stream));
    }
    if (tc) {
        if (int rc = gather64(b, b->clog, b->clog_cap * rows, lanes, count, tc)) return rc;
    }
    if (tt) {
        if (int rc = gather64(b, b->clog_t, b->clog_cap * 3, lanes, count, tt)) return rc;
    }
    // The drained logs start over.
    if (int rc = stage_reserve(b, (size_t)count * 8, count)) return rc;
    HIPCHK(hipMemcpyAsync(b->stage_idx, lanes, (size_t)count * 4, hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipMemsetAsync(b->stage, 0, (size_t)count * 4, b->stream));
    unsigned *dst = b->clog_n, *src = (unsigned *)b->stage, *idx = b->stage_idx;
    void *params[] = {&dst, &src, &idx, &count};
    HIPCHK(launch_raw(b, b->sys->k_scatter32, count, params));
    HIPCHK(hipStreamSynchronize(b->stream));
    return 0;
}

} // extern "C"
