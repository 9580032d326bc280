// Internal declarations shared by the runtime's translation units.
#ifndef HYX_INTERNAL_H
#define HYX_INTERNAL_H

#include <cstdint>
#include <string>
#include <vector>

namespace hyx
{

// propagate_until over caller-owned device arrays with every scratch buffer carved out of
// `scratch` (cap bytes): the body of hyx_propagate_until_noalloc (hyx_noalloc.hip), which
// owns the fixed device array. Enqueues its kernels on `stream` and allocates nothing.
struct NoAllocArgs {
    uint64_t n_inst;
    double *state;
    const double *pars;
    double *t_hi, *t_lo;
    const double *tf_hi, *tf_lo, *max_dt;
    uint64_t max_steps;
    int64_t *outcome;
    double *min_h, *max_h;
    uint64_t *nsteps;
    void *stream;
};

// Internal opcodes of the twin form (never in a user op-list): a per-lane degree-0 binding
// (lane A: operand 0, lane B: operand 1; aux bit 0 / 1 negates it) and a partner exchange
// of a twin slot (aux 0: lane A's value on both lanes, 1: lane B's).
enum { HYX_OP_TWIN_BIND = 100, HYX_OP_TWIN_XCH = 101 };
// Internal opcodes of the quad form: a per-component degree-0 binding (x lanes: operand 0,
// y lanes: operand 1) and the partner component's value of a node (lanes x <-> y).
enum { HYX_OP_QUAD_BIND = 102, HYX_OP_QUAD_SWAP = 103 };

// A parsed op-list (include/hyx_oplist.h).
struct Sys {
    uint32_t n = 0, npar = 0, nnodes = 0, nconsts = 0;
    std::vector<uint32_t> op, aux;               // per node (state variables: 0)
    std::vector<std::vector<uint32_t>> args;     // per node, packed operands
    std::vector<uint32_t> rhs;                   // per state variable
    std::vector<double> consts;
    std::vector<int> deg;                        // highest possibly non-zero order, per node
};

void compute_degrees(Sys &S);

int parse_oplist(const uint32_t *w, uint32_t n_words, const double *consts, uint32_t n_consts, Sys &S,
                 std::string &err);

// HIP source of hyx_taylor_step<WTC>() for S at the given order, unrolled or compact.
// split_min: fma chains of at least this many terms use two accumulators (0: never).
// sub: lanes per instance of the compact form (1, 2, 4 or 8).
// With events (E: the event op-list, nev events) the step takes an extra
// `unsigned long long &hyx_evm`: per event, 1 when its root scan has to run this step.
// lds_rows (unrolled form): 1 = state rows in LDS, 0 = in registers, -1 = chosen from the
// register budget.
std::string generate_prelude(const Sys &S, int order, bool high_acc, bool compact, int split_min = 0, int sub = 1,
                             const Sys *E = nullptr, uint32_t nev = 0, int lds_rows = -1);
// Twin form (hyx_codegen.cpp): two lanes per instance for systems whose right-hand sides
// add isomorphic subexpressions; the same arithmetic as the unrolled form, bit for bit.
// has_twin: whether the system admits it (the unrolled form then compiles without implicit
// fusion); twin_oplist: its op-list (internal opcodes); generate_twin_prelude: its step
// source ("" with the reason in why when it does not apply).
bool has_twin(const Sys &S);
Sys twin_oplist(const Sys &S, std::string &why);
std::string generate_twin_prelude(const Sys &S, int order, bool high_acc, int split_min, std::string &why);
// Quad form (hyx_codegen.cpp): four lanes per instance for twin systems whose twin slots
// split once more under a pairing of the state variables (the planar CR3BP: x <-> y,
// vx <-> vy); the same arithmetic as the unrolled form, bit for bit. "" (reason in why)
// when it does not apply.
std::string generate_quad_prelude(const Sys &S, int order, bool high_acc, int split_min, std::string &why);
// Number of nodes whose coefficient history the recurrences read.
int history_nodes(const Sys &S);
// Default choice of code form: compact when the histories exceed the register file.
bool choose_compact(const Sys &S, int order);

// Group form: a template cluster evaluated by each lane of a group of G lanes
// per instance, with per-lane bindings (see plan_groups in hyx_codegen.cpp).
struct GroupPlan {
    int G = 0, ncl = 0, nin = 0, nout = 0, nkc = 1, own = 0;
    int os = 0;                     // LDS row stride of the cluster outputs (G + 1)
    Sys tmpl;                       // cluster template: inputs are its state variables 0..nin-1
    std::vector<char> perlane;      // template constant held per lane
    std::vector<uint32_t> outs;     // template nodes feeding right-hand sides, by output slot
    std::vector<int> iv;            // [G][nin] state variable of each input
    std::vector<double> kc;         // [G][nkc] constants
    // Owned state variables: lane g, slot o owns vo[g][o] (-1: none). Its right-hand
    // side is rc + sum of output terms (LDS output index, sign) + sum of state terms
    // (variable index, sign); per slot, nto[o] / nts[o] terms (padded: output index
    // os * nout and variable index n read zeros).
    std::vector<int> vo;            // [G][own]
    std::vector<double> rc;         // [G][own]
    std::vector<int> nto, nts;      // [own]
    std::vector<std::vector<int>> tio, tis;          // [own][G * nto[o]] / [own][G * nts[o]]
    std::vector<std::vector<double>> tso, tss;       // signs, same shapes
};

bool plan_groups(const Sys &S, GroupPlan &gp, std::string &why);
// pipeline: one LDS exchange per order when every cluster input is a lagged state variable.
// With events (E, nev) the step takes `unsigned long long &hyx_evm, double &hyx_mx`
// (per-event root-scan flags, and the group's max |x| at the start of the step).
std::string generate_group_prelude(const Sys &S, const GroupPlan &gp, int order, bool high_acc, int split_min = 0,
                                   bool pipeline = true, const Sys *E = nullptr, uint32_t nev = 0);

// Wave form: G lanes per instance (16, 32 or 64) over a static schedule of the nodes,
// every coefficient row in LDS; the same hyx_gstep contract as the group form, so the
// group kernels run it. Returns "" with `why` set when the form does not apply.
std::string generate_wave_prelude(const Sys &S, int order, bool high_acc, int G, const Sys *E, uint32_t nev,
                                  int split_min, std::string &why);

// Compiled function: the order-0 values of the first n_out right-hand sides (hyx_cfunc_eval).
std::string generate_cfunc_prelude(const Sys &S, uint32_t n_out);

// Event detection: directions (-1, 0, 1) of every event (terminal ones first),
// the terminal events' cooldowns (< 0: automatic) and the integrator tolerance.
struct EventSpec {
    uint32_t n_te = 0;
    std::vector<int> dir;
    std::vector<double> cooldown;
    std::vector<int> stopping; // terminal events without a callback: always stop the instance
    double tol = 0.0;
};

// Event prelude: hyx_event_jet() (the event op-list E's nodes at orders 0..P from the
// step's state coefficients) and hyx_ev_coeff(); E's rows 0..n_sys-1 are the state
// variables and its right-hand sides 0..nev-1 the event functions.
std::string generate_events_prelude(const Sys &E, uint32_t n_sys, int P, const EventSpec &ev, bool high_acc);
// Placed before the step prelude of a system with events: HYX_EVENTS and the declaration
// of hyx_ev_norms(), which the step calls to fold the event functions into its norms.
std::string events_forward_decl();

// Exact hexadecimal double literal.
std::string lit(double v);

} // namespace hyx

struct hyx_sys;
// Enqueues the propagation of NoAllocArgs on its stream with every scratch buffer inside
// [scratch, scratch + cap) (a device address); returns 0 or a negative HYX_E* code.
int hyx_noalloc_propagate(hyx_sys *s, const hyx::NoAllocArgs &na, unsigned char *scratch, uint64_t cap);
// Sets the thread's hyx_last_error() message and returns code.
int hyx_internal_fail(int code, const std::string &msg);

#endif
