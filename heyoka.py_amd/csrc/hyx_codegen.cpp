// Per-system code generation: op-list -> HIP source of hyx_taylor_step<WTC>().
//
// The generated function runs, for one instance held by the calling lane, the
// normalised-derivative recurrences of every node of the Taylor decomposition
// (doc/tut_taylor_method.rst:22-81; Jorba & Zou 2005), picks the step size
// from the infinity norms of the last two orders with heyoka's absolute /
// relative switch (doc/notebooks/Customising the adaptive integrator.ipynb:23-31)
// and advances the state by Horner's rule (compensated summation in
// high-accuracy mode).
//
// Unrolled form: every coefficient is a scalar SSA value and every loop is
// unrolled, so the register allocator keeps exactly the histories that later
// orders read. Operands whose series is known to stop early (numbers,
// parameters, time, nodes built only from those) contribute no terms past
// their degree, so e.g. a product with a parameter costs one multiply per order.
// Divisions inside the recurrences become multiplications by reciprocals
// computed once per step; 1/k and j/k factors are literals.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <set>
#include <sstream>
#include <stdexcept>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "../../include/hyx_oplist.h"
#include "hyx_internal.h"

namespace hyx
{

static const int INF_DEG = 1 << 20;

// Systems with events: the event functions' coefficients of the last two orders
// join the state's in the step-size norms (heyoka includes its event functions in
// the Taylor jet whose last two orders set the step size; the step counts and
// step sizes of doc/notebooks/Event detection.ipynb, cell 21, follow from it).

void compute_degrees(Sys &S)
{
    // Highest order with a possibly non-zero coefficient; state variables are full series.
    S.deg.assign(S.nnodes, INF_DEG);
    auto od = [&](uint32_t a) {
        return HYX_OPERAND_KIND(a) == HYX_KIND_NODE ? S.deg[HYX_OPERAND_INDEX(a)] : 0;
    };
    for (uint32_t j = S.n; j < S.nnodes; ++j) {
        const auto &A = S.args[j];
        int d = 0;
        switch (S.op[j]) {
        case HYX_OP_TIME: d = 1; break;
        case HYX_OP_ADD:
        case HYX_OP_SUB:
        case HYX_OP_SUM:
        case HYX_OP_NEG:
            for (auto a : A) d = std::max(d, od(a));
            break;
        case HYX_OP_MUL: d = std::min(INF_DEG, od(A[0]) + od(A[1])); break;
        case HYX_OP_SQUARE: d = std::min(INF_DEG, 2 * od(A[0])); break;
        case HYX_OP_DIV: d = (od(A[1]) == 0) ? od(A[0]) : INF_DEG; break;
        case HYX_OP_TWIN_BIND: d = 0; break;
        case HYX_OP_TWIN_XCH: d = od(A[0]); break;
        case HYX_OP_QUAD_BIND: d = 0; break;
        case HYX_OP_QUAD_SWAP: d = od(A[0]); break;
        default: d = (od(A[0]) == 0) ? 0 : INF_DEG;
        }
        S.deg[j] = d;
    }
}

int parse_oplist(const uint32_t *w, uint32_t n_words, const double *consts, uint32_t n_consts, Sys &S,
                 std::string &err)
{
    if (n_words < 6 || w[0] != HYX_OPLIST_MAGIC) {
        err = "invalid op-list: bad magic or size";
        return -1;
    }
    S.n = w[1];
    S.npar = w[2];
    S.nnodes = w[3];
    S.nconsts = w[4];
    if (S.n == 0 || S.nnodes < S.n || S.nconsts != n_consts) {
        err = "invalid op-list header";
        return -1;
    }
    S.consts.assign(consts, consts + n_consts);
    S.op.assign(S.nnodes, 0);
    S.aux.assign(S.nnodes, 0);
    S.args.assign(S.nnodes, {});
    uint32_t off = 6;
    for (uint32_t j = S.n; j < S.nnodes; ++j) {
        if (off + 3 > n_words) {
            err = "invalid op-list: truncated node record";
            return -1;
        }
        S.op[j] = w[off];
        const uint32_t na = w[off + 1];
        S.aux[j] = w[off + 2];
        if (off + 3 + na > n_words) {
            err = "invalid op-list: truncated arguments";
            return -1;
        }
        for (uint32_t q = 0; q < na; ++q) {
            const uint32_t a = w[off + 3 + q];
            const uint32_t kind = HYX_OPERAND_KIND(a), idx = HYX_OPERAND_INDEX(a);
            if ((kind == HYX_KIND_NODE && idx >= j) || (kind == HYX_KIND_CONST && idx >= S.nconsts)
                || (kind == HYX_KIND_PAR && idx >= S.npar) || kind > HYX_KIND_PAR) {
                err = "invalid op-list: operand out of range";
                return -1;
            }
            S.args[j].push_back(a);
        }
        const uint32_t op = S.op[j];
        const size_t need = (op == HYX_OP_TIME) ? 0
                            : (op == HYX_OP_ADD || op == HYX_OP_SUB || op == HYX_OP_MUL || op == HYX_OP_DIV) ? 2
                            : (op == HYX_OP_SUM) ? S.args[j].size()
                                                 : 1;
        if (op < HYX_OP_ADD || op > HYX_OP_SUM || S.args[j].size() != need || (op == HYX_OP_SUM && need < 2)) {
            err = "invalid op-list: bad opcode or arity at node " + std::to_string(j);
            return -1;
        }
        if (op == HYX_OP_POW && S.aux[j] >= S.nconsts) {
            err = "invalid op-list: pow exponent out of range";
            return -1;
        }
        if ((op == HYX_OP_SIN || op == HYX_OP_COS)) {
            const uint32_t p = S.aux[j];
            if (p < S.n || p >= S.nnodes) {
                err = "invalid op-list: sin/cos partner out of range";
                return -1;
            }
        }
        off += 3 + na;
    }
    if (off != w[5] || off + S.n != n_words) {
        err = "invalid op-list: rhs offset mismatch";
        return -1;
    }
    S.rhs.assign(w + off, w + off + S.n);
    for (uint32_t i = 0; i < S.n; ++i) {
        const uint32_t kind = HYX_OPERAND_KIND(S.rhs[i]), idx = HYX_OPERAND_INDEX(S.rhs[i]);
        if ((kind == HYX_KIND_NODE && idx >= S.nnodes) || (kind == HYX_KIND_CONST && idx >= S.nconsts)
            || (kind == HYX_KIND_PAR && idx >= S.npar) || kind > HYX_KIND_PAR) {
            err = "invalid op-list: rhs operand out of range";
            return -1;
        }
    }
    // sin/cos partners must point at each other with the same argument.
    for (uint32_t j = S.n; j < S.nnodes; ++j) {
        if (S.op[j] == HYX_OP_SIN || S.op[j] == HYX_OP_COS) {
            const uint32_t p = S.aux[j];
            const uint32_t want = (S.op[j] == HYX_OP_SIN) ? HYX_OP_COS : HYX_OP_SIN;
            if (S.op[p] != want || S.aux[p] != j || S.args[p] != S.args[j]) {
                err = "invalid op-list: inconsistent sin/cos pair";
                return -1;
            }
        }
    }
    compute_degrees(S);
    return 0;
}

std::string lit(double v)
{
    if (std::isnan(v)) return "(__builtin_nan(\"\"))";
    if (std::isinf(v)) return v > 0 ? "(__builtin_inf())" : "(-__builtin_inf())";
    char buf[64];
    std::snprintf(buf, sizeof(buf), "(%a)", v);
    return buf;
}

namespace
{

struct Gen {
    const Sys &S;
    const int P;
    const bool high_acc;
    std::ostringstream o;
    std::set<std::pair<uint32_t, int>> zset; // (node, k) known to be exactly zero
    std::set<std::pair<uint32_t, int>> dvars; // generated j * a^[j] values
    std::vector<char> perlane;                // constants held per lane (kc_prefix[]) instead of literals
    std::string kc_prefix = "kc";
    // State coefficients in LDS: each lane keeps the rows x^[k] of its instance in
    // xs[(k * n + i) * 64] (a per-block __shared__ array, the lane's column), and reads
    // the orders below the one being generated back from there, instead of holding
    // n * (p + 1) doubles in registers until the state update.
    bool lds_state = false;
    int cur_k = 0;
    // pow recurrence with literal coefficients (no i c^[i] history). Measured: the lane
    // form's CR3BP step drops from 4,232 to 3,853 instructions and 411 to 341 registers;
    // the group form's 6-body step runs 2.6 % slower with it (its register count stays
    // above the two-wave threshold either way), so the group form keeps the other form.
    bool literal_pow = true;

    // Events (unrolled form): the event op-list, whose rows 0..n-1 are this system's state
    // variables; its nodes are generated by a second Gen (base = this) in a nested scope.
    const Sys *E = nullptr;
    uint32_t nev = 0;
    const Gen *base = nullptr;
    // Twin form (HYX_OP_TWIN_BIND / HYX_OP_TWIN_XCH nodes, see plan_twin): two lanes per
    // instance, the lane flag hyx_twin_b and the pair broadcasts hyx_twin_from_a / _b().
    bool twin = false;
    // Quad form (with twin set; HYX_OP_QUAD_BIND / HYX_OP_QUAD_SWAP nodes, see QuadPlanner):
    // four lanes per instance, q = 2 * primary + component. Lane flags hyx_twin_b (primary B)
    // and hyx_quad_y (component y); hyx_twin_from_a / _b broadcast the same component's lane
    // of primary A / B, hyx_quad_swap the other component's lane of the same primary. The
    // state is held in lane-relative order: variable i is x[i] on x lanes and x[qsig[i]] on
    // y lanes (qsig: the pairing of the state variables, an involution).
    bool quad = false;
    std::vector<uint32_t> qsig;
    // No implicit fusion of a product into a following add inside the step: the lane
    // form of a system that has a twin form must round exactly as the twin does, and
    // the exchange changes which products the compiler would fuse.
    bool no_contract = false;

    // Group form events: the state variables' coefficients are read from the LDS rows
    // fl[lds_xb + k * lds_nx + lds_col[i]] (lds_xb < 0: not used).
    int lds_xb = -1, lds_nx = 0;
    std::vector<int> lds_col; // column of each state variable (empty: the variable's index)

    Gen(const Sys &s, int p, bool ha) : S(s), P(p), high_acc(ha), perlane(s.nconsts, 0) {}

    static std::string cv(uint32_t node, int k) { return "c" + std::to_string(node) + "_" + std::to_string(k); }

    std::string xs(uint32_t i, int k) const { return "xs[" + std::to_string(((uint32_t)k * S.n + i) * 64u) + "]"; }

    // Coefficient k of node idx: a register, or for a state variable in LDS an older row.
    std::string nv(uint32_t idx, int k) const
    {
        if (lds_xb >= 0 && idx < S.n)
            return "fl[" + std::to_string(lds_xb + k * lds_nx + (lds_col.empty() ? (int)idx : lds_col[idx])) + "]";
        if (base && idx < base->S.n) return base->nv(idx, k);
        if (lds_state && idx < S.n && k != cur_k) return xs(idx, k);
        return cv(idx, k);
    }

    bool zero(uint32_t w, int k) const
    {
        const uint32_t kind = HYX_OPERAND_KIND(w), idx = HYX_OPERAND_INDEX(w);
        if (kind != HYX_KIND_NODE)
            return k > 0 || (kind == HYX_KIND_CONST && !perlane[idx] && S.consts[idx] == 0.0);
        if (k > S.deg[idx]) return true;
        return zset.count({idx, k}) != 0;
    }

    // Degree-0 operand: a number, a parameter or a node whose series is constant.
    bool scalar(uint32_t w) const
    {
        const uint32_t kind = HYX_OPERAND_KIND(w);
        return kind != HYX_KIND_NODE || S.deg[HYX_OPERAND_INDEX(w)] == 0;
    }

    std::string val(uint32_t w, int k) const
    {
        const uint32_t kind = HYX_OPERAND_KIND(w), idx = HYX_OPERAND_INDEX(w);
        if (kind == HYX_KIND_NODE) return nv(idx, k);
        if (k > 0) return "0.0";
        if (kind == HYX_KIND_CONST)
            return perlane[idx] ? kc_prefix + "[" + std::to_string(idx) + "]" : lit(S.consts[idx]);
        return "pv[" + std::to_string(idx) + "]";
    }

    // j * a^[j], shared by the exp / sin / cos recurrences over the same argument.
    std::string dval(uint32_t w, int j)
    {
        if (j == 1) return val(w, 1);
        const uint32_t idx = HYX_OPERAND_INDEX(w);
        const std::string nm = "d" + std::to_string(idx) + "_" + std::to_string(j);
        if (!dvars.count({idx, j})) {
            dvars.insert({idx, j});
            o << "    const double " << nm << " = " << lit((double)j) << " * " << val(w, j) << ";\n";
        }
        return nm;
    }

    // i * c^[i] of node j's own history (pow recurrence), formed once.
    std::string sval(uint32_t j, int i)
    {
        if (i == 1) return cv(j, 1);
        const std::string nm = "z" + std::to_string(j) + "_" + std::to_string(i);
        if (!dvars.count({j | 0x80000000u, i})) {
            dvars.insert({j | 0x80000000u, i});
            o << "    const double " << nm << " = " << lit((double)i) << " * " << cv(j, i) << ";\n";
        }
        return nm;
    }

    // True when every pow coefficient k alpha - i (alpha + 1), k <= P, i < k, is exact with
    // zero low 32 bits (usable as a 32-bit literal of a double operand).
    bool short_pow_coeffs(double al) const
    {
        for (int k = 1; k <= P; ++k)
            for (int i = 0; i < k; ++i) {
                const double cf = (double)k * al - (double)i * (al + 1.0);
                uint64_t b;
                std::memcpy(&b, &cf, 8);
                if ((b & 0xffffffffull) != 0) return false;
            }
        return true;
    }

    void set_zero(uint32_t node, int k)
    {
        zset.insert({node, k});
        o << "    const double " << cv(node, k) << " = 0.0;\n";
    }

    // acc = sum of products (terms), as an fma chain; "" when empty. With split_min > 0,
    // chains of at least split_min terms run as two interleaved accumulators (even and
    // odd terms) added at the end: more independent FP64 work in flight per wave, at the
    // price of a different summation order.
    int split_min = 0;
    std::string fma_chain(const std::vector<std::pair<std::string, std::string>> &t, const std::string &nm,
                          std::ostringstream &o) const
    {
        if (t.empty()) return "";
        if (split_min > 0 && (int)t.size() >= split_min) {
            const std::string a = nm + "a", b = nm + "b";
            o << "    double " << a << " = " << t[0].first << " * " << t[0].second << ";\n";
            o << "    double " << b << " = " << t[1].first << " * " << t[1].second << ";\n";
            for (size_t q = 2; q < t.size(); ++q) {
                const std::string &acc = (q % 2 == 0) ? a : b;
                o << "    " << acc << " = __builtin_fma(" << t[q].first << ", " << t[q].second << ", " << acc << ");\n";
            }
            o << "    const double " << nm << " = " << a << " + " << b << ";\n";
            return nm;
        }
        o << "    double " << nm << " = " << t[0].first << " * " << t[0].second << ";\n";
        for (size_t q = 1; q < t.size(); ++q)
            o << "    " << nm << " = __builtin_fma(" << t[q].first << ", " << t[q].second << ", " << nm << ");\n";
        return nm;
    }

    void order0(uint32_t j)
    {
        const auto &A = S.args[j];
        const std::string c = cv(j, 0);
        auto v = [&](size_t q) { return val(A[q], 0); };
        switch (S.op[j]) {
        case HYX_OP_ADD: o << "    const double " << c << " = " << v(0) << " + " << v(1) << ";\n"; break;
        case HYX_OP_SUB: o << "    const double " << c << " = " << v(0) << " - " << v(1) << ";\n"; break;
        case HYX_OP_NEG: o << "    const double " << c << " = -" << v(0) << ";\n"; break;
        case HYX_OP_SUM: {
            o << "    const double " << c << " = " << v(0);
            for (size_t q = 1; q < A.size(); ++q) o << " + " << v(q);
            o << ";\n";
            break;
        }
        case HYX_OP_MUL: o << "    const double " << c << " = " << v(0) << " * " << v(1) << ";\n"; break;
        case HYX_OP_DIV:
            o << "    const double " << c << " = " << v(0) << " / " << v(1) << ";\n";
            if (S.deg[j] > 0) o << "    const double rb" << j << " = 1.0 / " << v(1) << ";\n";
            break;
        case HYX_OP_SQUARE: o << "    const double " << c << " = " << v(0) << " * " << v(0) << ";\n"; break;
        case HYX_OP_SQRT:
            o << "    const double " << c << " = sqrt(" << v(0) << ");\n";
            if (S.deg[j] > 0) o << "    const double rs" << j << " = 0.5 / " << c << ";\n";
            break;
        case HYX_OP_POW:
            o << "    const double " << c << " = pow(" << v(0) << ", " << lit(S.consts[S.aux[j]]) << ");\n";
            if (S.deg[j] > 0) {
                const double al = S.consts[S.aux[j]];
                o << "    const double ra" << j << " = 1.0 / " << v(0) << ";\n";
                o << "    const double raa" << j << " = ra" << j << " * " << lit(al) << ";\n";
                o << "    const double rab" << j << " = ra" << j << " * " << lit(al + 1.0) << ";\n";
            }
            break;
        case HYX_OP_EXP: o << "    const double " << c << " = exp(" << v(0) << ");\n"; break;
        case HYX_OP_LOG:
            o << "    const double " << c << " = log(" << v(0) << ");\n";
            if (S.deg[j] > 0) o << "    const double ra" << j << " = 1.0 / " << v(0) << ";\n";
            break;
        case HYX_OP_SIN: {
            // The pair shares one sincos() call, issued at the sin node.
            const uint32_t p = S.aux[j];
            if (p > j) {
                o << "    double " << c << ", " << cv(p, 0) << ";\n";
                o << "    sincos(" << v(0) << ", &" << c << ", &" << cv(p, 0) << ");\n";
            }
            break;
        }
        case HYX_OP_COS: {
            const uint32_t p = S.aux[j];
            if (p > j) {
                o << "    double " << cv(p, 0) << ", " << c << ";\n";
                o << "    sincos(" << v(0) << ", &" << cv(p, 0) << ", &" << c << ");\n";
            }
            break;
        }
        case HYX_OP_TIME: o << "    const double " << c << " = t0;\n"; break;
        case HYX_OP_TWIN_BIND: {
            // The lane's own degree-0 value: lane A the first operand, lane B the second,
            // each negated when its aux bit says so.
            const std::string a = ((S.aux[j] & 1u) ? "-" : "") + v(0), b = ((S.aux[j] & 2u) ? "-" : "") + v(1);
            o << "    const double " << c << " = hyx_twin_b ? (" << b << ") : (" << a << ");\n";
            break;
        }
        case HYX_OP_TWIN_XCH: xch(j, 0); break;
        case HYX_OP_QUAD_BIND:
            // The component's own degree-0 value: x lanes operand 0, y lanes operand 1.
            o << "    const double " << c << " = hyx_quad_y ? (" << v(1) << ") : (" << v(0) << ");\n";
            break;
        case HYX_OP_QUAD_SWAP: qswap(j, 0); break;
        }
    }

    // The other component's value of a quad node at order k (lanes x <-> y of a primary).
    void qswap(uint32_t j, int k)
    {
        const uint32_t w = S.args[j][0];
        if (zero(w, k)) {
            set_zero(j, k);
            return;
        }
        o << "    const double " << cv(j, k) << " = hyx_quad_swap(" << val(w, k) << ");\n";
    }

    // Exchange node j at order k: aux 0 gives lane A's value of the slot on both lanes,
    // aux 1 lane B's: a DPP broadcast from that lane of the pair.
    void xch(uint32_t j, int k)
    {
        const uint32_t w = S.args[j][0];
        if (zero(w, k)) {
            set_zero(j, k);
            return;
        }
        o << "    const double " << cv(j, k) << " = " << (S.aux[j] ? "hyx_twin_from_b(" : "hyx_twin_from_a(")
          << val(w, k) << ");\n";
    }

    void orderk(uint32_t j, int k)
    {
        if (k > S.deg[j]) return; // zero by degree; never referenced
        const auto &A = S.args[j];
        const std::string c = cv(j, k);
        const std::string ks = std::to_string(k);
        std::vector<std::pair<std::string, std::string>> t;
        switch (S.op[j]) {
        case HYX_OP_ADD:
        case HYX_OP_SUB:
        case HYX_OP_NEG:
        case HYX_OP_SUM: {
            std::string e;
            bool first = true;
            for (size_t q = 0; q < A.size(); ++q) {
                if (zero(A[q], k)) continue;
                const bool neg = (S.op[j] == HYX_OP_NEG) || (S.op[j] == HYX_OP_SUB && q == 1);
                if (first)
                    e = (neg ? "-" : "") + val(A[q], k);
                else
                    e += (neg ? " - " : " + ") + val(A[q], k);
                first = false;
            }
            if (first)
                set_zero(j, k);
            else
                o << "    const double " << c << " = " << e << ";\n";
            break;
        }
        case HYX_OP_MUL: {
            for (int q = 0; q <= k; ++q)
                if (!zero(A[0], q) && !zero(A[1], k - q)) t.push_back({val(A[0], q), val(A[1], k - q)});
            if (t.empty())
                set_zero(j, k);
            else if (t.size() == 1)
                o << "    const double " << c << " = " << t[0].first << " * " << t[0].second << ";\n";
            else {
                fma_chain(t, "m" + std::to_string(j) + "_" + ks, o);
                o << "    const double " << c << " = m" << j << "_" << ks << ";\n";
            }
            break;
        }
        case HYX_OP_DIV: {
            // c^[k] = (a^[k] - sum_{i=1}^{k} b^[i] c^[k-i]) / b^[0]
            for (int q = 1; q <= k; ++q)
                if (!zero(A[1], q) && !zero(HYX_OPERAND(HYX_KIND_NODE, j), k - q)) t.push_back({val(A[1], q), cv(j, k - q)});
            const std::string acc = fma_chain(t, "s" + std::to_string(j) + "_" + ks, o);
            const std::string a = zero(A[0], k) ? "0.0" : val(A[0], k);
            if (acc.empty() && a == "0.0")
                set_zero(j, k);
            else
                o << "    const double " << c << " = (" << a << (acc.empty() ? "" : " - " + acc) << ") * rb" << j << ";\n";
            break;
        }
        case HYX_OP_SQUARE: {
            for (int q = 0; 2 * q < k; ++q)
                if (!zero(A[0], q) && !zero(A[0], k - q)) t.push_back({val(A[0], q), val(A[0], k - q)});
            const std::string acc = fma_chain(t, "q" + std::to_string(j) + "_" + ks, o);
            const bool mid = (k % 2 == 0) && !zero(A[0], k / 2);
            if (acc.empty() && !mid)
                set_zero(j, k);
            else if (acc.empty())
                o << "    const double " << c << " = " << val(A[0], k / 2) << " * " << val(A[0], k / 2) << ";\n";
            else if (!mid)
                o << "    const double " << c << " = " << acc << " + " << acc << ";\n";
            else
                o << "    const double " << c << " = __builtin_fma(" << val(A[0], k / 2) << ", " << val(A[0], k / 2)
                  << ", " << acc << " + " << acc << ");\n";
            break;
        }
        case HYX_OP_SQRT: {
            // c^[k] = (a^[k] - sum_{i=1}^{k-1} c^[i] c^[k-i]) / (2 c^[0])
            for (int q = 1; 2 * q < k; ++q) t.push_back({cv(j, q), cv(j, k - q)});
            std::string acc = fma_chain(t, "r" + std::to_string(j) + "_" + ks, o);
            std::string tot;
            if (!acc.empty()) tot = acc + " + " + acc;
            if (k % 2 == 0 && k >= 2) {
                const std::string m = cv(j, k / 2);
                tot = tot.empty() ? (m + " * " + m) : ("__builtin_fma(" + m + ", " + m + ", " + tot + ")");
            }
            const std::string a = zero(A[0], k) ? "0.0" : val(A[0], k);
            o << "    const double " << c << " = (" << a << (tot.empty() ? "" : " - (" + tot + ")") << ") * rs" << j
              << ";\n";
            break;
        }
        case HYX_OP_POW: {
            // c^[k] = 1/(k a^[0]) sum_{i=0}^{k-1} (k alpha - i (alpha + 1)) a^[k-i] c^[i].
            const double al = S.consts[S.aux[j]];
            if (literal_pow && short_pow_coeffs(al)) {
                // Every coefficient k alpha - i (alpha + 1) is a double whose low 32 bits are
                // zero (alpha a multiple of 1/2, as for r^-3/2): it is the 32-bit literal
                // operand of v_fmac_f64, so the sum is one product and one fma per term with
                // no extra history (the form below keeps i c^[i] for every order).
                std::string acc;
                const std::string nm = "w" + std::to_string(j) + "_" + ks;
                for (int q = 0; q < k; ++q) {
                    if (zero(A[0], k - q)) continue;
                    const double cf = (double)k * al - (double)q * (al + 1.0);
                    if (cf == 0.0) continue;
                    const std::string prod = "(" + val(A[0], k - q) + " * " + cv(j, q) + ")";
                    if (acc.empty()) {
                        o << "    double " << nm << " = " << lit(cf) << " * " << prod << ";\n";
                        acc = nm;
                    } else
                        o << "    " << nm << " = __builtin_fma(" << lit(cf) << ", " << prod << ", " << nm << ");\n";
                }
                if (acc.empty())
                    set_zero(j, k);
                else
                    o << "    const double " << c << " = " << acc << " * (ra" << j << " * " << lit(1.0 / k) << ");\n";
                break;
            }
            // General alpha: (alpha / a^[0]) S1 - ((alpha + 1) / a^[0]) (1/k) S2,
            // S1 = sum a^[k-i] c^[i], S2 = sum a^[k-i] (i c^[i]): two fma chains without
            // per-term coefficients (i c^[i] is formed once per order).
            std::vector<std::pair<std::string, std::string>> t2;
            for (int q = 0; q < k; ++q) {
                if (zero(A[0], k - q)) continue;
                t.push_back({val(A[0], k - q), cv(j, q)});
                if (q > 0) t2.push_back({val(A[0], k - q), sval(j, q)});
            }
            const std::string s1 = fma_chain(t, "w" + std::to_string(j) + "_" + ks, o);
            const std::string s2 = fma_chain(t2, "u" + std::to_string(j) + "_" + ks, o);
            if (s1.empty())
                set_zero(j, k);
            else if (s2.empty())
                o << "    const double " << c << " = " << s1 << " * raa" << j << ";\n";
            else
                o << "    const double " << c << " = __builtin_fma(" << s1 << ", raa" << j << ", -(" << s2 << " * rab" << j
                  << ") * " << lit(1.0 / k) << ");\n";
            break;
        }
        case HYX_OP_EXP: {
            // c^[k] = (1/k) sum_{i=1}^{k} i a^[i] c^[k-i]
            for (int q = 1; q <= k; ++q)
                if (!zero(A[0], q)) t.push_back({dval(A[0], q), cv(j, k - q)});
            const std::string acc = fma_chain(t, "e" + std::to_string(j) + "_" + ks, o);
            if (acc.empty())
                set_zero(j, k);
            else
                o << "    const double " << c << " = " << acc << " * " << lit(1.0 / k) << ";\n";
            break;
        }
        case HYX_OP_LOG: {
            // c^[k] = (a^[k] - (1/k) sum_{i=1}^{k-1} i c^[i] a^[k-i]) / a^[0]
            for (int q = 1; q < k; ++q)
                if (!zero(A[0], k - q)) t.push_back({lit((double)q / k), "(" + cv(j, q) + " * " + val(A[0], k - q) + ")"});
            const std::string acc = fma_chain(t, "l" + std::to_string(j) + "_" + ks, o);
            const std::string a = zero(A[0], k) ? "0.0" : val(A[0], k);
            o << "    const double " << c << " = (" << a << (acc.empty() ? "" : " - " + acc) << ") * ra" << j << ";\n";
            break;
        }
        case HYX_OP_SIN:
        case HYX_OP_COS: {
            // s^[k] = (1/k) sum i a^[i] c^[k-i];  c^[k] = -(1/k) sum i a^[i] s^[k-i]
            const uint32_t p = S.aux[j];
            for (int q = 1; q <= k; ++q)
                if (!zero(A[0], q)) t.push_back({dval(A[0], q), cv(p, k - q)});
            const std::string acc = fma_chain(t, "g" + std::to_string(j) + "_" + ks, o);
            if (acc.empty())
                set_zero(j, k);
            else
                o << "    const double " << c << " = " << acc << " * " << lit(S.op[j] == HYX_OP_SIN ? 1.0 / k : -1.0 / k)
                  << ";\n";
            break;
        }
        case HYX_OP_TIME:
            if (k == 1)
                o << "    const double " << c << " = 1.0;\n";
            else
                set_zero(j, k);
            break;
        case HYX_OP_TWIN_XCH: xch(j, k); break;
        case HYX_OP_QUAD_SWAP: qswap(j, k); break;
        }
    }

    std::string run()
    {
        const uint32_t n = S.n;
        o << "#define HYX_N " << n << "\n";
        o << "#define HYX_NPAR " << std::max<uint32_t>(S.npar, 1) << "\n";
        o << "#define HYX_P " << P << "\n";
        o << "typedef unsigned long long hyx_u64;\n";
        if (quad) {
            // Four lanes per instance (lane q = 2 * primary + component): identical stores,
            // the quad nodes split between them.
            o << "#define HYX_SUB 4\n#define HYX_TWIN 1\n#define HYX_QUAD 1\n";
            o << "#ifndef HYX_CPU_EMU\n";
            o << "__device__ __forceinline__ unsigned hyx_quad_lane() { return __lane_id() & 3u; }\n";
            // DPP quad permutations [0, 1, 0, 1] (primary A, same component), [2, 3, 2, 3]
            // (primary B, same component) and [1, 0, 3, 2] (same primary, other component).
            o << "template <int CTL>\n__device__ __forceinline__ double hyx_twin_bcast(double v)\n{\n"
                 "    const hyx_u64 b = __builtin_bit_cast(hyx_u64, v);\n"
                 "    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTL, 0xF, 0xF, true);\n"
                 "    const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTL, 0xF, 0xF, true);\n"
                 "    return __builtin_bit_cast(double, ((hyx_u64)(unsigned)hi << 32) | (hyx_u64)(unsigned)lo);\n}\n"
                 "__device__ __forceinline__ double hyx_twin_from_a(double v) { return hyx_twin_bcast<0x44>(v); }\n"
                 "__device__ __forceinline__ double hyx_twin_from_b(double v) { return hyx_twin_bcast<0xEE>(v); }\n"
                 "__device__ __forceinline__ double hyx_quad_swap(double v) { return hyx_twin_bcast<0xB1>(v); }\n";
            o << "#endif\n";
        } else if (twin) {
            // Two lanes per instance (adjacent, lane B odd): identical state and stores, the
            // twin nodes split between them.
            o << "#define HYX_SUB 2\n#define HYX_TWIN 1\n";
            o << "#ifndef HYX_CPU_EMU\n";
            o << "__device__ __forceinline__ bool hyx_twin_lane_b() { return (__lane_id() & 1u) != 0u; }\n";
            // A lane's value to both lanes of its pair: DPP quad permutations [0, 0, 2, 2]
            // (from lane A) and [1, 1, 3, 3] (from lane B).
            o << "template <int CTL>\n__device__ __forceinline__ double hyx_twin_bcast(double v)\n{\n"
                 "    const hyx_u64 b = __builtin_bit_cast(hyx_u64, v);\n"
                 "    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTL, 0xF, 0xF, true);\n"
                 "    const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTL, 0xF, 0xF, true);\n"
                 "    return __builtin_bit_cast(double, ((hyx_u64)(unsigned)hi << 32) | (hyx_u64)(unsigned)lo);\n}\n"
                 "__device__ __forceinline__ double hyx_twin_from_a(double v) { return hyx_twin_bcast<0xA0>(v); }\n"
                 "__device__ __forceinline__ double hyx_twin_from_b(double v) { return hyx_twin_bcast<0xF5>(v); }\n";
            o << "#endif\n";
        }
        if (lds_state) {
            o << "#define HYX_LDS_STATE 1\n";
            o << "#ifndef HYX_CPU_EMU\n";
            o << "__shared__ double hyx_xs[" << (uint64_t)(P + 1) * n * 64 << "];\n";
            o << "__device__ __forceinline__ double *hyx_xs_lane() { return hyx_xs + threadIdx.x; }\n";
            o << "#endif\n";
        }
        o << "template <bool WTC>\n"
             "__device__ __forceinline__ void hyx_taylor_step(double (&x)[HYX_N], const double (&pv)[HYX_NPAR], "
             "double t0, double max_dt, double &h_out, bool &nf_out, double *__restrict__ tc, hyx_u64 tcs, "
             "double *__restrict__ tape, hyx_u64 tps"
          << (E ? ", unsigned long long &hyx_evm" : "") << ")\n{\n";
        if (no_contract) o << "#pragma clang fp contract(off)\n";
        o << "    (void)pv; (void)t0; (void)tape; (void)tps; (void)tc; (void)tcs;\n";
        if (quad)
            o << "    const unsigned hyx_q = hyx_quad_lane();\n    const bool hyx_twin_b = (hyx_q & 2u) != 0u;\n"
                 "    const bool hyx_quad_y = (hyx_q & 1u) != 0u;\n";
        else if (twin)
            o << "    const bool hyx_twin_b = hyx_twin_lane_b();\n";
        if (lds_state) o << "    double *const xs = hyx_xs_lane();\n";
        // The rows of order k go to LDS as soon as they exist; the empty asm at the
        // end of each order keeps the compiler from forwarding them in registers.
        auto store_rows = [&](int k) {
            if (!lds_state) return;
            for (uint32_t i = 0; i < n; ++i) o << "    " << xs(i, k) << " = " << cv(i, k) << ";\n";
            o << "    __asm__ volatile(\"\" ::: \"memory\");\n";
        };
        // order 0
        cur_k = 0;
        for (uint32_t i = 0; i < n; ++i) {
            if (quad && qsig[i] != i)
                o << "    const double " << cv(i, 0) << " = hyx_quad_y ? x[" << qsig[i] << "] : x[" << i << "];\n";
            else
                o << "    const double " << cv(i, 0) << " = x[" << i << "];\n";
        }
        store_rows(0);
        for (uint32_t j = n; j < S.nnodes; ++j) order0(j);
        for (int k = 1; k <= P; ++k) {
            o << "    // order " << k << "\n";
            cur_k = k;
            for (uint32_t i = 0; i < n; ++i) {
                const uint32_t r = S.rhs[i];
                if (zero(r, k - 1))
                    set_zero(i, k);
                else {
                    // The right-hand side's order k-1 is a node of the previous order.
                    cur_k = k - 1;
                    const std::string rv = val(r, k - 1);
                    cur_k = k;
                    o << "    const double " << cv(i, k) << " = " << rv << " * " << lit(1.0 / k) << ";\n";
                }
            }
            store_rows(k);
            if (k < P)
                for (uint32_t j = n; j < S.nnodes; ++j) orderk(j, k);
        }
        cur_k = P;
        // Taylor coefficients of the state, when asked for.
        o << "    if constexpr (WTC) {\n";
        for (uint32_t i = 0; i < n; ++i)
            for (int k = 0; k <= P; ++k) {
                if (quad && qsig[i] != i)
                    o << "        tc[(hyx_u64)(hyx_quad_y ? " << (qsig[i] * (P + 1) + k) << " : " << (i * (P + 1) + k)
                      << ") * tcs] = " << nv(i, k) << ";\n";
                else
                    o << "        tc[(hyx_u64)" << (i * (P + 1) + k) << " * tcs] = " << nv(i, k) << ";\n";
            }
        o << "    }\n";
        if (E) {
            // The event functions' coefficients, from the state's registers: a nested scope
            // (the event op-list's names shadow the system's), results in hyx_evg.
            Gen eg(*E, P, high_acc);
            eg.base = this;
            eg.literal_pow = literal_pow;
            o << "    double hyx_evg[" << nev << "][" << (P + 1) << "];\n    {\n";
            for (uint32_t j = E->n; j < E->nnodes; ++j) eg.order0(j);
            for (int k = 1; k <= P; ++k)
                for (uint32_t j = E->n; j < E->nnodes; ++j) eg.orderk(j, k);
            for (uint32_t e = 0; e < nev; ++e)
                for (int k = 0; k <= P; ++k)
                    eg.o << "    hyx_evg[" << e << "][" << k << "] = " << (eg.zero(E->rhs[e], k) ? "0.0" : eg.val(E->rhs[e], k))
                         << ";\n";
            o << eg.o.str() << "    }\n";
        }
        // Step size: infinity norms of the state and of the last two orders.
        o << "    double mx = 0.0, mo = 0.0, mom1 = 0.0, ab;\n";
        for (uint32_t i = 0; i < n; ++i) {
            o << "    ab = fabs(" << nv(i, 0) << "); mx = (ab > mx) ? ab : mx;\n";
            o << "    ab = fabs(" << nv(i, P) << "); mo = (ab > mo) ? ab : mo;\n";
            o << "    ab = fabs(" << nv(i, P - 1) << "); mom1 = (ab > mom1) ? ab : mom1;\n";
        }
        for (uint32_t e = 0; E && e < nev; ++e) {
            o << "    ab = fabs(hyx_evg[" << e << "][" << P << "]); mo = (ab > mo) ? ab : mo;\n";
            o << "    ab = fabs(hyx_evg[" << e << "][" << (P - 1) << "]); mom1 = (ab > mom1) ? ab : mom1;\n";
        }
        o << "    const double num = (mx <= 1.0) ? 1.0 : mx;\n";
        o << "    const double rho_o = pow(num / mo, " << lit(1.0 / P) << ");\n";
        o << "    const double rho_om1 = pow(num / mom1, " << lit(1.0 / (P - 1)) << ");\n";
        o << "    const double rho_m = (rho_om1 < rho_o) ? rho_om1 : rho_o;\n";
        const double e1 = std::exp(1.0);
        const double rhofac = std::exp(-7.0 / 10.0 / (double)(P - 1)) / (e1 * e1);
        o << "    double h = rho_m * " << lit(rhofac) << ";\n";
        o << "    const double lim = fabs(max_dt);\n";
        o << "    h = (h < lim) ? h : lim;\n";
        o << "    if (max_dt < 0.0) h = -h;\n";
        if (E) {
            // Fast exclusion per event over the step (the same test the root scan starts
            // with); only steps it cannot clear keep their coefficients in memory.
            const uint64_t off = (uint64_t)n * (P + 1) + (uint64_t)(E->nnodes - E->n) * (P + 1);
            o << "    hyx_evm = 0ull;\n";
            o << "#pragma unroll\n    for (int e = 0; e < " << nev << "; ++e) {\n"
                 "        double hk = 1.0, mag = 0.0, lo = 0.0, hi = 0.0;\n"
                 "#pragma unroll\n        for (int k = 0; k <= " << P << "; ++k) {\n"
                 "            const double c = hyx_evg[e][k] * hk;\n            hk *= h;\n            mag += fabs(c);\n"
                 "            if (k > 0) { lo += fmin(c, 0.0); hi += fmax(c, 0.0); }\n        }\n"
                 "        const double slack = " << lit(4.0 * (P + 1) * 1.1102230246251565e-16) << " * mag;\n"
                 "        const double c0 = hyx_evg[e][0];\n"
                 "        if (mag > 0.0 && isfinite(mag) && !(c0 + lo > slack || c0 + hi < -slack)) hyx_evm |= 1ull << e;\n"
                 "    }\n";
            o << "    if (hyx_evm != 0ull) {\n";
            for (uint32_t e = 0; e < nev; ++e)
                for (int k = 0; k <= P; ++k)
                    o << "        tc[(hyx_u64)" << (off + (uint64_t)e * (P + 1) + k) << " * tcs] = hyx_evg[" << e << "][" << k
                      << "];\n";
            o << "        if constexpr (!WTC) {\n";
            for (uint32_t i = 0; i < n; ++i)
                for (int k = 0; k <= P; ++k)
                    o << "            tc[(hyx_u64)" << (i * (P + 1) + k) << " * tcs] = " << nv(i, k) << ";\n";
            o << "        }\n    }\n";
        }
        o << "    bool nf = false;\n";
        // Quad form: the new state in lane-relative order first, then back to x's order.
        const std::string xo = quad ? "hyx_xr" : "x";
        if (quad) o << "    double hyx_xr[HYX_N];\n";
        for (uint32_t i = 0; i < n; ++i) {
            if (!high_acc) {
                o << "    { double r = " << nv(i, P) << ";\n";
                for (int k = P - 1; k >= 0; --k) o << "      r = __builtin_fma(r, h, " << nv(i, k) << ");\n";
                o << "      " << xo << "[" << i << "] = r; nf = nf || !isfinite(r); }\n";
            } else {
                // Kahan-compensated sum of x^[k] h^k, k = 0..p.
                o << "    { double sm = " << nv(i, 0) << ", cp = 0.0, hk = 1.0;\n";
                for (int k = 1; k <= P; ++k)
                    o << "      { hk *= h; const double y = " << nv(i, k)
                      << " * hk - cp; const double tt = sm + y; cp = (tt - sm) - y; sm = tt; }\n";
                o << "      " << xo << "[" << i << "] = sm; nf = nf || !isfinite(sm); }\n";
            }
        }
        if (quad)
            for (uint32_t i = 0; i < n; ++i)
                o << "    x[" << i << "] = hyx_quad_y ? hyx_xr[" << qsig[i] << "] : hyx_xr[" << i << "];\n";
        o << "    h_out = h;\n    nf_out = nf;\n}\n";
        return o.str();
    }
};


// Compact form: a runtime loop over the orders with the node list unrolled
// inside it and the Taylor tape in device memory, laid out
// tape[(node * (P + 1) + k) * tps + lane] so each access of a wave is one
// contiguous 512-byte row. Code size grows with the node count only, not with
// node count x order^2, which keeps hipRTC compile times to seconds for
// systems with hundreds of nodes.
struct CompactGen {
    const Sys &S;
    const int P;
    const bool high_acc;
    std::ostringstream o;

    // Lanes per instance: each sums every sub-th term of a convolution, and a butterfly
    // of __shfl_xor adds the partial sums (every lane gets the same value, because the
    // two operands of each addition only swap between partner lanes).
    int sub = 1;
    // The system has events: the step takes hyx_evm and always writes tc (see run()).
    bool events = false;

    CompactGen(const Sys &s, int p, bool ha) : S(s), P(p), high_acc(ha) {}

    // "for (int q = lo + hyx_sub; cond; q += HYX_SUB)" and its reduction.
    static std::string loop(int lo, const std::string &cond)
    {
        return "for (int q = " + std::to_string(lo) + " + hyx_sub; " + cond + "; q += HYX_SUB)";
    }

    static std::string T(uint32_t node, const std::string &k)
    {
        return "T(" + std::to_string(node) + ", " + k + ")";
    }

    bool is_node(uint32_t w) const { return HYX_OPERAND_KIND(w) == HYX_KIND_NODE; }
    uint32_t idx(uint32_t w) const { return HYX_OPERAND_INDEX(w); }

    // Order-0 value of an operand.
    std::string v0(uint32_t w) const
    {
        const uint32_t kind = HYX_OPERAND_KIND(w), i = HYX_OPERAND_INDEX(w);
        if (kind == HYX_KIND_NODE) return T(i, "0");
        if (kind == HYX_KIND_CONST) return lit(S.consts[i]);
        return "pv[" + std::to_string(i) + "]";
    }

    // Order-k (k >= 1, runtime) value of an operand: 0 for numbers/parameters.
    std::string vk(uint32_t w, const std::string &k) const
    {
        if (!is_node(w)) return "0.0";
        return T(idx(w), k);
    }

    bool scalar(uint32_t w) const { return !is_node(w) || S.deg[idx(w)] == 0; }

    void order0(uint32_t j)
    {
        const auto &A = S.args[j];
        const std::string c = T(j, "0");
        switch (S.op[j]) {
        case HYX_OP_ADD: o << "    " << c << " = " << v0(A[0]) << " + " << v0(A[1]) << ";\n"; break;
        case HYX_OP_SUB: o << "    " << c << " = " << v0(A[0]) << " - " << v0(A[1]) << ";\n"; break;
        case HYX_OP_NEG: o << "    " << c << " = -" << v0(A[0]) << ";\n"; break;
        case HYX_OP_SUM:
            o << "    " << c << " = " << v0(A[0]);
            for (size_t q = 1; q < A.size(); ++q) o << " + " << v0(A[q]);
            o << ";\n";
            break;
        case HYX_OP_MUL: o << "    " << c << " = " << v0(A[0]) << " * " << v0(A[1]) << ";\n"; break;
        case HYX_OP_DIV:
            o << "    " << c << " = " << v0(A[0]) << " / " << v0(A[1]) << ";\n";
            o << "    const double rb" << j << " = 1.0 / " << v0(A[1]) << ";\n";
            break;
        case HYX_OP_SQUARE: o << "    { const double a = " << v0(A[0]) << "; " << c << " = a * a; }\n"; break;
        case HYX_OP_SQRT:
            o << "    " << c << " = sqrt(" << v0(A[0]) << ");\n";
            o << "    const double rs" << j << " = 0.5 / " << c << ";\n";
            break;
        case HYX_OP_POW:
            o << "    " << c << " = pow(" << v0(A[0]) << ", " << lit(S.consts[S.aux[j]]) << ");\n";
            o << "    const double ra" << j << " = 1.0 / " << v0(A[0]) << ";\n";
            break;
        case HYX_OP_EXP: o << "    " << c << " = exp(" << v0(A[0]) << ");\n"; break;
        case HYX_OP_LOG:
            o << "    " << c << " = log(" << v0(A[0]) << ");\n";
            o << "    const double ra" << j << " = 1.0 / " << v0(A[0]) << ";\n";
            break;
        case HYX_OP_SIN:
        case HYX_OP_COS: {
            const uint32_t p = S.aux[j];
            if (p > j) {
                const uint32_t si = S.op[j] == HYX_OP_SIN ? j : p, co = S.op[j] == HYX_OP_SIN ? p : j;
                o << "    { double s_, c_; sincos(" << v0(A[0]) << ", &s_, &c_); " << T(si, "0") << " = s_; "
                  << T(co, "0") << " = c_; }\n";
            }
            break;
        }
        case HYX_OP_TIME: o << "    " << c << " = t0;\n"; break;
        }
    }

    void orderk(uint32_t j)
    {
        const auto &A = S.args[j];
        const std::string c = T(j, "k");
        const int d = S.deg[j];
        if (d == 0) {
            o << "        " << c << " = 0.0;\n";
            return;
        }
        std::string guard_open, guard_close;
        if (d < INF_DEG) {
            o << "        if (k > " << d << ") { " << c << " = 0.0; } else {\n";
            guard_close = "        }\n";
        }
        switch (S.op[j]) {
        case HYX_OP_ADD:
        case HYX_OP_SUB:
        case HYX_OP_NEG:
        case HYX_OP_SUM: {
            std::string e;
            bool first = true;
            for (size_t q = 0; q < A.size(); ++q) {
                if (!is_node(A[q]) || S.deg[idx(A[q])] == 0) continue;
                const bool neg = (S.op[j] == HYX_OP_NEG) || (S.op[j] == HYX_OP_SUB && q == 1);
                e += first ? (neg ? "-" : "") + vk(A[q], "k") : (neg ? " - " : " + ") + vk(A[q], "k");
                first = false;
            }
            o << "        " << c << " = " << (first ? "0.0" : e) << ";\n";
            break;
        }
        case HYX_OP_MUL: {
            const bool s0 = scalar(A[0]), s1 = scalar(A[1]);
            if (s0 && s1)
                o << "        " << c << " = 0.0;\n";
            else if (s0)
                o << "        " << c << " = " << vk(A[1], "k") << " * " << v0(A[0]) << ";\n";
            else if (s1)
                o << "        " << c << " = " << vk(A[0], "k") << " * " << v0(A[1]) << ";\n";
            else
                o << "        { double acc = 0.0; " << loop(0, "q <= k") << " acc = __builtin_fma(" << T(idx(A[0]), "q")
                  << ", " << T(idx(A[1]), "k - q") << ", acc); " << c << " = hyx_sub_sum(acc); }\n";
            break;
        }
        case HYX_OP_DIV: {
            const std::string a = scalar(A[0]) ? "0.0" : vk(A[0], "k");
            if (scalar(A[1]))
                o << "        " << c << " = " << a << " * rb" << j << ";\n";
            else
                o << "        { double acc = 0.0; " << loop(1, "q <= k") << " acc = __builtin_fma(" << T(idx(A[1]), "q")
                  << ", " << T(j, "k - q") << ", acc); acc = hyx_sub_sum(acc); " << c << " = (" << a << " - acc) * rb" << j
                  << "; }\n";
            break;
        }
        case HYX_OP_SQUARE: {
            const uint32_t a = idx(A[0]);
            o << "        { double acc = 0.0; " << loop(0, "2 * q < k") << " acc = __builtin_fma(" << T(a, "q") << ", "
              << T(a, "k - q") << ", acc); acc = hyx_sub_sum(acc); acc = acc + acc; if ((k & 1) == 0) { const double m = "
              << T(a, "k >> 1")
              << "; acc = __builtin_fma(m, m, acc); } " << c << " = acc; }\n";
            break;
        }
        case HYX_OP_SQRT: {
            const std::string a = scalar(A[0]) ? "0.0" : vk(A[0], "k");
            o << "        { double acc = 0.0; " << loop(1, "2 * q < k") << " acc = __builtin_fma(" << T(j, "q") << ", "
              << T(j, "k - q") << ", acc); acc = hyx_sub_sum(acc); acc = acc + acc; if ((k & 1) == 0) { const double m = "
              << T(j, "k >> 1")
              << "; acc = __builtin_fma(m, m, acc); } " << c << " = (" << a << " - acc) * rs" << j << "; }\n";
            break;
        }
        case HYX_OP_POW: {
            const double al = S.consts[S.aux[j]];
            const uint32_t a = idx(A[0]);
            o << "        { double acc = 0.0; const double ka = (double)k * " << lit(al)
              << "; " << loop(0, "q < k") << " acc = __builtin_fma(__builtin_fma(-(double)q, " << lit(al + 1.0)
              << ", ka) * " << T(a, "k - q") << ", " << T(j, "q") << ", acc); " << c << " = hyx_sub_sum(acc) * (ra" << j
              << " * hyx_inv[k]); }\n";
            break;
        }
        case HYX_OP_EXP: {
            const uint32_t a = idx(A[0]);
            o << "        { double acc = 0.0; " << loop(1, "q <= k") << " acc = __builtin_fma((double)q * " << T(a, "q")
              << ", " << T(j, "k - q") << ", acc); " << c << " = hyx_sub_sum(acc) * hyx_inv[k]; }\n";
            break;
        }
        case HYX_OP_LOG: {
            const uint32_t a = idx(A[0]);
            o << "        { double acc = 0.0; " << loop(1, "q < k") << " acc = __builtin_fma((double)q * " << T(j, "q")
              << ", " << T(a, "k - q") << ", acc); acc = hyx_sub_sum(acc); " << c << " = (" << T(a, "k")
              << " - acc * hyx_inv[k]) * ra" << j << "; }\n";
            break;
        }
        case HYX_OP_SIN:
        case HYX_OP_COS: {
            const uint32_t a = idx(A[0]), p = S.aux[j];
            o << "        { double acc = 0.0; " << loop(1, "q <= k") << " acc = __builtin_fma((double)q * " << T(a, "q")
              << ", " << T(p, "k - q") << ", acc); " << c << " = hyx_sub_sum(acc) * " << (S.op[j] == HYX_OP_SIN ? "" : "-")
              << "hyx_inv[k]; }\n";
            break;
        }
        case HYX_OP_TIME: o << "        " << c << " = (k == 1) ? 1.0 : 0.0;\n"; break;
        }
        o << guard_close;
    }

    std::string run()
    {
        const uint32_t n = S.n;
        o << "#define HYX_N " << n << "\n";
        o << "#define HYX_NPAR " << std::max<uint32_t>(S.npar, 1) << "\n";
        o << "#define HYX_P " << P << "\n";
        o << "#define HYX_COMPACT 1\n";
        o << "#define HYX_SUB " << sub << "\n";
        if (sub > 1) {
            o << "__device__ __forceinline__ double hyx_sub_sum(double v)\n{\n"
                 "#pragma unroll\n    for (int off = 1; off < HYX_SUB; off <<= 1) v += __shfl_xor(v, off);\n"
                 "    return v;\n}\n";
        } else {
            o << "#define hyx_sub_sum(v) (v)\n";
        }
        o << "#define HYX_TAPE_ROWS " << (S.nnodes * (uint64_t)(P + 1)) << "ull\n";
        o << "typedef unsigned long long hyx_u64;\n";
        o << "__constant__ double hyx_inv[" << (P + 1) << "] = {0.0";
        for (int k = 1; k <= P; ++k) o << ", " << lit(1.0 / k);
        o << "};\n";
        o << "#define T(node, kk) tape[((hyx_u64)(node) * (HYX_P + 1) + (hyx_u64)(kk)) * tps]\n";
        o << "template <bool WTC>\n"
             "__device__ __forceinline__ void hyx_taylor_step(double (&x)[HYX_N], const double (&pv)[HYX_NPAR], "
             "double t0, double max_dt, double &h_out, bool &nf_out, double *__restrict__ tc, hyx_u64 tcs, "
             "double *__restrict__ tape, hyx_u64 tps"
          << (events ? ", unsigned long long &hyx_evm" : "") << ")\n{\n";
        o << "    (void)pv; (void)t0; (void)tc; (void)tcs;\n";
        // With events every step keeps its coefficients and forms the event functions'
        // from them (hyx_ev_norms); the root scan decides per event.
        if (events) o << "    hyx_evm = ~0ull;\n";
        if (sub > 1)
            o << "    const int hyx_sub = (int)(__lane_id() & (HYX_SUB - 1));\n";
        else
            o << "    const int hyx_sub = 0;\n";
        for (uint32_t i = 0; i < n; ++i) o << "    " << T(i, "0") << " = x[" << i << "];\n";
        for (uint32_t j = n; j < S.nnodes; ++j) order0(j);
        o << "    for (int k = 1; k <= HYX_P; ++k) {\n";
        for (uint32_t i = 0; i < n; ++i) {
            const uint32_t r = S.rhs[i];
            if (HYX_OPERAND_KIND(r) == HYX_KIND_NODE)
                o << "        " << T(i, "k") << " = " << T(idx(r), "k - 1") << " * hyx_inv[k];\n";
            else
                o << "        " << T(i, "k") << " = (k == 1) ? " << v0(r) << " : 0.0;\n";
        }
        o << "        if (k == HYX_P) break;\n";
        for (uint32_t j = n; j < S.nnodes; ++j) orderk(j);
        o << "    }\n";
        o << (events ? "    {\n" : "    if constexpr (WTC) {\n");
        o << "        for (int i = 0; i < HYX_N; ++i)\n"
             "            for (int k = 0; k <= HYX_P; ++k) tc[(hyx_u64)(i * (HYX_P + 1) + k) * tcs] = T(i, k);\n";
        o << "    }\n";
        o << "    double mx = 0.0, mo = 0.0, mom1 = 0.0;\n";
        o << "    for (int i = 0; i < HYX_N; ++i) {\n"
             "        double ab = fabs(T(i, 0)); mx = (ab > mx) ? ab : mx;\n"
             "        ab = fabs(T(i, HYX_P)); mo = (ab > mo) ? ab : mo;\n"
             "        ab = fabs(T(i, HYX_P - 1)); mom1 = (ab > mom1) ? ab : mom1;\n"
             "    }\n";
        if (events) o << "    hyx_ev_norms(tc, tcs, pv, t0, mo, mom1);\n";
        o << "    const double num = (mx <= 1.0) ? 1.0 : mx;\n";
        o << "    const double rho_o = pow(num / mo, " << lit(1.0 / P) << ");\n";
        o << "    const double rho_om1 = pow(num / mom1, " << lit(1.0 / (P - 1)) << ");\n";
        o << "    const double rho_m = (rho_om1 < rho_o) ? rho_om1 : rho_o;\n";
        const double e1 = std::exp(1.0);
        const double rhofac = std::exp(-7.0 / 10.0 / (double)(P - 1)) / (e1 * e1);
        o << "    double h = rho_m * " << lit(rhofac) << ";\n";
        o << "    const double lim = fabs(max_dt);\n";
        o << "    h = (h < lim) ? h : lim;\n";
        o << "    if (max_dt < 0.0) h = -h;\n";
        o << "    bool nf = false;\n";
        o << "#pragma unroll\n    for (int i = 0; i < HYX_N; ++i) {\n";
        if (!high_acc) {
            o << "        double r = T(i, HYX_P);\n"
                 "        for (int k = HYX_P - 1; k >= 0; --k) r = __builtin_fma(r, h, T(i, k));\n";
        } else {
            o << "        double r = T(i, 0), cp = 0.0, hk = 1.0;\n"
                 "        for (int k = 1; k <= HYX_P; ++k) { hk *= h; const double y = T(i, k) * hk - cp; "
                 "const double tt = r + y; cp = (tt - r) - y; r = tt; }\n";
        }
        o << "        x[i] = r; nf = nf || !isfinite(r);\n    }\n";
        o << "    h_out = h;\n    nf_out = nf;\n}\n#undef T\n";
        return o.str();
    }
};

} // namespace


// ---------------------------------------------------------------------------------------------
// Group form planning.
//
// The right-hand side of every state variable is flattened, through linear
// nodes (add, sub, neg, sum), into signed terms: state variables, numbers and
// "output" nodes. The remaining nodes are split into clusters, the connected
// components of the dependency graph restricted to them. If all clusters are
// isomorphic (same ops in the same order, same exponent constants, same output
// positions; inputs and other constants may differ), one template cluster is
// generated and each lane of a group evaluates one cluster with its own input
// and constant bindings.

namespace
{

struct UF {
    std::vector<uint32_t> p;
    explicit UF(uint32_t n) : p(n)
    {
        for (uint32_t i = 0; i < n; ++i) p[i] = i;
    }
    uint32_t find(uint32_t x)
    {
        while (p[x] != x) x = p[x] = p[p[x]];
        return x;
    }
    void join(uint32_t a, uint32_t b) { p[find(a)] = find(b); }
};

bool is_linear(uint32_t op) { return op == HYX_OP_ADD || op == HYX_OP_SUB || op == HYX_OP_NEG || op == HYX_OP_SUM; }

} // namespace

bool plan_groups(const Sys &S, GroupPlan &gp, std::string &why)
{
    const uint32_t n = S.n, NN = S.nnodes;
    if (NN == n) {
        why = "no nodes";
        return false;
    }
    std::vector<char> internal(NN, 0);
    std::vector<std::vector<std::pair<uint32_t, double>>> terms(n);
    std::vector<double> rconst(n, 0.0);
    for (uint32_t i = 0; i < n; ++i) {
        std::vector<std::pair<uint32_t, double>> stack{{S.rhs[i], 1.0}};
        while (!stack.empty()) {
            auto [w, sg] = stack.back();
            stack.pop_back();
            const uint32_t kind = HYX_OPERAND_KIND(w), idx = HYX_OPERAND_INDEX(w);
            if (kind == HYX_KIND_CONST) {
                rconst[i] += sg * S.consts[idx];
                continue;
            }
            if (kind == HYX_KIND_PAR) {
                why = "a runtime parameter appears as a right-hand-side term";
                return false;
            }
            if (idx < n || !is_linear(S.op[idx]) || S.deg[idx] == 0) {
                terms[i].push_back({w, sg});
                continue;
            }
            internal[idx] = 1;
            const auto &A = S.args[idx];
            for (size_t q = A.size(); q-- > 0;) {
                const bool neg = S.op[idx] == HYX_OP_NEG || (S.op[idx] == HYX_OP_SUB && q == 1);
                stack.push_back({A[q], neg ? -sg : sg});
            }
        }
    }
    for (uint32_t j = n; j < NN; ++j) {
        if (internal[j]) continue;
        for (auto a : S.args[j])
            if (HYX_OPERAND_KIND(a) == HYX_KIND_NODE && internal[HYX_OPERAND_INDEX(a)]) {
                why = "a linear right-hand-side node is also used inside a cluster";
                return false;
            }
    }
    UF uf(NN);
    for (uint32_t j = n; j < NN; ++j) {
        if (internal[j]) continue;
        for (auto a : S.args[j]) {
            const uint32_t idx = HYX_OPERAND_INDEX(a);
            if (HYX_OPERAND_KIND(a) == HYX_KIND_NODE && idx >= n) uf.join(j, idx);
        }
        if (S.op[j] == HYX_OP_SIN || S.op[j] == HYX_OP_COS) uf.join(j, S.aux[j]);
    }
    // Clusters, ordered by their first node.
    std::vector<int> cid(NN, -1);
    std::vector<std::vector<uint32_t>> cl;
    for (uint32_t j = n; j < NN; ++j) {
        if (internal[j]) continue;
        const uint32_t r = uf.find(j);
        if (cid[r] < 0) {
            cid[r] = (int)cl.size();
            cl.push_back({});
        }
        cid[j] = cid[r];
        cl[cid[j]].push_back(j);
    }
    const int ncl = (int)cl.size();
    if (ncl < 2) {
        why = "fewer than two clusters";
        return false;
    }
    // Output nodes: node terms of any right-hand side.
    std::vector<char> is_out(NN, 0);
    for (uint32_t i = 0; i < n; ++i)
        for (auto &t : terms[i]) {
            const uint32_t idx = HYX_OPERAND_INDEX(t.first);
            if (HYX_OPERAND_KIND(t.first) == HYX_KIND_NODE && idx >= n) is_out[idx] = 1;
        }
    // Signature and bindings of each cluster.
    struct Bind {
        std::vector<uint32_t> inputs;   // state variables by input slot
        std::vector<uint32_t> kconsts;  // const indices by constant slot
        std::vector<uint32_t> outs;     // output nodes by output slot
        std::string sig;
    };
    std::vector<Bind> B(ncl);
    for (int c = 0; c < ncl; ++c) {
        auto &b = B[c];
        std::vector<int> pos(NN, -1);
        for (size_t q = 0; q < cl[c].size(); ++q) pos[cl[c][q]] = (int)q;
        std::ostringstream sg;
        for (uint32_t j : cl[c]) {
            sg << S.op[j] << ":";
            for (auto a : S.args[j]) {
                const uint32_t kind = HYX_OPERAND_KIND(a), idx = HYX_OPERAND_INDEX(a);
                if (kind == HYX_KIND_NODE && idx >= n) {
                    sg << "L" << pos[idx];
                } else if (kind == HYX_KIND_NODE) {
                    size_t s2 = std::find(b.inputs.begin(), b.inputs.end(), idx) - b.inputs.begin();
                    if (s2 == b.inputs.size()) b.inputs.push_back(idx);
                    sg << "I" << s2;
                } else if (kind == HYX_KIND_CONST) {
                    size_t s2 = std::find(b.kconsts.begin(), b.kconsts.end(), idx) - b.kconsts.begin();
                    if (s2 == b.kconsts.size()) b.kconsts.push_back(idx);
                    sg << "K" << s2;
                } else {
                    sg << "P" << idx;
                }
                sg << ",";
            }
            if (S.op[j] == HYX_OP_POW) sg << "E" << lit(S.consts[S.aux[j]]);
            if (S.op[j] == HYX_OP_SIN || S.op[j] == HYX_OP_COS) sg << "S" << pos[S.aux[j]];
            sg << "|";
        }
        sg << "O";
        for (uint32_t j : cl[c])
            if (is_out[j]) {
                b.outs.push_back(j);
                sg << pos[j] << ",";
            }
        b.sig = sg.str();
    }
    for (int c = 1; c < ncl; ++c)
        if (B[c].sig != B[0].sig) {
            why = "the clusters are not isomorphic";
            return false;
        }
    int G = 1;
    while (G < ncl) G <<= 1;
    if (G > 64) {
        why = "more than 64 clusters";
        return false;
    }
    const int own = (int)((n + G - 1) / G);
    if (own > 8) {
        why = "too many state variables per lane";
        return false;
    }
    gp = GroupPlan{};
    gp.G = G;
    gp.ncl = ncl;
    gp.nin = (int)B[0].inputs.size();
    gp.nout = (int)B[0].outs.size();
    gp.own = own;
    // Template: cluster 0 with inputs as state variables 0..nin-1.
    {
        Sys &T = gp.tmpl;
        const auto &c0 = cl[0];
        const uint32_t nin = gp.nin;
        T.n = nin;
        T.npar = S.npar;
        T.nnodes = nin + (uint32_t)c0.size();
        T.op.assign(T.nnodes, 0);
        T.aux.assign(T.nnodes, 0);
        T.args.assign(T.nnodes, {});
        std::vector<int> pos(NN, -1);
        for (size_t q = 0; q < c0.size(); ++q) pos[c0[q]] = (int)q;
        // Constants: per-lane slots first, then the pow exponents.
        T.consts.clear();
        for (size_t q = 0; q < B[0].kconsts.size(); ++q) T.consts.push_back(S.consts[B[0].kconsts[q]]);
        gp.perlane.assign(T.consts.size(), 1);
        for (size_t q = 0; q < c0.size(); ++q) {
            const uint32_t j = c0[q], tj = nin + (uint32_t)q;
            T.op[tj] = S.op[j];
            for (auto a : S.args[j]) {
                const uint32_t kind = HYX_OPERAND_KIND(a), idx = HYX_OPERAND_INDEX(a);
                if (kind == HYX_KIND_NODE && idx >= n)
                    T.args[tj].push_back(HYX_OPERAND(HYX_KIND_NODE, nin + pos[idx]));
                else if (kind == HYX_KIND_NODE)
                    T.args[tj].push_back(HYX_OPERAND(
                        HYX_KIND_NODE, std::find(B[0].inputs.begin(), B[0].inputs.end(), idx) - B[0].inputs.begin()));
                else if (kind == HYX_KIND_CONST)
                    T.args[tj].push_back(HYX_OPERAND(
                        HYX_KIND_CONST,
                        std::find(B[0].kconsts.begin(), B[0].kconsts.end(), idx) - B[0].kconsts.begin()));
                else
                    T.args[tj].push_back(a);
            }
            if (S.op[j] == HYX_OP_POW) {
                T.aux[tj] = (uint32_t)T.consts.size();
                T.consts.push_back(S.consts[S.aux[j]]);
                gp.perlane.push_back(0);
            }
            if (S.op[j] == HYX_OP_SIN || S.op[j] == HYX_OP_COS) T.aux[tj] = nin + pos[S.aux[j]];
        }
        T.nconsts = (uint32_t)T.consts.size();
        compute_degrees(T);
        for (uint32_t j : B[0].outs) gp.outs.push_back(nin + pos[j]);
    }
    gp.nkc = std::max<int>(1, (int)gp.tmpl.nconsts);
    // Per-lane bindings.
    gp.iv.assign((size_t)G * std::max(gp.nin, 1), 0);
    gp.kc.assign((size_t)G * gp.nkc, 0.0);
    for (int c = 0; c < ncl; ++c) {
        for (int q = 0; q < gp.nin; ++q) gp.iv[(size_t)c * std::max(gp.nin, 1) + q] = (int)B[c].inputs[q];
        for (size_t q = 0; q < B[c].kconsts.size(); ++q) gp.kc[(size_t)c * gp.nkc + q] = S.consts[B[c].kconsts[q]];
        for (size_t q = B[c].kconsts.size(); q < gp.tmpl.nconsts; ++q) gp.kc[(size_t)c * gp.nkc + q] = gp.tmpl.consts[q];
    }
    // Output q of lane c lives at LDS index q * (G + 1) + c: the lanes of a group write
    // one output row with consecutive 8-byte words, and the odd row stride puts the
    // outputs of one cluster in different banks, so owners summing several outputs of
    // the same cluster in one load do not collide. Index (G + 1) * nout reads zero.
    const int OS = G + 1;
    gp.os = OS;
    std::vector<int> outslot(NN, -1);
    for (int c = 0; c < ncl; ++c)
        for (size_t q = 0; q < B[c].outs.size(); ++q) outslot[B[c].outs[q]] = (int)q * OS + c;
    // Ownership: variables sorted by (output terms, state terms), dealt to lanes slot
    // by slot, so that each slot's term counts are as uniform as possible.
    std::vector<uint32_t> order(n);
    std::vector<int> n_out(n, 0), n_st(n, 0);
    for (uint32_t v = 0; v < n; ++v) {
        order[v] = v;
        for (auto &t : terms[v]) (HYX_OPERAND_INDEX(t.first) < n ? n_st[v] : n_out[v])++;
    }
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
        return std::make_pair(n_out[a], n_st[a]) > std::make_pair(n_out[b], n_st[b]);
    });
    gp.vo.assign((size_t)G * own, -1);
    gp.rc.assign((size_t)G * own, 0.0);
    gp.nto.assign(own, 0);
    gp.nts.assign(own, 0);
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t v = order[i];
        const int g = (int)(i % G), o = (int)(i / G);
        gp.vo[(size_t)g * own + o] = (int)v;
        gp.rc[(size_t)g * own + o] = rconst[v];
        gp.nto[o] = std::max(gp.nto[o], n_out[v]);
        gp.nts[o] = std::max(gp.nts[o], n_st[v]);
    }
    gp.tio.assign(own, {});
    gp.tis.assign(own, {});
    gp.tso.assign(own, {});
    gp.tss.assign(own, {});
    for (int o = 0; o < own; ++o) {
        gp.tio[o].assign((size_t)G * gp.nto[o], OS * gp.nout);
        gp.tso[o].assign((size_t)G * gp.nto[o], 1.0);
        gp.tis[o].assign((size_t)G * gp.nts[o], (int)n);
        gp.tss[o].assign((size_t)G * gp.nts[o], 1.0);
    }
    for (int g = 0; g < G; ++g)
        for (int o = 0; o < own; ++o) {
            const int v = gp.vo[(size_t)g * own + o];
            if (v < 0) continue;
            int a = 0, b = 0;
            for (auto &t : terms[v]) {
                const uint32_t idx = HYX_OPERAND_INDEX(t.first);
                if (idx < n) {
                    gp.tis[o][(size_t)g * gp.nts[o] + b] = (int)idx;
                    gp.tss[o][(size_t)g * gp.nts[o] + b] = t.second;
                    ++b;
                } else {
                    if (outslot[idx] < 0) {
                        why = "a right-hand-side term is not a cluster output";
                        return false;
                    }
                    gp.tio[o][(size_t)g * gp.nto[o] + a] = outslot[idx];
                    gp.tso[o][(size_t)g * gp.nto[o] + a] = t.second;
                    ++a;
                }
            }
        }
    return true;
}

// LDS column of each state variable in the group form's x rows (column n is the zero
// column). A double at LDS word address a occupies the 4-byte banks 2a and 2a + 1 of 32,
// and a 64-bit read is served 16 lanes (one group) per pass: lanes whose addresses differ
// but agree mod 16 serialise. Every column read of the step is the same table column for
// the 16 lanes of a group (reads[r], with its count per step), so the map is chosen to
// spread those columns over the 16 double banks, by a deterministic local search (pair
// swaps kept when they do not raise the serialised passes). Measured on the 6-body group
// kernel (65,536 instances, one dt = 25 slice), with the output slots of lds_output_map:
// SQ_LDS_BANK_CONFLICT 1.50e10 -> 7.33e9, the same instructions, 147.0 -> 144.0 ms.
static std::vector<int> lds_column_map(int n, int G, const std::vector<std::pair<std::vector<int>, int>> &reads)
{
    std::vector<int> col((size_t)n + 1);
    for (int v = 0; v <= n; ++v) col[(size_t)v] = v;
    if (G != 16 || reads.empty()) return col;
    auto cost = [&](const std::vector<int> &c) {
        long tot = 0;
        for (const auto &r : reads) {
            int cnt[16] = {0};
            int seen[16][16];
            int worst = 1;
            for (int v : r.first) {
                const int a = c[(size_t)v], b = a & 15;
                bool dup = false;
                for (int j = 0; j < cnt[b]; ++j) dup = dup || seen[b][j] == a;
                if (!dup) {
                    seen[b][cnt[b]++] = a;
                    worst = std::max(worst, cnt[b]);
                }
            }
            tot += (long)(worst - 1) * r.second;
        }
        return tot;
    };
    long best = cost(col);
    uint64_t s = 0x9e3779b97f4a7c15ull;
    for (int it = 0; it < 20000 && best > 0; ++it) {
        s = s * 6364136223846793005ull + 1442695040888963407ull;
        const int i = (int)((s >> 33) % (uint64_t)(n + 1));
        s = s * 6364136223846793005ull + 1442695040888963407ull;
        const int j = (int)((s >> 33) % (uint64_t)(n + 1));
        if (i == j) continue;
        std::swap(col[(size_t)i], col[(size_t)j]);
        const long c = cost(col);
        if (c <= best)
            best = c;
        else
            std::swap(col[(size_t)i], col[(size_t)j]);
    }
    return col;
}

// Slots of the cluster outputs in the group form's LDS block: output q of lane c at
// row[q] + perm[c], the zero slot at zero. One lane permutation for every row keeps a
// lane's output addresses at constant distances (the compiler pairs its writes into
// ds_write2_b64); the rows start at chosen residues mod 16, 16 doubles apart or more. The
// owners' loads read one output each per lane (reads[r]: the 16 addresses of the
// identity layout q * OS + c, OS * nout the zero slot, with the count per step); perm and
// the row residues are chosen, like the column map, to spread each load over the 16
// double banks. With G != 16 the identity layout is kept.
struct OutputSlots {
    std::vector<int> perm, row;
    int zero = 0;
    int map(int a, int OS, int nout) const
    {
        if (a >= OS * nout) return zero;
        return row[(size_t)(a / OS)] + perm[(size_t)(a % OS)];
    }
};

static void place_rows(OutputSlots &m, const std::vector<int> &res, int G)
{
    int next = 0;
    for (size_t q = 0; q < res.size(); ++q) {
        int r = next + ((res[q] - next) % 16 + 16) % 16;
        m.row[q] = r;
        next = r + G;
    }
    m.zero = next;
}

static OutputSlots lds_output_map(int G, int OS, int nout, const std::vector<std::pair<std::vector<int>, int>> &reads)
{
    OutputSlots m;
    m.perm.resize((size_t)G);
    m.row.resize((size_t)nout);
    for (int c = 0; c < G; ++c) m.perm[(size_t)c] = c;
    for (int q = 0; q < nout; ++q) m.row[(size_t)q] = q * OS;
    m.zero = OS * nout;
    if (G != 16 || nout == 0 || reads.empty()) return m;
    std::vector<int> res((size_t)nout);
    for (int q = 0; q < nout; ++q) res[(size_t)q] = (q * OS) % 16;
    auto cost = [&](const OutputSlots &t) {
        long tot = 0;
        for (const auto &r : reads) {
            int cnt[16] = {0};
            int seen[16][16];
            int worst = 1;
            for (int a0 : r.first) {
                const int a = t.map(a0, OS, nout), b = a & 15;
                bool dup = false;
                for (int j = 0; j < cnt[b]; ++j) dup = dup || seen[b][j] == a;
                if (!dup) {
                    seen[b][cnt[b]++] = a;
                    worst = std::max(worst, cnt[b]);
                }
            }
            tot += (long)(worst - 1) * r.second;
        }
        return tot;
    };
    OutputSlots cur = m;
    place_rows(cur, res, G);
    const long base = cost(m);
    long best = cost(cur);
    uint64_t s = 0x243f6a8885a308d3ull;
    auto rnd = [&](int k) {
        s = s * 6364136223846793005ull + 1442695040888963407ull;
        return (int)((s >> 33) % (uint64_t)k);
    };
    for (int it = 0; it < 60000 && best > 0; ++it) {
        if (rnd(2) == 0) {
            const int i = rnd(G), j = rnd(G);
            if (i == j) continue;
            std::swap(cur.perm[(size_t)i], cur.perm[(size_t)j]);
            const long c = cost(cur);
            if (c <= best)
                best = c;
            else
                std::swap(cur.perm[(size_t)i], cur.perm[(size_t)j]);
        } else {
            const int q = rnd(nout), old = res[(size_t)q];
            res[(size_t)q] = rnd(16);
            place_rows(cur, res, G);
            const long c = cost(cur);
            if (c <= best) {
                best = c;
            } else {
                res[(size_t)q] = old;
                place_rows(cur, res, G);
            }
        }
    }
    return best < base ? cur : m;
}

std::string generate_group_prelude(const Sys &S, const GroupPlan &gp, int P, bool high_acc, int split_min,
                                   bool pipeline, const Sys *E, uint32_t nev)
{
    const int G = gp.G, own = gp.own, nin = gp.nin, nout = gp.nout;
    const int n = (int)S.n, nx = n + 1;
    const int OS = gp.os; // output row stride of the identity layout (the tio tables)
    std::vector<std::pair<std::vector<int>, int>> oreads;
    for (int o = 0; o < own; ++o)
        for (int t = 0; t < gp.nto[o]; ++t) {
            std::vector<int> as;
            for (int g = 0; g < G; ++g) as.push_back(gp.tio[o][(size_t)g * gp.nto[o] + t]);
            oreads.push_back({as, P});
        }
    const OutputSlots oslot = lds_output_map(G, OS, nout, oreads);
    const int xb = std::max(oslot.zero, OS) + 1; // LDS: outputs, zero slot, then x[P+1][n+1]
    const int flat = xb + (P + 1) * nx;
    std::ostringstream h;
    h << "#define HYX_N " << n << "\n#define HYX_NPAR " << std::max<uint32_t>(S.npar, 1) << "\n#define HYX_P " << P
      << "\n#define HYX_GROUP 1\n#define HYX_G " << G << "\n#define HYX_I " << (64 / G) << "\n#define HYX_NIN "
      << std::max(nin, 1) << "\n#define HYX_NOUT " << std::max(nout, 1) << "\n#define HYX_NKC " << gp.nkc
      << "\n#define HYX_OWN " << own << "\n#define HYX_NX " << nx << "\n#define HYX_XB " << xb << "\n#define HYX_FLAT "
      << flat << "\n";
    h << "typedef unsigned long long hyx_u64;\n";
    h << "#ifndef HYX_CPU_EMU\n"
         "// A block is one wavefront, and the LDS operations of a wavefront complete in issue\n"
         "// order: the per-order exchanges only need the compiler to keep memory accesses on\n"
         "// their side of the fence (no s_waitcnt, no s_barrier).\n"
         "__device__ __forceinline__ void hyx_wsync()\n{\n"
         "    __builtin_amdgcn_fence(__ATOMIC_RELEASE, \"wavefront\");\n"
         "    __builtin_amdgcn_wave_barrier();\n"
         "    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, \"wavefront\");\n}\n"
         "#define HYX_EMU_SYNC() hyx_wsync()\n"
         "#endif\n";
    auto itab = [&](const std::string &nm, const std::vector<int> &v) {
        h << "__constant__ int " << nm << "[" << std::max<size_t>(v.size(), 1) << "] = {";
        for (size_t q = 0; q < v.size(); ++q) h << (q ? "," : "") << v[q];
        if (v.empty()) h << "0";
        h << "};\n";
    };
    auto dtab = [&](const std::string &nm, const std::vector<double> &v) {
        h << "__constant__ double " << nm << "[" << std::max<size_t>(v.size(), 1) << "] = {";
        for (size_t q = 0; q < v.size(); ++q) h << (q ? "," : "") << lit(v[q]);
        if (v.empty()) h << "0.0";
        h << "};\n";
    };
    // Pipelined schedule applies when every cluster input is a "lagged" state variable:
    // its right-hand side has no cluster-output term. Lane tables then give, per input,
    // the state terms (variable, scale) of that right-hand side, in the owner's order.
    std::vector<std::pair<int, int>> owner(n, {-1, -1});
    for (int g = 0; g < G; ++g)
        for (int oo = 0; oo < own; ++oo) {
            const int v = gp.vo[(size_t)g * own + oo];
            if (v >= 0 && v < n) owner[v] = {g, oo};
        }
    bool lagged = pipeline && nin > 0 && nout > 0;
    int nls = 0;
    for (int g = 0; g < G && lagged; ++g)
        for (int q = 0; q < nin && lagged; ++q) {
            const int v = gp.iv[(size_t)g * nin + q];
            if (v < 0 || v >= n || owner[v].first < 0) {
                lagged = false;
                break;
            }
            const int g2 = owner[v].first, oo = owner[v].second;
            for (int t = 0; t < gp.nto[oo]; ++t)
                if (gp.tio[oo][(size_t)g2 * gp.nto[oo] + t] != OS * nout) lagged = false;
            nls = std::max(nls, gp.nts[oo]);
        }
    if (nls == 0) lagged = false;
    std::vector<int> lis;
    std::vector<double> lss, lrc;
    std::vector<std::string> lss_sign;
    bool lrc_nonzero = false, lss_table = false;
    if (lagged) {
        lis.assign((size_t)G * nin * nls, n);
        lss.assign((size_t)G * nin * nls, 1.0);
        lrc.assign((size_t)G * nin, 0.0);
        for (int g = 0; g < G; ++g)
            for (int q = 0; q < nin; ++q) {
                const int v = gp.iv[(size_t)g * nin + q];
                const int g2 = owner[v].first, oo = owner[v].second;
                for (int t = 0; t < gp.nts[oo]; ++t) {
                    lis[((size_t)g * nin + q) * nls + t] = gp.tis[oo][(size_t)g2 * gp.nts[oo] + t];
                    lss[((size_t)g * nin + q) * nls + t] = gp.tss[oo][(size_t)g2 * gp.nts[oo] + t];
                }
                lrc[(size_t)g * nin + q] = gp.rc[(size_t)g2 * own + oo];
                lrc_nonzero = lrc_nonzero || lrc[(size_t)g * nin + q] != 0.0;
            }
        for (int qt = 0; qt < nin * nls; ++qt) {
            bool same = true;
            for (int g = 1; g < G; ++g) same = same && lss[(size_t)g * nin * nls + qt] == lss[qt];
            const double v = lss[qt];
            lss_sign.push_back(!same ? std::string() : v == 1.0 ? std::string("+") : v == -1.0 ? std::string("-") : lit(v));
            lss_table = lss_table || !same;
        }
    }
    h << "#define HYX_PIPELINED " << (lagged ? 1 : 0) << "\n";
    // Column reads of one step, per group: the owners' rows (initial write, Horner, the
    // coefficient log), the cluster inputs (lagged: their state terms), the owners' state
    // terms.
    std::vector<std::pair<std::vector<int>, int>> creads;
    auto add_reads = [&](const std::vector<int> &tabv, int cnt, int weight) {
        for (int q = 0; q < cnt; ++q) {
            std::vector<int> vs;
            for (int g = 0; g < G; ++g) {
                const int v = tabv[(size_t)g * cnt + q];
                vs.push_back(v >= 0 && v < n ? v : n);
            }
            creads.push_back({vs, weight});
        }
    };
    add_reads(gp.vo, own, P + 3);
    if (lagged)
        add_reads(lis, nin * nls, P);
    else
        add_reads(gp.iv, nin, P + 1);
    for (int o = 0; o < own; ++o) add_reads(gp.tis[o], gp.nts[o], P + 1);
    const std::vector<int> col = lds_column_map(n, G, creads);
    auto cols = [&](const std::vector<int> &tabv) {
        std::vector<int> r(tabv.size());
        for (size_t q = 0; q < tabv.size(); ++q) r[q] = col[(size_t)(tabv[q] >= 0 && tabv[q] < n ? tabv[q] : n)];
        return r;
    };
    itab("hyx_tab_iv", cols(gp.iv));
    auto oslots = [&](const std::vector<int> &tabv) {
        std::vector<int> r(tabv.size());
        for (size_t q = 0; q < tabv.size(); ++q) r[q] = oslot.map(tabv[q], OS, nout);
        return r;
    };
    itab("hyx_tab_wp", oslot.perm);
    dtab("hyx_tab_kc", gp.kc);
    if (lagged) {
        itab("hyx_tab_lis", cols(lis));
        if (lss_table) dtab("hyx_tab_lss", lss);
        if (lrc_nonzero) dtab("hyx_tab_lrc", lrc);
    }
    itab("hyx_tab_vo", gp.vo);
    itab("hyx_tab_vc", cols(gp.vo));
    dtab("hyx_tab_rc", gp.rc);
    // Per-slot term tables; a sign uniform over the lanes becomes a literal.
    std::vector<std::vector<std::string>> sgn_o(own), sgn_s(own);
    for (int o = 0; o < own; ++o) {
        itab("hyx_tab_tio" + std::to_string(o), oslots(gp.tio[o]));
        itab("hyx_tab_tis" + std::to_string(o), cols(gp.tis[o]));
        dtab("hyx_tab_tso" + std::to_string(o), gp.tso[o]);
        dtab("hyx_tab_tss" + std::to_string(o), gp.tss[o]);
        auto uni = [&](const std::vector<double> &v, int cnt, int t, const std::string &arr) {
            bool same = true;
            for (int g = 1; g < G; ++g) same = same && v[(size_t)g * cnt + t] == v[t];
            return same ? (v[t] == 1.0 ? std::string("+") : v[t] == -1.0 ? std::string("-") : lit(v[t]))
                        : "tab." + arr + "[" + std::to_string(t) + "]";
        };
        for (int t = 0; t < gp.nto[o]; ++t) sgn_o[o].push_back(uni(gp.tso[o], gp.nto[o], t, "tso" + std::to_string(o)));
        for (int t = 0; t < gp.nts[o]; ++t) sgn_s[o].push_back(uni(gp.tss[o], gp.nts[o], t, "tss" + std::to_string(o)));
    }
    // The per-lane table struct and its loader.
    h << "struct hyx_gtab {\n    int iv[HYX_NIN];\n    double kc[HYX_NKC];\n    int vo[HYX_OWN];\n    int vc[HYX_OWN];\n"
         "    double rc[HYX_OWN];\n    int wp;\n";
    if (lagged) {
        h << "    int lis[" << nin * nls << "];\n";
        if (lss_table) h << "    double lss[" << nin * nls << "];\n";
        if (lrc_nonzero) h << "    double lrc[HYX_NIN];\n";
    }
    for (int o = 0; o < own; ++o)
        h << "    int tio" << o << "[" << std::max(gp.nto[o], 1) << "];\n    int tis" << o << "["
          << std::max(gp.nts[o], 1) << "];\n    double tso" << o << "[" << std::max(gp.nto[o], 1) << "];\n    double tss"
          << o << "[" << std::max(gp.nts[o], 1) << "];\n";
    h << "};\n__device__ __forceinline__ void hyx_gtab_load(hyx_gtab &t, int g)\n{\n";
    h << "    for (int q = 0; q < HYX_NIN; ++q) t.iv[q] = hyx_tab_iv[g * HYX_NIN + q];\n";
    h << "    t.wp = hyx_tab_wp[g];\n";
    h << "    for (int q = 0; q < HYX_NKC; ++q) t.kc[q] = hyx_tab_kc[g * HYX_NKC + q];\n";
    h << "    for (int q = 0; q < HYX_OWN; ++q) { t.vo[q] = hyx_tab_vo[g * HYX_OWN + q]; t.vc[q] = hyx_tab_vc[g * HYX_OWN + q]; "
         "t.rc[q] = hyx_tab_rc[g * HYX_OWN + q]; }\n";
    if (lagged) {
        h << "    for (int q = 0; q < " << nin * nls << "; ++q) t.lis[q] = hyx_tab_lis[g * " << nin * nls << " + q];\n";
        if (lss_table) h << "    for (int q = 0; q < " << nin * nls << "; ++q) t.lss[q] = hyx_tab_lss[g * " << nin * nls << " + q];\n";
        if (lrc_nonzero) h << "    for (int q = 0; q < HYX_NIN; ++q) t.lrc[q] = hyx_tab_lrc[g * HYX_NIN + q];\n";
    }
    for (int o = 0; o < own; ++o) {
        if (gp.nto[o])
            h << "    for (int q = 0; q < " << gp.nto[o] << "; ++q) { t.tio" << o << "[q] = hyx_tab_tio" << o << "[g * "
              << gp.nto[o] << " + q]; t.tso" << o << "[q] = hyx_tab_tso" << o << "[g * " << gp.nto[o] << " + q]; }\n";
        if (gp.nts[o])
            h << "    for (int q = 0; q < " << gp.nts[o] << "; ++q) { t.tis" << o << "[q] = hyx_tab_tis" << o << "[g * "
              << gp.nts[o] << " + q]; t.tss" << o << "[q] = hyx_tab_tss" << o << "[g * " << gp.nts[o] << " + q]; }\n";
    }
    h << "}\n";

    Gen gen(gp.tmpl, P, high_acc);
    gen.split_min = split_min;
    gen.literal_pow = false;
    gen.perlane = gp.perlane;
    gen.kc_prefix = "tab.kc";
    auto &o = gen.o;
    auto xv = [](int oo, int k) { return "x" + std::to_string(oo) + "_" + std::to_string(k); };
    auto X = [&](int k, const std::string &col) { return "fl[" + std::to_string(xb + k * nx) + " + " + col + "]"; };
    auto vcol = [&](int oo) { return "vc" + std::to_string(oo); };
    o << "template <bool WTC, class TAB>\n"
         "__device__ __forceinline__ void hyx_gstep(double (&xo)[HYX_OWN], const double (&pv)[HYX_NPAR], double t0, "
         "double max_dt, double &h_out, bool &nf_out, int g, double *fl, const TAB &tab, double *__restrict__ tc, "
         "hyx_u64 tcs"
      << (E ? ", unsigned long long &hyx_evm, double &hyx_mx" : "") << ")\n{\n";
    o << "    (void)pv; (void)t0; (void)tc; (void)tcs; (void)g;\n";
    // Column of each owned variable in the x rows (column n reads zero for empty slots).
    for (int oo = 0; oo < own; ++oo)
        o << "    const int " << vcol(oo) << " = tab.vc[" << oo << "];\n";
    o << "    hyx_wsync();\n";
    for (int oo = 0; oo < own; ++oo) {
        o << "    const double " << xv(oo, 0) << " = xo[" << oo << "];\n";
        // Empty slots hold +0.0 and write it into the zero column: no branch needed.
        o << "    " << X(0, vcol(oo)) << " = " << xv(oo, 0) << ";\n";
    }
    o << "    hyx_wsync();\n";
    // Owner sum of slot oo for x^[k+1], from the term values t_o (outputs) and t_s (state).
    auto owner_expr = [&](int oo, int k, const std::vector<std::string> &t_o, const std::vector<std::string> &t_s) {
        std::string r = (k == 0) ? "tab.rc[" + std::to_string(oo) + "]" : "";
        auto add = [&](const std::string &sg, const std::string &val) {
            if (sg == "+")
                r = r.empty() ? val : "(" + r + " + " + val + ")";
            else if (sg == "-")
                r = r.empty() ? "(-" + val + ")" : "(" + r + " - " + val + ")";
            else
                r = r.empty() ? "(" + sg + " * " + val + ")" : "__builtin_fma(" + sg + ", " + val + ", " + r + ")";
        };
        for (int t = 0; t < gp.nto[oo]; ++t) add(sgn_o[oo][t], t_o[t]);
        for (int t = 0; t < gp.nts[oo]; ++t) add(sgn_s[oo][t], t_s[t]);
        if (r.empty()) r = "0.0";
        return r + " * " + lit(1.0 / (k + 1));
    };
    auto cluster_order = [&](int k) {
        for (uint32_t j = gp.tmpl.n; j < gp.tmpl.nnodes; ++j) {
            if (k == 0)
                gen.order0(j);
            else
                gen.orderk(j, k);
        }
        for (int q = 0; q < nout; ++q) {
            const uint32_t w = HYX_OPERAND(HYX_KIND_NODE, gp.outs[q]);
            o << "    fl[" << oslot.row[(size_t)q] << " + tab.wp] = " << (gen.zero(w, k) ? "0.0" : gen.val(w, k)) << ";\n";
        }
    };
    if (lagged) {
        // Pipelined schedule, one exchange per order: every cluster input is a state
        // variable whose right-hand side holds state terms only (N-body positions,
        // x' = v), so a lane forms its order-k inputs itself from the order k-1 rows
        // instead of waiting for the owners' order-k rows. In the phase of order k the
        // owners' loads (outputs of order k-1) and the cluster-input loads are issued
        // together, the owners write x^[k], then the clusters compute order k and
        // overwrite the outputs: the owners read them first, and one wavefront runs
        // the whole group, whose LDS accesses complete in issue order.
        o << "    // order 0\n";
        for (int q = 0; q < nin; ++q)
            o << "    const double " << Gen::cv((uint32_t)q, 0) << " = " << X(0, "tab.iv[" + std::to_string(q) + "]") << ";\n";
        cluster_order(0);
        o << "    hyx_wsync();\n";
        for (int k = 1; k <= P; ++k) {
            o << "    // owner loads of order " << k - 1 << (k < P ? ", cluster order " + std::to_string(k) : "") << "\n";
            std::vector<std::vector<std::string>> to(own), ts(own);
            for (int oo = 0; oo < own; ++oo) {
                for (int t = 0; t < gp.nto[oo]; ++t) {
                    const std::string nm = "ot" + std::to_string(oo) + "_" + std::to_string(t) + "_" + std::to_string(k);
                    o << "    const double " << nm << " = fl[tab.tio" << oo << "[" << t << "]];\n";
                    to[oo].push_back(nm);
                }
                for (int t = 0; t < gp.nts[oo]; ++t) {
                    const std::string nm = "os" + std::to_string(oo) + "_" + std::to_string(t) + "_" + std::to_string(k);
                    o << "    const double " << nm << " = " << X(k - 1, "tab.tis" + std::to_string(oo) + "[" + std::to_string(t) + "]")
                      << ";\n";
                    ts[oo].push_back(nm);
                }
            }
            if (k < P)
                for (int q = 0; q < nin; ++q) {
                    std::string r = (k == 1 && lrc_nonzero) ? "tab.lrc[" + std::to_string(q) + "]" : "";
                    for (int t = 0; t < nls; ++t) {
                        const std::string val = X(k - 1, "tab.lis[" + std::to_string(q * nls + t) + "]");
                        const std::string &sg = lss_sign[(size_t)q * nls + t];
                        const std::string sgv = sg.empty() ? "tab.lss[" + std::to_string(q * nls + t) + "]" : sg;
                        if (sgv == "+")
                            r = r.empty() ? val : "(" + r + " + " + val + ")";
                        else if (sgv == "-")
                            r = r.empty() ? "(-" + val + ")" : "(" + r + " - " + val + ")";
                        else
                            r = r.empty() ? "(" + sgv + " * " + val + ")" : "__builtin_fma(" + sgv + ", " + val + ", " + r + ")";
                    }
                    if (r.empty()) r = "0.0";
                    o << "    const double " << Gen::cv((uint32_t)q, k) << " = " << r << " * " << lit(1.0 / k) << ";\n";
                }
            // The owners finish order k before the cluster work, so no sum stays live across it.
            for (int oo = 0; oo < own; ++oo) {
                o << "    const double " << xv(oo, k) << " = " << owner_expr(oo, k - 1, to[oo], ts[oo]) << ";\n";
                o << "    " << X(k, vcol(oo)) << " = " << xv(oo, k) << ";\n";
            }
            if (k < P) {
                o << "    HYX_EMU_SYNC();\n";
                cluster_order(k);
            }
            o << "    hyx_wsync();\n";
        }
    } else {
        for (int k = 0; k < P; ++k) {
            o << "    // order " << k << "\n";
            for (int q = 0; q < nin; ++q)
                o << "    const double " << Gen::cv((uint32_t)q, k) << " = " << X(k, "tab.iv[" + std::to_string(q) + "]") << ";\n";
            cluster_order(k);
            o << "    hyx_wsync();\n";
            for (int oo = 0; oo < own; ++oo) {
                std::vector<std::string> t_o, t_s;
                for (int t = 0; t < gp.nto[oo]; ++t) t_o.push_back("fl[tab.tio" + std::to_string(oo) + "[" + std::to_string(t) + "]]");
                for (int t = 0; t < gp.nts[oo]; ++t)
                    t_s.push_back(X(k, "tab.tis" + std::to_string(oo) + "[" + std::to_string(t) + "]"));
                o << "    const double " << xv(oo, k + 1) << " = " << owner_expr(oo, k, t_o, t_s) << ";\n";
                o << "    " << X(k + 1, vcol(oo)) << " = " << xv(oo, k + 1) << ";\n";
            }
            o << "    hyx_wsync();\n";
        }
    }
    o << "    if constexpr (WTC) if (tc) {\n";
    for (int oo = 0; oo < own; ++oo) {
        o << "        if (tab.vo[" << oo << "] >= 0)\n";
        o << "            for (int k = 0; k <= HYX_P; ++k) tc[((hyx_u64)tab.vo[" << oo << "] * " << (P + 1)
          << " + k) * tcs] = fl[HYX_XB + k * HYX_NX + " << vcol(oo) << "];\n";
    }
    o << "    }\n";
    if (E) {
        // Event functions: every lane of the group forms the whole jet from the state
        // rows in LDS (broadcast reads), so the group agrees without an exchange.
        Gen eg(*E, P, high_acc);
        eg.lds_xb = xb;
        eg.lds_nx = nx;
        eg.lds_col = col;
        o << "    double hyx_evg[" << nev << "][" << (P + 1) << "];\n    {\n";
        for (uint32_t j = E->n; j < E->nnodes; ++j) eg.order0(j);
        for (int k = 1; k <= P; ++k)
            for (uint32_t j = E->n; j < E->nnodes; ++j) eg.orderk(j, k);
        for (uint32_t e = 0; e < nev; ++e)
            for (int k = 0; k <= P; ++k)
                eg.o << "    hyx_evg[" << e << "][" << k << "] = " << (eg.zero(E->rhs[e], k) ? "0.0" : eg.val(E->rhs[e], k))
                     << ";\n";
        o << eg.o.str() << "    }\n";
    }
    o << "    double mx = 0.0, mo = 0.0, mom1 = 0.0, ab;\n";
    for (int oo = 0; oo < own; ++oo) {
        o << "    ab = fabs(" << xv(oo, 0) << "); mx = (ab > mx) ? ab : mx;\n";
        o << "    ab = fabs(" << xv(oo, P) << "); mo = (ab > mo) ? ab : mo;\n";
        o << "    ab = fabs(" << xv(oo, P - 1) << "); mom1 = (ab > mom1) ? ab : mom1;\n";
    }
    for (uint32_t e = 0; E && e < nev; ++e) {
        o << "    ab = fabs(hyx_evg[" << e << "][" << P << "]); mo = (ab > mo) ? ab : mo;\n";
        o << "    ab = fabs(hyx_evg[" << e << "][" << (P - 1) << "]); mom1 = (ab > mom1) ? ab : mom1;\n";
    }
    o << "#pragma unroll\n    for (int off = 1; off < HYX_G; off <<= 1) {\n"
         "        ab = __shfl_xor(mx, off); mx = (ab > mx) ? ab : mx;\n"
         "        ab = __shfl_xor(mo, off); mo = (ab > mo) ? ab : mo;\n"
         "        ab = __shfl_xor(mom1, off); mom1 = (ab > mom1) ? ab : mom1;\n    }\n";
    o << "    const double num = (mx <= 1.0) ? 1.0 : mx;\n";
    o << "    const double rho_o = pow(num / mo, " << lit(1.0 / P) << ");\n";
    o << "    const double rho_om1 = pow(num / mom1, " << lit(1.0 / (P - 1)) << ");\n";
    o << "    const double rho_m = (rho_om1 < rho_o) ? rho_om1 : rho_o;\n";
    const double e1 = std::exp(1.0);
    const double rhofac = std::exp(-7.0 / 10.0 / (double)(P - 1)) / (e1 * e1);
    o << "    double h = rho_m * " << lit(rhofac) << ";\n";
    o << "    const double lim = fabs(max_dt);\n    h = (h < lim) ? h : lim;\n    if (max_dt < 0.0) h = -h;\n";
    if (E) {
        // The same fast exclusion as the lane forms; the coefficients of the steps it
        // cannot clear go to tc (event rows by every lane, identical values; state rows
        // by their owners).
        const uint64_t off = (uint64_t)n * (P + 1) + (uint64_t)(E->nnodes - E->n) * (P + 1);
        o << "    hyx_evm = 0ull;\n";
        o << "#pragma unroll\n    for (int e = 0; e < " << nev << "; ++e) {\n"
             "        double hk = 1.0, emag = 0.0, lo = 0.0, hi = 0.0;\n"
             "#pragma unroll\n        for (int k = 0; k <= " << P << "; ++k) {\n"
             "            const double c = hyx_evg[e][k] * hk;\n            hk *= h;\n            emag += fabs(c);\n"
             "            if (k > 0) { lo += fmin(c, 0.0); hi += fmax(c, 0.0); }\n        }\n"
             "        const double slack = " << lit(4.0 * (P + 1) * 1.1102230246251565e-16) << " * emag;\n"
             "        const double c0 = hyx_evg[e][0];\n"
             "        if (emag > 0.0 && isfinite(emag) && !(c0 + lo > slack || c0 + hi < -slack)) hyx_evm |= 1ull << e;\n"
             "    }\n";
        o << "    if (hyx_evm != 0ull && tc) {\n";
        for (uint32_t e = 0; e < nev; ++e)
            for (int k = 0; k <= P; ++k)
                o << "        tc[(hyx_u64)" << (off + (uint64_t)e * (P + 1) + k) << " * tcs] = hyx_evg[" << e << "][" << k
                  << "];\n";
        o << "        if constexpr (!WTC) {\n";
        for (int oo = 0; oo < own; ++oo) {
            o << "            if (tab.vo[" << oo << "] >= 0)\n";
            o << "                for (int k = 0; k <= HYX_P; ++k) tc[((hyx_u64)tab.vo[" << oo << "] * " << (P + 1)
              << " + k) * tcs] = fl[HYX_XB + k * HYX_NX + " << vcol(oo) << "];\n";
        }
        o << "        }\n    }\n";
        o << "    hyx_mx = mx;\n";
    }
    o << "    int nf = 0;\n";
    // State update from the coefficient rows in LDS.
    for (int oo = 0; oo < own; ++oo) {
        if (!high_acc) {
            o << "    { double r = " << xv(oo, P) << ";\n";
            o << "#pragma unroll\n      for (int k = HYX_P - 1; k >= 0; --k) r = __builtin_fma(r, h, fl[HYX_XB + k * HYX_NX + "
              << vcol(oo) << "]);\n";
        } else {
            o << "    { double r = " << xv(oo, 0) << ", cp = 0.0, hk = 1.0;\n";
            o << "#pragma unroll\n      for (int k = 1; k <= HYX_P; ++k) { hk *= h; const double y = fl[HYX_XB + k * HYX_NX + "
              << vcol(oo) << "] * hk - cp; const double tt = r + y; cp = (tt - r) - y; r = tt; }\n";
        }
        o << "      xo[" << oo << "] = r; nf |= (tab.vo[" << oo << "] >= 0 && !isfinite(r)) ? 1 : 0; }\n";
    }
    o << "#pragma unroll\n    for (int off = 1; off < HYX_G; off <<= 1) nf |= __shfl_xor(nf, off);\n";
    o << "    h_out = h;\n    nf_out = nf != 0;\n}\n";
    return h.str() + o.str();
}

std::string generate_cfunc_prelude(const Sys &S, uint32_t n_out)
{
    if (n_out > S.n) throw std::invalid_argument("a compiled function has at most as many outputs as op-list rows");
    Gen g(S, 2, false);
    std::ostringstream &o = g.o;
    o << "#define HYX_N " << S.n << "\n#define HYX_NPAR " << std::max<uint32_t>(S.npar, 1) << "\n#define HYX_NOUT "
      << n_out << "\ntypedef unsigned long long hyx_u64;\n";
    o << "__device__ __forceinline__ void hyx_cfunc_eval(const double (&x)[HYX_N], const double (&pv)[HYX_NPAR], "
         "double t0, double *__restrict__ out, hyx_u64 os)\n{\n    (void)pv; (void)t0;\n";
    for (uint32_t i = 0; i < S.n; ++i) o << "    const double " << Gen::cv(i, 0) << " = x[" << i << "];\n";
    for (uint32_t j = S.n; j < S.nnodes; ++j) g.order0(j);
    for (uint32_t q = 0; q < n_out; ++q) o << "    out[(hyx_u64)" << q << " * os] = " << g.val(S.rhs[q], 0) << ";\n";
    o << "}\n";
    return o.str();
}

std::string generate_events_prelude(const Sys &E, uint32_t n_sys, int P, const EventSpec &ev, bool high_acc)
{
    const uint32_t nev = (uint32_t)ev.dir.size();
    if (nev == 0 || ev.n_te > nev) throw std::invalid_argument("invalid event specification");
    if (nev > E.n) throw std::invalid_argument("the event op-list has fewer rows than events");
    if (E.n < n_sys) throw std::invalid_argument("the event op-list has fewer rows than the system has state variables");
    CompactGen g(E, P, high_acc);
    std::ostringstream &o = g.o;
    const uint32_t nrows = E.n, nodes = E.nnodes - nrows;
    o << "#define HYX_NEV " << nev << "\n#define HYX_NTE " << ev.n_te << "\n";
    o << "#define HYX_EV_ROWS " << (uint64_t)nodes * (P + 1) << "ull\n";
    o << "#define HYX_TOL " << lit(ev.tol) << "\n";
    if (high_acc) o << "#define HYX_HIGH_ACC 1\n";
    o << "#ifndef HYX_SUB\n#define HYX_SUB 1\n#endif\n";
    o << "#ifndef HYX_COMPACT\n#define hyx_sub_sum(v) (v)\n";
    o << "__constant__ double hyx_inv[" << (P + 1) << "] = {0.0";
    for (int k = 1; k <= P; ++k) o << ", " << lit(1.0 / k);
    o << "};\n#endif\n";
    o << "__constant__ int hyx_ev_dir[" << nev << "] = {";
    for (uint32_t e = 0; e < nev; ++e) o << (e ? ", " : "") << ev.dir[e];
    o << "};\n";
    o << "__constant__ double hyx_ev_cd[" << std::max<uint32_t>(ev.n_te, 1) << "] = {";
    for (uint32_t e = 0; e < std::max<uint32_t>(ev.n_te, 1); ++e)
        o << (e ? ", " : "") << lit(e < ev.n_te ? ev.cooldown[e] : -1.0);
    o << "};\n";
    o << "__constant__ int hyx_ev_stop[" << std::max<uint32_t>(ev.n_te, 1) << "] = {";
    for (uint32_t e = 0; e < std::max<uint32_t>(ev.n_te, 1); ++e)
        o << (e ? ", " : "") << ((e < ev.n_te && e < ev.stopping.size()) ? ev.stopping[e] : 1);
    o << "};\n";
    // Event-function jet: the event op-list's nodes at orders 0..P from the state
    // coefficients of the step (rows 0..n-1 of tc); node rows live in et.
    o << "#define T(node, kk) (((node) < " << n_sys << "u) ? tc[((hyx_u64)(node) * (HYX_P + 1) + (hyx_u64)(kk)) * ld] : "
      << "et[((hyx_u64)((node) - " << nrows << "u) * (HYX_P + 1) + (hyx_u64)(kk)) * ld])\n";
    o << "__device__ __forceinline__ void hyx_event_jet(double *__restrict__ tc, double *__restrict__ et, "
         "hyx_u64 ld, const double *__restrict__ pv, double t0)\n{\n";
    o << "    (void)pv; (void)t0; (void)tc; (void)et; (void)ld;\n";
    o << "    const int hyx_sub = (HYX_SUB > 1) ? (int)(__lane_id() & (HYX_SUB - 1)) : 0;\n    (void)hyx_sub;\n";
    for (uint32_t j = nrows; j < E.nnodes; ++j) g.order0(j);
    if (nodes > 0) {
        o << "    for (int k = 1; k <= HYX_P; ++k) {\n";
        for (uint32_t j = nrows; j < E.nnodes; ++j) g.orderk(j);
        o << "    }\n";
    }
    o << "}\n";
    // Event e's coefficient k from the jet: a node row, a state row, or a number / parameter (order 0 only).
    o << "__device__ __forceinline__ double hyx_ev_src(int e, int k, const double *__restrict__ tc, "
         "const double *__restrict__ et, hyx_u64 ld, const double *__restrict__ pv)\n{\n";
    o << "    (void)tc; (void)et; (void)ld; (void)pv;\n    switch (e) {\n";
    for (uint32_t e = 0; e < nev; ++e) {
        const uint32_t w = E.rhs[e], kind = HYX_OPERAND_KIND(w), i = HYX_OPERAND_INDEX(w);
        o << "    case " << e << ": return ";
        if (kind == HYX_KIND_NODE) {
            if (i >= n_sys && i < nrows) throw std::invalid_argument("an event function refers to a padding row");
            o << "T(" << i << "u, k)";
        } else if (kind == HYX_KIND_CONST)
            o << "(k == 0) ? " << lit(E.consts[i]) << " : 0.0";
        else
            o << "(k == 0) ? pv[" << i << "] : 0.0";
        o << ";\n";
    }
    o << "    }\n    return 0.0;\n}\n#undef T\n";
    // Rows in tc after the state's: the event nodes' (compact form), then the event
    // functions' coefficients [e][k], which the root scan reads (hyx_ev_coeff).
    o << "#define HYX_EV_OUT_ROW ((hyx_u64)HYX_N * (HYX_P + 1) + HYX_EV_ROWS)\n";
    o << "__device__ __forceinline__ double *hyx_ev_rows(double *tc, hyx_u64 ld) { return tc + (hyx_u64)HYX_N * "
         "(HYX_P + 1) * ld; }\n";
    o << "__device__ __forceinline__ double hyx_ev_coeff(int e, int k, const double *__restrict__ tc, hyx_u64 ld)\n"
         "{\n    return tc[(HYX_EV_OUT_ROW + (hyx_u64)e * (HYX_P + 1) + (hyx_u64)k) * ld];\n}\n";
    o << "__device__ __forceinline__ void hyx_ev_norms(double *tc, hyx_u64 ld, const double *pv, double t0, double &mo, "
         "double &mom1)\n{\n    double *et = hyx_ev_rows(tc, ld);\n    hyx_event_jet(tc, et, ld, pv, t0);\n"
         "    for (int e = 0; e < HYX_NEV; ++e) {\n"
         "        for (int k = 0; k <= HYX_P; ++k)\n"
         "            tc[(HYX_EV_OUT_ROW + (hyx_u64)e * (HYX_P + 1) + (hyx_u64)k) * ld] = hyx_ev_src(e, k, tc, et, ld, pv);\n"
         "        const double a = fabs(hyx_ev_coeff(e, HYX_P, tc, ld));\n"
         "        const double b = fabs(hyx_ev_coeff(e, HYX_P - 1, tc, ld));\n"
         "        mo = (a > mo) ? a : mo;\n        mom1 = (b > mom1) ? b : mom1;\n    }\n}\n";
    return o.str();
}

std::string events_forward_decl()
{
    return "#define HYX_EVENTS 1\ntypedef unsigned long long hyx_u64;\n"
           "__device__ __forceinline__ void hyx_ev_norms(double *tc, hyx_u64 ld, const double *pv, double t0, "
           "double &mo, double &mom1);\n";
}

int history_nodes(const Sys &S)
{
    // Nodes whose coefficient history later orders read: state variables, the
    // node operands of convolutions and the self-referencing recurrences.
    std::vector<char> h(S.nnodes, 0);
    for (uint32_t i = 0; i < S.n; ++i) h[i] = 1;
    for (uint32_t j = S.n; j < S.nnodes; ++j) {
        if (S.deg[j] == 0) continue;
        const uint32_t op = S.op[j];
        const bool conv = op == HYX_OP_MUL || op == HYX_OP_DIV || op == HYX_OP_SQUARE || op == HYX_OP_SQRT
                          || op == HYX_OP_POW || op == HYX_OP_EXP || op == HYX_OP_LOG || op == HYX_OP_SIN
                          || op == HYX_OP_COS;
        if (!conv) continue;
        int nn = 0;
        for (auto a : S.args[j])
            if (HYX_OPERAND_KIND(a) == HYX_KIND_NODE && S.deg[HYX_OPERAND_INDEX(a)] > 0) ++nn;
        if (op == HYX_OP_MUL && nn < 2) continue;
        if (op == HYX_OP_DIV && !(HYX_OPERAND_KIND(S.args[j][1]) == HYX_KIND_NODE && S.deg[HYX_OPERAND_INDEX(S.args[j][1])] > 0)) continue;
        for (auto a : S.args[j])
            if (HYX_OPERAND_KIND(a) == HYX_KIND_NODE) h[HYX_OPERAND_INDEX(a)] = 1;
        if (op != HYX_OP_MUL && op != HYX_OP_SQUARE) h[j] = 1;
    }
    int c = 0;
    for (auto v : h) c += v;
    return c;
}

bool choose_compact(const Sys &S, int order)
{
    // Unrolled code keeps every history in registers (256 doubles per lane, spilling a little beyond).
    return (long)history_nodes(S) * (order + 1) > 320;
}

std::string generate_prelude(const Sys &S, int order, bool high_acc, bool compact, int split_min, int sub,
                             const Sys *E, uint32_t nev, int lds_rows)
{
    if (order < 2) throw std::invalid_argument("the Taylor order must be at least 2");
    if (compact) {
        CompactGen g(S, order, high_acc);
        g.sub = sub;
        g.events = E != nullptr;
        return g.run();
    }
    Gen g(S, order, high_acc);
    g.split_min = split_min;
    g.E = E;
    g.nev = nev;
    g.no_contract = E == nullptr && has_twin(S);
    // State rows go to LDS when that brings the remaining histories within about 200
    // VGPRs (two waves per SIMD) and the rows of a 64-lane block fit in 64 KB. When the
    // node histories alone exceed that, the extra LDS traffic buys no occupancy.
    const long hist = history_nodes(S), nst = S.n;
    const bool rows_fit = (long)(order + 1) * S.n * 64 * 8 <= 65536;
    if (lds_rows == 1 && !rows_fit) throw std::invalid_argument("the state rows of a 64-lane block exceed 64 KB of LDS");
    g.lds_state = lds_rows >= 0 ? lds_rows == 1
                                : hist * (order + 1) > 160 && (hist - nst) * (order + 1) <= 96 && rows_fit;
    return g.run();
}

// ---- Twin form ----------------------------------------------------------------------------
//
// A latency form for small systems whose right-hand sides add two isomorphic
// subexpressions (the planar CR3BP: the pull of each primary). Two adjacent lanes hold
// one instance. The isomorphic pairs of nodes become twin slots: lane A evaluates the
// first member of each pair, lane B the second, with the same instructions; where the
// members differ only in a degree-0 operand (a parameter, a number, a node of constant
// series) the slot reads a per-lane binding (HYX_OP_TWIN_BIND). Every other node (shared
// operands, the linear sums that combine the pairs, the state variables, the step size
// and the update) is evaluated identically by both lanes; a consumer of a twin member
// reads it through HYX_OP_TWIN_XCH (lane A's or lane B's value, broadcast over the pair by DPP).
// Each node keeps the operands, operand order and recurrence of the lane form, so with
// implicit fusion off in both forms the twin form's arithmetic is the lane form's bit for
// bit: the persistent kernel can hand its last, longest-running instances to the twin
// form, which then runs each step on two lanes in about half the instructions.

namespace
{

struct TwinPlanner {
    const Sys &S;
    std::vector<int> partner; // node -> partner (itself for a shared node), -1 unpaired
    std::vector<char> side;   // 1: lane A member, 2: lane B member
    std::vector<uint64_t> hs;
    std::string why;

    explicit TwinPlanner(const Sys &s) : S(s), partner(s.nnodes, -1), side(s.nnodes, 0), hs(s.nnodes, 0) {}

    bool node(uint32_t w) const { return HYX_OPERAND_KIND(w) == HYX_KIND_NODE; }
    bool k0(uint32_t w) const { return !node(w) || (HYX_OPERAND_INDEX(w) >= S.n && S.deg[HYX_OPERAND_INDEX(w)] == 0); }

    // The operand position holding the degree-0 value of an add / sub / mul with exactly one
    // such operand (-1: none; sub only as x - K).
    int kpos(uint32_t j) const
    {
        const uint32_t op = S.op[j];
        if (op != HYX_OP_ADD && op != HYX_OP_SUB && op != HYX_OP_MUL) return -1;
        const auto &A = S.args[j];
        const bool a0 = k0(A[0]), a1 = k0(A[1]);
        if (a0 == a1) return -1;
        if (op == HYX_OP_SUB && a0) return -1;
        return a0 ? 0 : 1;
    }

    // Shape class of a node for pairing: add and sub with one degree-0 operand are one
    // class (x + K), products with one degree-0 operand another (K * x).
    uint64_t cls(uint32_t j) const
    {
        const int q = kpos(j);
        if (q >= 0) return S.op[j] == HYX_OP_MUL ? 1000 : 1001;
        return S.op[j];
    }

    uint64_t mix(uint64_t h, uint64_t v) const { return (h ^ v) * 1099511628211ull + 0x9e3779b97f4a7c15ull; }

    // Structural hash: states by index, degree-0 operands as one wildcard.
    uint64_t hash(uint32_t w)
    {
        if (k0(w)) return 7;
        const uint32_t i = HYX_OPERAND_INDEX(w);
        if (i < S.n) return mix(11, i);
        if (hs[i]) return hs[i];
        uint64_t h = mix(13, cls(i));
        const int q = kpos(i);
        const auto &A = S.args[i];
        for (size_t a = 0; a < A.size(); ++a)
            if ((int)a != q) h = mix(h, hash(A[a]));
        if (S.op[i] == HYX_OP_POW) {
            uint64_t b;
            std::memcpy(&b, &S.consts[S.aux[i]], 8);
            h = mix(h, b);
        }
        if (S.op[i] == HYX_OP_SIN || S.op[i] == HYX_OP_COS) h = mix(h, S.op[i]);
        return hs[i] = h | 1;
    }

    bool fail(const std::string &m)
    {
        if (why.empty()) why = m;
        return false;
    }

    bool pair_operand(uint32_t wa, uint32_t wb)
    {
        if (k0(wa) || k0(wb)) return (k0(wa) && k0(wb)) ? true : fail("a degree-0 operand pairs with a series");
        const uint32_t a = HYX_OPERAND_INDEX(wa), b = HYX_OPERAND_INDEX(wb);
        if (a < S.n || b < S.n) return a == b ? true : fail("twins read different state variables");
        return pair(a, b);
    }

    bool pair(uint32_t a, uint32_t b)
    {
        if (a == b) {
            if (side[a]) return fail("a node is both shared and a twin");
            partner[a] = (int)a;
            return true;
        }
        if (partner[a] >= 0 || partner[b] >= 0) {
            return (partner[a] == (int)b && side[a] == 1 && side[b] == 2) ? true : fail("inconsistent twin pairing");
        }
        if (cls(a) != cls(b) || S.args[a].size() != S.args[b].size()) return fail("twins differ in shape");
        if (S.op[a] == HYX_OP_POW && S.consts[S.aux[a]] != S.consts[S.aux[b]]) return fail("different exponents");
        if (S.op[a] == HYX_OP_TIME) return fail("time in a twin");
        partner[a] = (int)b;
        partner[b] = (int)a;
        side[a] = 1;
        side[b] = 2;
        const int q = kpos(a), qb = kpos(b);
        const auto &A = S.args[a], &B = S.args[b];
        if (q >= 0) {
            // x + K / K * x: the series operands pair, the degree-0 ones become a binding.
            if (!pair_operand(A[1 - q], B[1 - qb])) return false;
        } else {
            for (size_t i = 0; i < A.size(); ++i)
                if (!pair_operand(A[i], B[i])) return false;
        }
        if (S.op[a] == HYX_OP_SIN || S.op[a] == HYX_OP_COS) return pair(S.aux[a], S.aux[b]);
        return true;
    }

    // Pair the isomorphic terms of each right-hand side's linear sum.
    bool plan()
    {
        for (uint32_t i = 0; i < S.n; ++i) {
            std::vector<uint32_t> terms;
            std::vector<uint32_t> st{S.rhs[i]};
            while (!st.empty()) {
                const uint32_t w = st.back();
                st.pop_back();
                if (!node(w)) continue;
                const uint32_t j = HYX_OPERAND_INDEX(w);
                if (j < S.n || S.deg[j] == 0) continue;
                const uint32_t op = S.op[j];
                if ((op == HYX_OP_ADD || op == HYX_OP_SUB || op == HYX_OP_SUM || op == HYX_OP_NEG) && kpos(j) < 0) {
                    for (auto a : S.args[j]) st.push_back(a);
                    continue;
                }
                terms.push_back(j);
            }
            std::reverse(terms.begin(), terms.end());
            std::vector<char> used(terms.size(), 0);
            for (size_t x = 0; x < terms.size(); ++x) {
                if (used[x] || side[terms[x]]) continue;
                for (size_t y = x + 1; y < terms.size(); ++y) {
                    if (used[y] || side[terms[y]] || terms[y] == terms[x]) continue;
                    if (hash(HYX_OPERAND(HYX_KIND_NODE, terms[x])) != hash(HYX_OPERAND(HYX_KIND_NODE, terms[y]))) continue;
                    // Lane A takes the first term unless its subtrees are already lane B's.
                    const auto p0 = partner;
                    const auto s0 = side;
                    if (!pair(terms[x], terms[y])) {
                        partner = p0;
                        side = s0;
                        why.clear();
                        if (!pair(terms[y], terms[x])) return false;
                    }
                    used[x] = used[y] = 1;
                    break;
                }
            }
        }
        // A shared node may not read a twin member directly (it would differ per lane).
        for (uint32_t j = S.n; j < S.nnodes; ++j) {
            if (partner[j] != (int)j) continue;
            for (auto a : S.args[j])
                if (node(a) && side[HYX_OPERAND_INDEX(a)]) return fail("a shared node reads a twin member");
        }
        // Worth it: the twins hold at least a third of the convolution work.
        long twin = 0, all = 0;
        for (uint32_t j = S.n; j < S.nnodes; ++j) {
            const uint32_t op = S.op[j];
            long w = 1;
            if (S.deg[j] > 0 && kpos(j) < 0 && op != HYX_OP_ADD && op != HYX_OP_SUB && op != HYX_OP_SUM &&
                op != HYX_OP_NEG && op != HYX_OP_TIME)
                w = 3;
            if (op == HYX_OP_POW) w = 4;
            all += w;
            if (side[j]) twin += w;
        }
        if (twin == 0) return fail("no isomorphic right-hand-side terms");
        if (3 * twin < all) return fail("the twin nodes hold too little of the work");
        return true;
    }
};

// The op-list of the twin form: twin pairs become slots, per-lane bindings and exchanges
// become nodes, in an order where every node follows its operands.
Sys build_twin(const Sys &S, const TwinPlanner &tp)
{
    // An operand of a proto node: a raw operand word of S (state variable, number,
    // parameter) or a proto node id.
    struct Ref {
        bool proto;
        uint32_t v;
        bool operator<(const Ref &o) const { return proto != o.proto ? proto < o.proto : v < o.v; }
    };
    struct Proto {
        uint32_t op, aux;
        std::vector<Ref> args;
        int partner = -1; // sin / cos: the partner's proto id
    };
    std::vector<Proto> pr;
    std::vector<int> pid(S.nnodes, -1); // S node -> proto (both members of a twin pair map to the slot)
    std::map<std::pair<uint32_t, uint32_t>, uint32_t> xch;
    std::map<std::tuple<uint32_t, uint32_t, uint32_t, uint32_t, uint32_t>, uint32_t> bind;
    auto addp = [&](Proto p) {
        pr.push_back(std::move(p));
        return (uint32_t)(pr.size() - 1);
    };
    std::function<uint32_t(uint32_t)> mk;
    // An operand as a non-twin node reads it: twin members through an exchange.
    auto opnd = [&](uint32_t w) -> Ref {
        if (HYX_OPERAND_KIND(w) != HYX_KIND_NODE || HYX_OPERAND_INDEX(w) < S.n) return Ref{false, w};
        const uint32_t i = HYX_OPERAND_INDEX(w);
        const uint32_t p = mk(i);
        if (!tp.side[i]) return Ref{true, p};
        const uint32_t aux = tp.side[i] == 1 ? 0u : 1u;
        auto it = xch.find({p, aux});
        if (it == xch.end()) it = xch.emplace(std::make_pair(p, aux), addp(Proto{HYX_OP_TWIN_XCH, aux, {Ref{true, p}}})).first;
        return Ref{true, it->second};
    };
    auto binding = [&](uint32_t wa, bool na, uint32_t wb, bool nb) -> Ref {
        const Ref a = opnd(wa), b = opnd(wb);
        const uint32_t aux = (na ? 1u : 0u) | (nb ? 2u : 0u);
        const auto key = std::make_tuple((uint32_t)a.proto, a.v, (uint32_t)b.proto, b.v, aux);
        auto it = bind.find(key);
        if (it == bind.end()) it = bind.emplace(key, addp(Proto{HYX_OP_TWIN_BIND, aux, {a, b}})).first;
        return Ref{true, it->second};
    };
    mk = [&](uint32_t j) -> uint32_t {
        if (pid[j] >= 0) return (uint32_t)pid[j];
        if (!tp.side[j]) {
            Proto p{S.op[j], S.aux[j], {}};
            for (auto w : S.args[j]) p.args.push_back(opnd(w));
            const uint32_t id = addp(p);
            pid[j] = (int)id;
            return id;
        }
        const uint32_t a = tp.side[j] == 1 ? j : (uint32_t)tp.partner[j], b = (uint32_t)tp.partner[a];
        const int q = tp.kpos(a), qb = tp.kpos(b);
        // A series operand of a slot: a slot, a shared node or a state variable.
        auto sop = [&](uint32_t w) -> Ref {
            if (HYX_OPERAND_KIND(w) != HYX_KIND_NODE || HYX_OPERAND_INDEX(w) < S.n) return Ref{false, w};
            return Ref{true, mk(HYX_OPERAND_INDEX(w))};
        };
        const auto &A = S.args[a], &B = S.args[b];
        Proto p{S.op[a], S.aux[a], {}};
        if (q >= 0) {
            // x + K (lane A add or sub x - K) or K * x: add / mul of the series slot and the
            // binding, in lane A's operand order (x - K as x + (-K), which rounds the same).
            const Ref x = sop(A[1 - q]);
            const Ref k = binding(A[q], S.op[a] == HYX_OP_SUB, B[qb], S.op[b] == HYX_OP_SUB);
            p.op = S.op[a] == HYX_OP_MUL ? (uint32_t)HYX_OP_MUL : (uint32_t)HYX_OP_ADD;
            p.aux = 0;
            p.args = (q == 0) ? std::vector<Ref>{k, x} : std::vector<Ref>{x, k};
        } else {
            for (size_t i = 0; i < A.size(); ++i) {
                if (tp.k0(A[i]))
                    p.args.push_back(A[i] == B[i] ? opnd(A[i]) : binding(A[i], false, B[i], false));
                else
                    p.args.push_back(sop(A[i]));
            }
        }
        const uint32_t id = addp(p);
        pid[a] = pid[b] = (int)id;
        return id;
    };
    for (uint32_t j = S.n; j < S.nnodes; ++j) mk(j);
    for (uint32_t j = S.n; j < S.nnodes; ++j)
        if (S.op[j] == HYX_OP_SIN || S.op[j] == HYX_OP_COS) pr[pid[j]].partner = pid[S.aux[j]];
    std::vector<Ref> rhs;
    for (uint32_t i = 0; i < S.n; ++i) rhs.push_back(opnd(S.rhs[i]));
    // Topological order, operands first.
    std::vector<int> state(pr.size(), 0), slot(pr.size(), -1);
    std::vector<uint32_t> order;
    std::function<void(uint32_t)> visit = [&](uint32_t id) {
        if (state[id] == 2) return;
        state[id] = 1;
        for (const Ref &r : pr[id].args)
            if (r.proto) visit(r.v);
        state[id] = 2;
        slot[id] = (int)order.size();
        order.push_back(id);
    };
    for (uint32_t id = 0; id < pr.size(); ++id) visit(id);
    Sys T;
    T.n = S.n;
    T.npar = S.npar;
    T.consts = S.consts;
    T.nconsts = S.nconsts;
    T.nnodes = S.n + (uint32_t)order.size();
    T.op.assign(T.nnodes, 0);
    T.aux.assign(T.nnodes, 0);
    T.args.assign(T.nnodes, {});
    auto word = [&](const Ref &r) { return r.proto ? HYX_OPERAND(HYX_KIND_NODE, S.n + (uint32_t)slot[r.v]) : r.v; };
    for (uint32_t r = 0; r < order.size(); ++r) {
        const Proto &p = pr[order[r]];
        const uint32_t j = S.n + r;
        T.op[j] = p.op;
        T.aux[j] = (p.op == HYX_OP_SIN || p.op == HYX_OP_COS) ? S.n + (uint32_t)slot[p.partner] : p.aux;
        for (const Ref &a : p.args) T.args[j].push_back(word(a));
    }
    for (const Ref &r : rhs) T.rhs.push_back(word(r));
    compute_degrees(T);
    return T;
}

} // namespace

bool has_twin(const Sys &S)
{
    TwinPlanner tp(S);
    return tp.plan();
}

Sys twin_oplist(const Sys &S, std::string &why)
{
    TwinPlanner tp(S);
    if (!tp.plan()) {
        why = tp.why;
        return Sys{};
    }
    return build_twin(S, tp);
}

std::string generate_twin_prelude(const Sys &S, int order, bool high_acc, int split_min, std::string &why)
{
    if (order < 2) throw std::invalid_argument("the Taylor order must be at least 2");
    TwinPlanner tp(S);
    if (!tp.plan()) {
        why = tp.why;
        return "";
    }
    const Sys T = build_twin(S, tp);
    Gen g(T, order, high_acc);
    g.split_min = split_min;
    g.twin = true;
    g.no_contract = true;
    return g.run();
}

// ---- Quad form ----------------------------------------------------------------------------
//
// A latency form one step past the twin: four lanes per instance, lane q = 2 * primary +
// component. It applies to a twin op-list T whose nodes split once more under a pairing
// sigma of the state variables (an involution without fixed points; the planar CR3BP's
// x <-> y, vx <-> vy), such that the right-hand sides of i and sigma(i) are isomorphic when
// every state variable s read by one is read as sigma(s) by the other. The lanes of
// component x evaluate the first member of each pair of nodes, the lanes of component y the
// second, with the same instructions. Each lane holds the state in its own order (variable
// i of the op-list is x[i] on the x lanes, x[sigma(i)] on the y lanes), so a node that reads
// s on the x lanes reads sigma(s) on the y lanes. Members that differ in a degree-0 operand
// read a per-component binding (HYX_OP_QUAD_BIND); a node s + K whose partner is the bare
// state variable sigma(s) adds K on the x lanes and -0.0 on the y lanes (v + -0.0 == v for
// every v, -0.0 included). Nodes common to both components (the twin's shared nodes, and
// + or * of the two members of a pair, such as x^2 + y^2: own value op the other
// component's) are evaluated by all four lanes; the right-hand side of sigma(i) is the
// other component's value of the right-hand side of i (HYX_OP_QUAD_SWAP). Operands, their
// order (up to swapping the operands of + and *, which round the same) and every recurrence
// stay the lane form's, so the quad form's arithmetic is the lane form's bit for bit, and
// each lane runs about the instructions of one component of one primary per order.

namespace
{

struct QuadPlanner {
    const Sys &T;
    std::vector<uint32_t> sg; // the pairing of the state variables
    std::vector<int> partner; // node -> partner node (itself: common), -1 unset
    std::vector<char> side;   // 1: x member, 2: y member
    std::vector<char> sym;    // common + / * of a pair: 1 operand 0 is the x member, 2 operand 1 is
    std::vector<char> addk;   // s + K paired with a bare state variable: 1 evaluated on x lanes, 2 on y
    std::string why;

    explicit QuadPlanner(const Sys &t) : T(t) {}

    bool node(uint32_t w) const { return HYX_OPERAND_KIND(w) == HYX_KIND_NODE; }
    uint32_t ix(uint32_t w) const { return HYX_OPERAND_INDEX(w); }
    bool k0(uint32_t w) const { return !node(w) || (ix(w) >= T.n && T.deg[ix(w)] == 0); }
    bool fail(const std::string &m)
    {
        if (why.empty()) why = m;
        return false;
    }

    // s + K or K + s (a state variable plus a degree-0 operand): the position of K, else -1.
    int kpos_state_add(uint32_t j) const
    {
        if (T.op[j] != HYX_OP_ADD || T.args[j].size() != 2) return -1;
        const uint32_t a = T.args[j][0], b = T.args[j][1];
        if (k0(a) && node(b) && ix(b) < T.n) return 0;
        if (k0(b) && node(a) && ix(a) < T.n) return 1;
        return -1;
    }

    // Operand wx on the x lanes pairs with wy on the y lanes.
    bool pair2(uint32_t wx, uint32_t wy)
    {
        const bool kx = k0(wx), ky = k0(wy);
        if (kx || ky) return (kx && ky) ? true : fail("a degree-0 operand pairs with a series");
        const uint32_t a = ix(wx), b = ix(wy);
        const bool sa = a < T.n, sb = b < T.n;
        if (sa && sb) return sg[a] == b ? true : fail("paired nodes read unpaired state variables");
        if (sa || sb) {
            const uint32_t nd = sa ? b : a, st = sa ? a : b;
            const int q = kpos_state_add(nd);
            if (q < 0) return fail("a state variable pairs with a node other than s + K");
            if (sg[ix(T.args[nd][1 - q])] != st) return fail("s + K pairs with an unpaired state variable");
            const char sd = sa ? 2 : 1;
            if (partner[nd] >= 0) return fail("s + K is also paired with a node");
            if (addk[nd] && addk[nd] != sd) return fail("s + K pairs on both components");
            addk[nd] = sd;
            return true;
        }
        return pairn(a, b);
    }

    bool pairn(uint32_t a, uint32_t b)
    {
        if (a == b) return common(a);
        if (addk[a] || addk[b]) return fail("s + K is also paired with a node");
        if (partner[a] >= 0 || partner[b] >= 0)
            return (partner[a] == (int)b && side[a] == 1 && side[b] == 2) ? true
                                                                        : fail("inconsistent component pairing");
        const uint32_t op = T.op[a];
        if (op != T.op[b] || T.args[a].size() != T.args[b].size()) return fail("paired nodes differ in shape");
        if (op == HYX_OP_TIME) return fail("time in a component pair");
        if (op == HYX_OP_POW && T.consts[T.aux[a]] != T.consts[T.aux[b]]) return fail("different exponents");
        if (op != HYX_OP_POW && op != HYX_OP_SIN && op != HYX_OP_COS && T.aux[a] != T.aux[b])
            return fail("paired nodes differ in shape");
        partner[a] = (int)b;
        partner[b] = (int)a;
        side[a] = 1;
        side[b] = 2;
        const auto &A = T.args[a], &B = T.args[b];
        for (size_t i = 0; i < A.size(); ++i)
            if (!pair2(A[i], B[i])) return false;
        if (op == HYX_OP_SIN || op == HYX_OP_COS) return pairn(T.aux[a], T.aux[b]);
        return true;
    }

    // Node j is evaluated alike by both components.
    bool common(uint32_t j)
    {
        if (partner[j] == (int)j) return true;
        if (partner[j] >= 0 || addk[j]) return fail("a node is both common and paired");
        partner[j] = (int)j;
        const uint32_t op = T.op[j];
        const auto &A = T.args[j];
        if ((op == HYX_OP_ADD || op == HYX_OP_MUL) && A.size() == 2 && !k0(A[0]) && !k0(A[1]) && A[0] != A[1]) {
            // The two members of a pair combined by a commutative op (x^2 + y^2): either
            // operand may be the x member.
            for (int o = 0; o < 2; ++o) {
                const auto p0 = partner;
                const auto s0 = side, y0 = sym, k0v = addk;
                const std::string w0 = why;
                if (pair2(A[o], A[1 - o])) {
                    sym[j] = (char)(1 + o);
                    return true;
                }
                partner = p0;
                side = s0;
                sym = y0;
                addk = k0v;
                why = w0;
            }
        }
        for (auto w : A) {
            if (k0(w)) continue;
            if (ix(w) < T.n) return fail("a common node reads a state variable");
            if (!common(ix(w))) return false;
        }
        if (op == HYX_OP_SIN || op == HYX_OP_COS) return common(T.aux[j]);
        return true;
    }

    bool attempt(const std::vector<uint32_t> &s)
    {
        sg = s;
        partner.assign(T.nnodes, -1);
        side.assign(T.nnodes, 0);
        sym.assign(T.nnodes, 0);
        addk.assign(T.nnodes, 0);
        for (uint32_t f = 0; f < T.n; ++f)
            if (f < sg[f] && !pair2(T.rhs[f], T.rhs[sg[f]])) return false;
        // Worth it: the component pairs hold convolution work.
        long split = 0;
        for (uint32_t j = T.n; j < T.nnodes; ++j) {
            const uint32_t op = T.op[j];
            if (side[j] && T.deg[j] > 0 && (op == HYX_OP_MUL || op == HYX_OP_SQUARE || op == HYX_OP_DIV ||
                                            op == HYX_OP_POW || op == HYX_OP_EXP || op == HYX_OP_SIN))
                ++split;
        }
        return split > 0 ? true : fail("no convolution splits between the components");
    }

    bool plan()
    {
        const uint32_t n = T.n;
        if (n % 2 != 0) return fail("an odd number of state variables");
        if (n > 12) return fail("more than 12 state variables");
        std::vector<uint32_t> cur(n, UINT32_MAX);
        std::function<bool(uint32_t)> rec = [&](uint32_t i) -> bool {
            while (i < n && cur[i] != UINT32_MAX) ++i;
            if (i == n) return attempt(cur);
            for (uint32_t j = i + 1; j < n; ++j) {
                if (cur[j] != UINT32_MAX) continue;
                cur[i] = j;
                cur[j] = i;
                if (rec(i + 1)) return true;
                cur[i] = cur[j] = UINT32_MAX;
            }
            return false;
        };
        if (rec(0)) {
            why.clear();
            return true;
        }
        if (why.empty()) why = "no pairing of the state variables";
        return false;
    }
};

// The op-list of the quad form: x members (and the twin's nodes common to both components)
// as nodes, per-component bindings and component swaps as nodes, operands first.
Sys build_quad(const Sys &T, const QuadPlanner &qp)
{
    struct Ref {
        bool proto;
        uint32_t v;
        bool operator<(const Ref &o) const { return proto != o.proto ? proto < o.proto : v < o.v; }
    };
    struct Proto {
        uint32_t op, aux;
        std::vector<Ref> args;
        int partner = -1; // sin / cos: the partner's T node
    };
    std::vector<Proto> pr;
    std::vector<int> pid(T.nnodes, -1), pk(T.nnodes, -1);
    std::map<uint32_t, uint32_t> swp;
    std::map<std::pair<Ref, Ref>, uint32_t> qb;
    Sys Q;
    Q.n = T.n;
    Q.npar = T.npar;
    Q.consts = T.consts;
    uint32_t negz = UINT32_MAX;
    auto addp = [&](Proto p) {
        pr.push_back(std::move(p));
        return (uint32_t)(pr.size() - 1);
    };
    auto swap_of = [&](uint32_t p) {
        auto it = swp.find(p);
        if (it == swp.end()) it = swp.emplace(p, addp(Proto{HYX_OP_QUAD_SWAP, 0, {Ref{true, p}}})).first;
        return it->second;
    };
    auto qbind = [&](Ref x, Ref y) -> Ref {
        auto it = qb.find({x, y});
        if (it == qb.end()) it = qb.emplace(std::make_pair(x, y), addp(Proto{HYX_OP_QUAD_BIND, 0, {x, y}})).first;
        return Ref{true, it->second};
    };
    std::function<uint32_t(uint32_t)> mk;
    auto k0ref = [&](uint32_t w) -> Ref { return qp.node(w) ? Ref{true, mk(qp.ix(w))} : Ref{false, w}; };
    auto state = [&](uint32_t i) { return Ref{false, HYX_OPERAND(HYX_KIND_NODE, i)}; };
    // s + K paired with a bare state variable.
    auto mkk = [&](uint32_t nd) -> uint32_t {
        if (pk[nd] >= 0) return (uint32_t)pk[nd];
        const int q = qp.kpos_state_add(nd);
        const uint32_t s = qp.ix(T.args[nd][1 - q]);
        if (negz == UINT32_MAX) {
            negz = (uint32_t)Q.consts.size();
            Q.consts.push_back(-0.0);
        }
        const Ref kr = k0ref(T.args[nd][q]), zr{false, HYX_OPERAND(HYX_KIND_CONST, negz)};
        const bool onx = qp.addk[nd] == 1;
        const Ref b = onx ? qbind(kr, zr) : qbind(zr, kr);
        // Evaluated on the x lanes it reads s; on the y lanes it reads s there, which is
        // relative variable sigma(s).
        const Ref st = state(onx ? s : qp.sg[s]);
        Proto p{HYX_OP_ADD, 0, {}};
        p.args = (q == 0) ? std::vector<Ref>{b, st} : std::vector<Ref>{st, b};
        pk[nd] = (int)addp(p);
        return (uint32_t)pk[nd];
    };
    // The operand pair (x lanes wx, y lanes wy).
    auto pref = [&](uint32_t wx, uint32_t wy) -> Ref {
        if (qp.k0(wx)) return wx == wy ? k0ref(wx) : qbind(k0ref(wx), k0ref(wy));
        const uint32_t a = qp.ix(wx), b = qp.ix(wy);
        if (a < T.n && b < T.n) return state(a);
        if (a < T.n || b < T.n) return Ref{true, mkk(a < T.n ? b : a)};
        return Ref{true, mk(a)};
    };
    mk = [&](uint32_t j) -> uint32_t {
        if (pid[j] >= 0) return (uint32_t)pid[j];
        if (qp.side[j] == 2) return mk((uint32_t)qp.partner[j]);
        Proto p{T.op[j], T.aux[j], {}};
        const auto &A = T.args[j];
        if (T.op[j] == HYX_OP_SIN || T.op[j] == HYX_OP_COS) p.partner = (int)T.aux[j];
        if (qp.side[j] == 1) {
            const auto &B = T.args[qp.partner[j]];
            for (size_t i = 0; i < A.size(); ++i) p.args.push_back(pref(A[i], B[i]));
        } else if (qp.sym[j]) {
            const int o = qp.sym[j] - 1; // operand o is the x member
            const Ref own = pref(A[o], A[1 - o]);
            const Ref oth = own.proto ? Ref{true, swap_of(own.v)} : state(qp.sg[own.v & 0x3fffffffu]);
            p.args = (o == 0) ? std::vector<Ref>{own, oth} : std::vector<Ref>{oth, own};
        } else {
            for (auto w : A) p.args.push_back(k0ref(w));
        }
        const uint32_t id = addp(p);
        pid[j] = (int)id;
        if (qp.side[j] == 1) pid[qp.partner[j]] = (int)id;
        return id;
    };
    std::vector<Ref> rhs(T.n);
    for (uint32_t f = 0; f < T.n; ++f) {
        const uint32_t g = qp.sg[f];
        if (f > g) continue;
        const uint32_t wx = T.rhs[f], wy = T.rhs[g];
        const Ref r = pref(wx, wy);
        rhs[f] = r;
        if (qp.k0(wx))
            rhs[g] = (wx == wy) ? r : qbind(k0ref(wy), k0ref(wx));
        else if (!r.proto)
            rhs[g] = state(qp.ix(wy));
        else if (qp.ix(wx) == qp.ix(wy))
            rhs[g] = r;
        else
            rhs[g] = Ref{true, swap_of(r.v)};
    }
    // Topological order, operands first.
    for (uint32_t j = T.n; j < T.nnodes; ++j)
        if (T.op[j] == HYX_OP_SIN || T.op[j] == HYX_OP_COS)
            if (pid[j] >= 0 && pr[pid[j]].partner >= 0) mk((uint32_t)pr[pid[j]].partner);
    std::vector<int> st(pr.size(), 0), slot(pr.size(), -1);
    std::vector<uint32_t> order;
    std::function<void(uint32_t)> visit = [&](uint32_t id) {
        if (st[id] == 2) return;
        st[id] = 1;
        for (const Ref &r : pr[id].args)
            if (r.proto) visit(r.v);
        st[id] = 2;
        slot[id] = (int)order.size();
        order.push_back(id);
    };
    for (uint32_t id = 0; id < pr.size(); ++id) visit(id);
    Q.nconsts = (uint32_t)Q.consts.size();
    Q.nnodes = Q.n + (uint32_t)order.size();
    Q.op.assign(Q.nnodes, 0);
    Q.aux.assign(Q.nnodes, 0);
    Q.args.assign(Q.nnodes, {});
    auto word = [&](const Ref &r) { return r.proto ? HYX_OPERAND(HYX_KIND_NODE, Q.n + (uint32_t)slot[r.v]) : r.v; };
    for (uint32_t r = 0; r < order.size(); ++r) {
        const Proto &p = pr[order[r]];
        const uint32_t j = Q.n + r;
        Q.op[j] = p.op;
        Q.aux[j] = (p.op == HYX_OP_SIN || p.op == HYX_OP_COS) ? Q.n + (uint32_t)slot[pid[p.partner]] : p.aux;
        for (const Ref &a : p.args) Q.args[j].push_back(word(a));
    }
    for (const Ref &r : rhs) Q.rhs.push_back(word(r));
    compute_degrees(Q);
    return Q;
}

} // namespace

std::string generate_quad_prelude(const Sys &S, int order, bool high_acc, int split_min, std::string &why)
{
    if (order < 2) throw std::invalid_argument("the Taylor order must be at least 2");
    TwinPlanner tp(S);
    if (!tp.plan()) {
        why = tp.why;
        return "";
    }
    const Sys T = build_twin(S, tp);
    QuadPlanner qp(T);
    if (!qp.plan()) {
        why = qp.why;
        return "";
    }
    const Sys Q = build_quad(T, qp);
    Gen g(Q, order, high_acc);
    g.split_min = split_min;
    g.twin = true;
    g.quad = true;
    g.qsig = qp.sg;
    g.no_contract = true;
    return g.run();
}

// ---- Wave form ----------------------------------------------------------------------------
//
// G lanes per instance (64: one instance per wavefront, or 32 / 16). The nodes of the
// decomposition are spread over the lanes of the group by a static schedule: each node
// gets a level above the levels of its node operands (state variables are level 0),
// chosen so that nodes of one kind share levels, and the nodes of one level and one
// recurrence kind form slots of at most G nodes,
// lane q of a slot evaluating its q-th node. Per order, the state rows come first, then
// the slots level by level, one wavefront-scope fence per level. Every coefficient lives
// in the instance's LDS block:
//   T[k][r] = fl[k * R + r]: orders 0..p-1 of the state variables (rows 0..n-1), a zero
//     row (n), the nodes whose history a recurrence reads, and the constants and
//     parameters a recurrence reads at higher (zero) orders; order p of the state rows
//     and the zero row. The state rows are the layout the group kernels' state update,
//     grid output and event code read.
//   cur[c] = fl[P * R + n + 1 + c]: the current order of every state variable (c = i),
//     the zero entry (c = n), every node (c = node + 1) and every constant or
//     parameter (zero above order 0), read by the linear nodes, the products with a
//     constant or parameter and the right-hand sides;
//   sv[q]: the value of every constant and parameter.
// A lane's operand rows come from per-lane tables loaded once into registers. Each
// recurrence at order k is split into the terms that read orders below k only, emitted
// one level ahead, and the same-order terms on the level's critical path; div, sqrt,
// exp, log and sin/cos keep the summation order of the register forms, products, squares
// and pow add their same-order terms last. Terms that the register forms skip because a
// factor is zero by degree are read here as exact zeros.
namespace
{
enum WaveKind { WK_LIN, WK_MUL, WK_MULS, WK_SQ, WK_POW, WK_DIV, WK_SQRT, WK_EXP, WK_LOG, WK_TRIG, WK_TIME, WK_COUNT };

struct WaveSlot {
    int kind = 0;
    int m = 0;                   // linear slots: operands per lane
    std::vector<uint32_t> nodes; // the node of each lane (lane q < nodes.size())
};
} // namespace

std::string generate_wave_prelude(const Sys &S, int P, bool high_acc, int G, const Sys *E, uint32_t nev,
                                  int split_min, std::string &why)
{
    const int n = (int)S.n, NN = (int)S.nnodes;
    if (NN == n) {
        why = "no nodes";
        return "";
    }
    if (G != 16 && G != 32 && G != 64) throw std::invalid_argument("wave form: G must be 16, 32 or 64");
    auto is_node = [](uint32_t w) { return HYX_OPERAND_KIND(w) == HYX_KIND_NODE; };
    std::vector<int> kind(NN, -1);
    for (int j = n; j < NN; ++j) {
        const auto &A = S.args[j];
        switch (S.op[j]) {
        case HYX_OP_ADD:
        case HYX_OP_SUB:
        case HYX_OP_NEG:
        case HYX_OP_SUM: kind[j] = WK_LIN; break;
        case HYX_OP_MUL: kind[j] = (is_node(A[0]) && is_node(A[1])) ? WK_MUL : WK_MULS; break;
        case HYX_OP_SQUARE: kind[j] = WK_SQ; break;
        case HYX_OP_POW: kind[j] = WK_POW; break;
        case HYX_OP_DIV: kind[j] = WK_DIV; break;
        case HYX_OP_SQRT: kind[j] = WK_SQRT; break;
        case HYX_OP_EXP: kind[j] = WK_EXP; break;
        case HYX_OP_LOG: kind[j] = WK_LOG; break;
        case HYX_OP_SIN:
        case HYX_OP_COS: kind[j] = WK_TRIG; break;
        case HYX_OP_TIME: kind[j] = WK_TIME; break;
        default: why = "an operation the wave form does not implement"; return "";
        }
        for (auto a : A)
            if (is_node(a) && (int)HYX_OPERAND_INDEX(a) >= j) {
                why = "a node operand that is not an earlier node";
                return "";
            }
    }
    // Derived nodes: a linear node or a product with a constant whose only node operand is
    // another node, and whose other operands are constants. The lane that computes that
    // operand (its root: the first node of the chain that is not derived) computes it too,
    // right after it, from the operand's value in a register and with the arithmetic of its
    // own recurrence: no level of its own, and no LDS round trip on the critical path.
    // dlist[r] lists the nodes derived from root r in index (dependency) order.
    std::vector<int> dsrc(NN, -1), root(NN, -1);
    std::vector<std::vector<int>> dlist(NN);
    for (int j = n; j < NN; ++j) {
        root[j] = j;
        if (kind[j] != WK_LIN && kind[j] != WK_MULS) continue;
        int cnt = 0, src = -1;
        bool consts = true;
        for (auto a : S.args[j]) {
            if (is_node(a)) {
                ++cnt;
                src = (int)HYX_OPERAND_INDEX(a);
            } else if (HYX_OPERAND_KIND(a) != HYX_KIND_CONST)
                consts = false;
        }
        if (cnt == 1 && src >= n && consts && S.args[j].size() <= 15) {
            dsrc[j] = src;
            root[j] = root[src];
            dlist[root[j]].push_back(j);
        }
    }
    // Position of a derived node's source in its root's chain: 0 = the root, d + 1 = step d.
    auto dpos = [&](int x) {
        const int r = root[x];
        if (dsrc[x] == r) return 0;
        for (size_t e = 0; e < dlist[r].size(); ++e)
            if (dlist[r][e] == dsrc[x]) return (int)e + 1;
        throw std::logic_error("wave form: a derived node without its source");
    };
    // The shape of a derived step: kind, operation and which operands are nodes (the
    // constants may differ between lanes); its source is matched separately.
    auto dshape = [&](int x) {
        std::string sg = kind[x] == WK_MULS ? "M" : "L" + std::to_string(S.op[x]) + ":";
        if (kind[x] == WK_LIN)
            for (auto a : S.args[x]) sg += is_node(a) ? "n" : "c";
        return sg;
    };
    // Rows with a history: node operands of every recurrence but the linear ones, and
    // the nodes whose recurrence reads their own (or their partner's) history.
    // A product with a constant or parameter reads its other operand's current order only.
    std::vector<char> hist(NN, 0);
    for (int j = n; j < NN; ++j) {
        if (kind[j] != WK_LIN && kind[j] != WK_MULS)
            for (auto a : S.args[j])
                if (is_node(a) && (int)HYX_OPERAND_INDEX(a) >= n) hist[HYX_OPERAND_INDEX(a)] = 1;
        const int k = kind[j];
        if (k == WK_POW || k == WK_DIV || k == WK_SQRT || k == WK_EXP || k == WK_LOG || k == WK_TRIG) hist[j] = 1;
    }
    // Constants and parameters used by nodes or right-hand sides.
    std::vector<uint32_t> scal;
    auto scal_index = [&](uint32_t w) {
        for (size_t q = 0; q < scal.size(); ++q)
            if (scal[q] == w) return (int)q;
        scal.push_back(w);
        return (int)scal.size() - 1;
    };
    for (int j = n; j < NN; ++j)
        for (auto a : S.args[j])
            if (!is_node(a)) scal_index(a);
    for (int i = 0; i < n; ++i)
        if (!is_node(S.rhs[i])) scal_index(S.rhs[i]);
    const int NS = (int)scal.size();
    // T rows and cur entries.
    std::vector<int> hrow(NN, -1);
    int R = n + 1;
    for (int j = n; j < NN; ++j)
        if (hist[j]) hrow[j] = R++;
    // Constants and parameters: a T row only where a recurrence reads their higher
    // (zero) orders; products with them read the value slot sv[q].
    std::vector<int> srow(NS, -1);
    for (int j = n; j < NN; ++j)
        if (kind[j] != WK_LIN && kind[j] != WK_MULS)
            for (auto a : S.args[j])
                if (!is_node(a) && srow[scal_index(a)] < 0) srow[scal_index(a)] = R++;
    // A sink row: the T stores of slot lanes whose node has no history row go there, so
    // a slot stores without a per-lane branch (nothing reads it).
    const int sink = R++;
    // T holds orders 0..P-1 of every row and order P of the state rows (and the zero row):
    // node coefficients are formed up to order P-1.
    const int CB = P * R + n + 1;
    const int NC = NN + 1 + NS + 1; // the last entry is a sink, like the sink row
    const int csink = NN + 1 + NS;
    const int SVB = CB + NC;
    const int flat = SVB + NS;
    const int ipb = 64 / G;
    if ((long)flat * ipb * 8 > 160 * 1024) {
        why = "the coefficient rows exceed the 160 KB of LDS of a compute unit";
        return "";
    }
    auto trow = [&](uint32_t w) -> int {
        const int idx = (int)HYX_OPERAND_INDEX(w);
        if (!is_node(w)) {
            if (srow[scal_index(w)] < 0) throw std::logic_error("wave form: constant without a row");
            return srow[scal_index(w)];
        }
        if (idx < n) return idx;
        if (hrow[idx] < 0) throw std::logic_error("wave form: operand without a history row");
        return hrow[idx];
    };
    auto cid = [&](uint32_t w) -> int {
        const int idx = (int)HYX_OPERAND_INDEX(w);
        if (!is_node(w)) return NN + 1 + scal_index(w);
        return idx < n ? idx : idx + 1;
    };
    // Levels. A node's level lies between its earliest, 1 + the highest earliest level of
    // its node operands, and its latest, 1 below the lowest latest level of the nodes that
    // read it (the last level for the nodes only right-hand sides read). Levels are
    // assigned in increasing order: at each level, a recurrence kind with ready nodes at
    // their latest level opens slots for them, and the spare lanes of those slots take the
    // kind's other ready nodes, the nearest latest level first; the ready nodes of a kind
    // that opens no slot wait. Fewer, fuller slots mean fewer instructions per order, and
    // the level of a node changes none of its arithmetic.
    // Derived nodes take their root's level; a node that reads a derived node depends on
    // its root.
    std::vector<std::vector<int>> eops(NN);
    for (int j = n; j < NN; ++j)
        if (dsrc[j] < 0)
            for (auto a : S.args[j])
                if (is_node(a) && (int)HYX_OPERAND_INDEX(a) >= n) eops[j].push_back(root[HYX_OPERAND_INDEX(a)]);
    std::vector<int> asap(NN, 0), alap(NN, 0), lvl(NN, 0);
    int maxl = 0;
    for (int j = n; j < NN; ++j) {
        if (dsrc[j] >= 0) continue;
        int l = 1;
        for (int a : eops[j]) l = std::max(l, asap[a] + 1);
        asap[j] = l;
        maxl = std::max(maxl, l);
    }
    for (int j = n; j < NN; ++j) alap[j] = maxl;
    for (int j = NN - 1; j >= n; --j)
        for (int a : eops[j]) alap[a] = std::min(alap[a], alap[j] - 1);
    for (int l = 1; l <= maxl; ++l) {
        std::map<int, std::vector<int>> must, spare;
        for (int j = n; j < NN; ++j) {
            if (lvl[j] || dsrc[j] >= 0) continue;
            bool ready = true;
            for (int a : eops[j]) ready = ready && lvl[a] >= 1 && lvl[a] < l;
            if (ready) (alap[j] == l ? must : spare)[kind[j]].push_back(j);
        }
        for (auto &mk : must) {
            auto &sp = spare[mk.first];
            std::stable_sort(sp.begin(), sp.end(), [&](int x, int y) { return alap[x] < alap[y]; });
            const size_t cap = (mk.second.size() + G - 1) / G * G;
            for (int j : mk.second) lvl[j] = l;
            for (size_t q = 0; q < sp.size() && mk.second.size() + q < cap; ++q) lvl[sp[q]] = l;
        }
    }
    for (int j = n; j < NN; ++j) {
        if (dsrc[j] >= 0) {
            lvl[j] = lvl[root[j]];
            continue;
        }
        if (lvl[j] < asap[j] || lvl[j] > alap[j]) throw std::logic_error("wave form: a node left without a level");
    }
    // Slots: per level and kind, the nodes' derived chains are merged into templates (a
    // sequence of step shapes, each with its source position); a chain fits a template
    // when its steps map in order onto template steps of the same shape whose sources map
    // onto its sources' steps. Lanes run every template step; a step that is not one of
    // their chain's computes on and stores to the sinks. A template grows by appending the
    // steps a chain does not find, up to 6 steps.
    std::vector<WaveSlot> slots;
    std::vector<std::vector<std::pair<std::string, int>>> tmpls; // per slot: (shape, source position)
    std::vector<std::vector<std::vector<int>>> tmap;              // per slot, lane, template step: node or -1
    auto embed = [&](const std::vector<int> &ch, std::vector<std::pair<std::string, int>> &t, bool grow,
                     std::vector<int> *pos) {
        std::vector<int> ps;
        size_t i = 0;
        auto tt = t;
        for (int x : ch) {
            const int sp = dpos(x);
            const std::pair<std::string, int> want(dshape(x), sp == 0 ? 0 : ps[sp - 1] + 1);
            while (i < tt.size() && tt[i] != want) ++i;
            if (i >= tt.size()) {
                if (!grow) return false;
                tt.push_back(want);
                i = tt.size() - 1;
            }
            ps.push_back((int)i++);
        }
        if (grow) {
            if (tt.size() > 6 && tt.size() > t.size()) return false;
            t = tt;
        }
        if (pos) *pos = ps;
        return true;
    };
    for (int l = 1; l <= maxl; ++l)
        for (int kd = 0; kd < WK_COUNT; ++kd) {
            std::vector<int> ns;
            for (int j = n; j < NN; ++j)
                if (lvl[j] == l && kind[j] == kd && dsrc[j] < 0) ns.push_back(j);
            if (ns.empty()) continue;
            // Templates from the longest chains down; nodes without derived nodes fill the
            // spare lanes of the templates' last slots, then slots of their own.
            std::vector<int> byl, bare;
            for (int j : ns) (dlist[j].empty() ? bare : byl).push_back(j);
            std::stable_sort(byl.begin(), byl.end(), [&](int x, int y) { return dlist[x].size() > dlist[y].size(); });
            std::vector<std::vector<std::pair<std::string, int>>> temps;
            std::vector<std::vector<int>> members;
            for (int j : byl) {
                int best = -1;
                size_t best_len = 0;
                for (size_t ti = 0; ti < temps.size(); ++ti) {
                    auto t2 = temps[ti];
                    if (embed(dlist[j], t2, true, nullptr) && (best < 0 || t2.size() < best_len)) {
                        best = (int)ti;
                        best_len = t2.size();
                    }
                }
                if (best < 0) {
                    temps.emplace_back();
                    members.emplace_back();
                    best = (int)temps.size() - 1;
                }
                embed(dlist[j], temps[best], true, nullptr);
                members[best].push_back(j);
            }
            for (size_t ti = 0; ti < temps.size(); ++ti) {
                auto &mb = members[ti];
                std::sort(mb.begin(), mb.end());
                while (!bare.empty() && mb.size() % G) {
                    mb.push_back(bare.front());
                    bare.erase(bare.begin());
                }
            }
            if (!bare.empty()) {
                temps.emplace_back();
                members.push_back(bare);
            }
            for (size_t ti = 0; ti < temps.size(); ++ti) {
                const auto &mb = members[ti];
                for (size_t q0 = 0; q0 < mb.size(); q0 += G) {
                    WaveSlot sl;
                    sl.kind = kd;
                    for (size_t q = q0; q < std::min(mb.size(), q0 + G); ++q) sl.nodes.push_back((uint32_t)mb[q]);
                    for (auto j : sl.nodes) sl.m = std::max(sl.m, (int)S.args[j].size());
                    if (kd == WK_LIN && sl.m > 15) {
                        why = "a sum of more than 15 terms";
                        return "";
                    }
                    std::vector<std::vector<int>> mp;
                    for (auto j : sl.nodes) {
                        std::vector<int> m(temps[ti].size(), -1), ps;
                        if (!embed(dlist[j], temps[ti], false, &ps)) throw std::logic_error("wave form: a chain outside its template");
                        for (size_t e = 0; e < ps.size(); ++e) m[ps[e]] = dlist[j][e];
                        mp.push_back(m);
                    }
                    slots.push_back(sl);
                    tmpls.push_back(temps[ti]);
                    tmap.push_back(mp);
                }
            }
        }
    const int own = (n + G - 1) / G;
    if (own > 8) {
        why = "too many state variables per lane";
        return "";
    }

    // Per-lane tables: int fields w[f][g], double fields d[f][g].
    std::vector<std::vector<int>> W;
    std::vector<std::vector<double>> D;
    auto addw = [&](std::vector<int> v) {
        v.resize(G, v.empty() ? 0 : v.back());
        W.push_back(v);
        return (int)W.size() - 1;
    };
    auto addd = [&](std::vector<double> v) {
        v.resize(G, 0.0);
        D.push_back(v);
        return (int)D.size() - 1;
    };
    std::vector<int> vo((size_t)G * own, -1);
    for (int i = 0; i < n; ++i) vo[(size_t)(i % G) * own + i / G] = i;
    // Right-hand side of each owned variable: its cur entry (the zero entry for none).
    std::vector<int> f_rhs(own);
    for (int o = 0; o < own; ++o) {
        std::vector<int> v(G, n);
        for (int g = 0; g < G; ++g)
            if (vo[(size_t)g * own + o] >= 0) v[g] = cid(S.rhs[vo[(size_t)g * own + o]]);
        f_rhs[o] = addw(v);
    }

    std::ostringstream h, o;
    h << "#define HYX_N " << n << "\n#define HYX_NPAR " << std::max<uint32_t>(S.npar, 1) << "\n#define HYX_P " << P
      << "\n#define HYX_GROUP 1\n#define HYX_WAVE 1\n#define HYX_G " << G << "\n#define HYX_I " << ipb
      << "\n#define HYX_OWN " << own << "\n#define HYX_NX " << R << "\n#define HYX_XB 0\n#define HYX_FLAT " << flat
      << "\n#define HYX_WAVE_SLOTS " << slots.size() << "\n";
    h << "typedef unsigned long long hyx_u64;\n";
    h << "#ifndef HYX_CPU_EMU\n"
         "// A block is one wavefront, and the LDS operations of a wavefront complete in issue\n"
         "// order: the exchanges between levels only need the compiler to keep memory accesses\n"
         "// on their side of the fence (no s_waitcnt, no s_barrier).\n"
         "__device__ __forceinline__ void hyx_wsync()\n{\n"
         "    __builtin_amdgcn_fence(__ATOMIC_RELEASE, \"wavefront\");\n"
         "    __builtin_amdgcn_wave_barrier();\n"
         "    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, \"wavefront\");\n}\n"
         "#endif\n";

    // Slot code. Operand rows are per-lane registers tab.w[f]; T[k][row] is fl[row + k * R].
    auto Tk = [&](const std::string &row, int k) {
        return k == 0 ? "fl[" + row + "]" : "fl[" + row + " + " + std::to_string(k * R) + "]";
    };
    auto cur = [&](const std::string &c) { return "fl[" + std::to_string(CB) + " + " + c + "]"; };
    struct SlotTab {
        int a = -1, b = -1, oc = -1, orow = -1, sgn = -1, al = -1, al1 = -1, cosf = -1;
        std::vector<int> lin;
        bool cur_any = false, hist_any = false;
    };
    std::vector<SlotTab> st(slots.size());
    // Readers of each node's cur entry: linear nodes and right-hand sides (a derived node
    // reads its source from a register).
    std::vector<char> cur_read(NC, 0);
    for (int j = n; j < NN; ++j)
        if (kind[j] == WK_LIN || kind[j] == WK_MULS)
            for (auto a : S.args[j])
                if (!(dsrc[j] >= 0 && is_node(a))) cur_read[cid(a)] = 1;
    for (int i = 0; i < n; ++i) cur_read[cid(S.rhs[i])] = 1;
    // Derived steps of each slot (the shape is the slot's; lanes past its width repeat the
    // first lane's, like its node): a product's constant per lane (d-field), a linear
    // node's constant operands per lane (their cur entries, read at order 0 only: they are
    // zero above), and the stores.
    struct DStep {
        int kd = 0, src = 0, op = 0, cval = -1, oc = -1, orow = -1;
        bool cur_any = false, hist_any = false;
        std::vector<int> cterm; // linear: cur entry field of each constant operand (-1: the source)
    };
    std::vector<std::vector<DStep>> dst(slots.size());
    for (size_t s = 0; s < slots.size(); ++s) {
        const auto &sl = slots[s];
        auto &t = st[s];
        const int w = (int)sl.nodes.size();
        auto lanes = [&](auto f) {
            std::vector<int> v(G, 0);
            for (int q = 0; q < w; ++q) v[q] = f(sl.nodes[q]);
            for (int q = w; q < G; ++q) v[q] = v[0];
            return v;
        };
        for (auto j : sl.nodes) {
            t.hist_any = t.hist_any || hist[j];
            t.cur_any = t.cur_any || cur_read[j + 1];
        }
        if (t.cur_any) t.oc = addw(lanes([&](uint32_t j) { return (int)j + 1; }));
        if (t.hist_any) t.orow = addw(lanes([&](uint32_t j) { return hist[j] ? hrow[j] : sink; }));
        switch (sl.kind) {
        case WK_LIN: {
            for (int q = 0; q < sl.m; ++q)
                t.lin.push_back(addw(lanes([&](uint32_t j) {
                    return q < (int)S.args[j].size() ? cid(S.args[j][q]) : n;
                })));
            t.sgn = addw(lanes([&](uint32_t j) {
                int m = 0;
                for (size_t q = 0; q < S.args[j].size(); ++q)
                    if (S.op[j] == HYX_OP_NEG || (S.op[j] == HYX_OP_SUB && q == 1)) m |= 1 << q;
                return m;
            }));
            break;
        }
        case WK_MULS:
            // a: the cur entry of the operand with a series (if any); b: the constant or parameter.
            t.a = addw(lanes([&](uint32_t j) { return cid(is_node(S.args[j][0]) ? S.args[j][0] : S.args[j][1]); }));
            t.b = addw(lanes([&](uint32_t j) { return scal_index(is_node(S.args[j][0]) ? S.args[j][1] : S.args[j][0]); }));
            break;
        case WK_TIME: break;
        default:
            t.a = addw(lanes([&](uint32_t j) { return trow(S.args[j][0]); }));
            if (sl.kind == WK_MUL || sl.kind == WK_DIV) t.b = addw(lanes([&](uint32_t j) { return trow(S.args[j][1]); }));
            if (sl.kind == WK_TRIG) {
                t.b = addw(lanes([&](uint32_t j) { return hrow[S.aux[j]]; }));
                t.cosf = addw(lanes([&](uint32_t j) { return S.op[j] == HYX_OP_COS ? 1 : 0; }));
            }
            if (sl.kind == WK_POW) {
                // Lanes past the slot's width repeat the first lane's exponent (they compute
                // its node again).
                std::vector<double> al(G, 0.0), al1(G, 0.0);
                for (int q = 0; q < G; ++q) {
                    al[q] = S.consts[S.aux[sl.nodes[q < w ? q : 0]]];
                    al1[q] = al[q] + 1.0;
                }
                t.al = addd(al);
                t.al1 = addd(al1);
            }
            break;
        }
        // Derived steps from the slot's template; lanes past the width repeat the first lane.
        const auto &T = tmpls[s];
        for (size_t d = 0; d < T.size(); ++d) {
            DStep ds;
            const std::string &shp = T[d].first;
            ds.kd = shp[0] == 'M' ? WK_MULS : WK_LIN;
            ds.src = T[d].second;
            auto at = [&](int q) { return tmap[s][q < w ? q : 0][d]; };
            auto dl = [&](auto f) {
                std::vector<int> v(G, 0);
                for (int q = 0; q < G; ++q) v[q] = f(at(q));
                return v;
            };
            for (int q = 0; q < w; ++q)
                if (at(q) >= 0) {
                    ds.cur_any = ds.cur_any || cur_read[at(q) + 1];
                    ds.hist_any = ds.hist_any || hist[at(q)];
                }
            if (ds.kd == WK_MULS) {
                std::vector<double> cv(G, 0.0);
                for (int q = 0; q < G; ++q) {
                    const int x = at(q);
                    if (x < 0) continue;
                    const uint32_t c = is_node(S.args[x][0]) ? S.args[x][1] : S.args[x][0];
                    cv[q] = S.consts[HYX_OPERAND_INDEX(c)];
                }
                ds.cval = addd(cv);
            } else {
                // "L<op>:<n|c per operand>"
                ds.op = std::atoi(shp.c_str() + 1);
                const std::string pat = shp.substr(shp.find(':') + 1);
                for (size_t q = 0; q < pat.size(); ++q)
                    ds.cterm.push_back(pat[q] == 'n' ? -1 : addw(dl([&](int x) { return x < 0 ? n : cid(S.args[x][q]); })));
            }
            if (ds.cur_any) ds.oc = addw(dl([&](int x) { return x >= 0 && cur_read[x + 1] ? x + 1 : csink; }));
            if (ds.hist_any) ds.orow = addw(dl([&](int x) { return x >= 0 && hist[x] ? hrow[x] : sink; }));
            dst[s].push_back(ds);
        }
    }
    // Table fields are unsigned 8-bit values, four per 32-bit word, when every value fits
    // (rows, entries and masks of systems up to ~250 nodes), else 16-bit, two per word:
    // fewer registers for the tables leave the scheduler room to batch a level's loads.
    int tb_bits = 8;
    for (auto &f : W)
        for (int v : f) {
            if (v < 0 || v > 65535) throw std::logic_error("wave form: table value out of 16-bit range");
            if (v > 255) tb_bits = 16;
        }
    const int tb_per = 32 / tb_bits;
    auto wv = [tb_bits, tb_per](int f) {
        const std::string w = "tab.w[" + std::to_string(f / tb_per) + "]";
        const int sh = tb_bits * (f % tb_per);
        if (sh + tb_bits == 32) return "(int)((unsigned)" + w + " >> " + std::to_string(sh) + ")";
        return "(int)(((unsigned)" + w + " >> " + std::to_string(sh) + ") & " + std::to_string((1u << tb_bits) - 1) + "u)";
    };
    auto dv = [](int f) { return "tab.d[" + std::to_string(f) + "]"; };
    // Per-slot registers that live across orders (reciprocals of order-0 values).
    auto rv = [](size_t s, const char *nm) { return std::string(nm) + std::to_string(s); };

    // Level of each slot; slots of one level are independent of each other.
    std::vector<int> slvl(slots.size());
    for (size_t s = 0; s < slots.size(); ++s) slvl[s] = lvl[slots[s].nodes[0]];
    // nm = the sum of the products tm, as one fma chain or, from split_min terms on, as two
    // interleaved chains (even and odd terms) added at the end: half the dependent FP64
    // latency per sum, at the price of a different summation order.
    auto chain = [&](const std::vector<std::pair<std::string, std::string>> &tm, const std::string &nm,
                     const std::string &ind) {
        if (split_min > 0 && (int)tm.size() >= split_min) {
            o << ind << "{\n" << ind << "    double e = " << tm[0].first << " * " << tm[0].second << ";\n" << ind
              << "    double f = " << tm[1].first << " * " << tm[1].second << ";\n";
            for (size_t q = 2; q < tm.size(); ++q)
                o << ind << "    " << (q % 2 ? "f" : "e") << " = __builtin_fma(" << tm[q].first << ", " << tm[q].second
                  << ", " << (q % 2 ? "f" : "e") << ");\n";
            o << ind << "    " << nm << " = e + f;\n" << ind << "}\n";
            return;
        }
        o << ind << nm << " = " << tm[0].first << " * " << tm[0].second << ";\n";
        for (size_t q = 1; q < tm.size(); ++q)
            o << ind << nm << " = __builtin_fma(" << tm[q].first << ", " << tm[q].second << ", " << nm << ");\n";
    };
    int cur_k = 0; // the order being generated (names of the per-order partial sums)
    auto ko = [&](size_t s) { return "ko" + std::to_string(s) + "_" + std::to_string(cur_k); };
    auto kq = [&](size_t s) { return "kq" + std::to_string(s) + "_" + std::to_string(cur_k); };
    // The part of a recurrence at order k >= 1 that reads orders below k only (no
    // same-order dependency): emitted one level ahead of the slot, so its LDS reads
    // overlap the previous level's exchange. Results in ko<s> (and kq<s>).
    // Returns whether ko<s> holds a value.
    std::vector<char> has_old(slots.size(), 0);
    auto emit_old = [&](size_t s, int k) {
        const auto &sl = slots[s];
        const auto &t = st[s];
        const std::string a = t.a >= 0 ? wv(t.a) : "", b = t.b >= 0 ? wv(t.b) : "";
        const std::string c = t.orow >= 0 ? wv(t.orow) : "";
        std::vector<std::pair<std::string, std::string>> tm, tm2;
        has_old[s] = 0;
        switch (sl.kind) {
        case WK_MUL:
            for (int q = 1; q < k; ++q) tm.push_back({Tk(a, q), Tk(b, k - q)});
            break;
        case WK_SQ:
            for (int q = 1; 2 * q < k; ++q) tm.push_back({Tk(a, q), Tk(a, k - q)});
            break;
        case WK_POW:
            for (int q = 1; q < k; ++q) tm.push_back({Tk(a, k - q), Tk(c, q)});
            break;
        case WK_DIV:
            for (int q = 1; q < k; ++q) tm.push_back({Tk(b, q), Tk(c, k - q)});
            break;
        case WK_SQRT:
            for (int q = 1; 2 * q < k; ++q) tm.push_back({Tk(c, q), Tk(c, k - q)});
            break;
        case WK_EXP:
        case WK_TRIG:
            for (int q = 1; q < k; ++q)
                tm.push_back({q == 1 ? Tk(a, 1) : "(" + lit((double)q) + " * " + Tk(a, q) + ")",
                              Tk(sl.kind == WK_EXP ? c : b, k - q)});
            break;
        case WK_LOG:
            for (int q = 1; q < k; ++q) tm.push_back({lit((double)q / k), "(" + Tk(c, q) + " * " + Tk(a, k - q) + ")"});
            break;
        default: break;
        }
        const bool sqrt_mid = sl.kind == WK_SQRT && k % 2 == 0;
        if (tm.empty() && !sqrt_mid) return;
        has_old[s] = 1;
        o << "    double " << ko(s) << " = 0.0";
        if (sl.kind == WK_POW && k >= 2) o << ", " << kq(s) << " = 0.0";
        o << ";\n    {\n";
        if (!tm.empty()) chain(tm, ko(s), "        ");
        if (sl.kind == WK_POW && k >= 2) {
            // S2 = sum_{q=1}^{k-1} a^[k-q] (q c^[q]).
            for (int q = 1; q < k; ++q) tm2.push_back({Tk(a, k - q), q == 1 ? Tk(c, 1) : "(" + lit((double)q) + " * " + Tk(c, q) + ")"});
            chain(tm2, kq(s), "        ");
        }
        if (sl.kind == WK_SQRT) {
            // tot = 2 sum c^[q] c^[k-q] (+ c^[k/2]^2 for even k).
            std::string tot = tm.empty() ? "" : ko(s) + " + " + ko(s);
            if (sqrt_mid) {
                const std::string m = Tk(c, k / 2);
                tot = tot.empty() ? (m + " * " + m) : ("__builtin_fma(" + m + ", " + m + ", " + tot + ")");
            }
            o << "        " << ko(s) << " = " << tot << ";\n";
        }
        o << "    }\n";
    };
    // The rest of the recurrence (the same-order terms) and the stores.
    auto emit_new = [&](size_t s, int k) {
        const auto &sl = slots[s];
        const auto &t = st[s];
        const std::string a = t.a >= 0 ? wv(t.a) : "", b = t.b >= 0 ? wv(t.b) : "";
        const std::string c = t.orow >= 0 ? wv(t.orow) : "";
        const bool old = has_old[s] != 0;
        const std::string O = ko(s);
        o << "    {\n";
        switch (sl.kind) {
        case WK_LIN: {
            o << "        const int sm = " << wv(t.sgn) << ";\n";
            for (int q = 0; q < sl.m; ++q) {
                const std::string x = cur(wv(t.lin[q]));
                o << "        const double l" << q << " = ((sm >> " << q << ") & 1) ? -" << x << " : " << x << ";\n";
            }
            o << "        double v = l0;\n";
            for (int q = 1; q < sl.m; ++q) o << "        v = v + l" << q << ";\n";
            break;
        }
        case WK_MUL:
            if (k == 0)
                o << "        const double v = " << Tk(a, 0) << " * " << Tk(b, 0) << ";\n";
            else
                o << "        const double v = __builtin_fma(" << Tk(a, k) << ", " << Tk(b, 0) << ", "
                  << (old ? "__builtin_fma(" + Tk(a, 0) + ", " + Tk(b, k) + ", " + O + ")" : Tk(a, 0) + " * " + Tk(b, k))
                  << ");\n";
            break;
        case WK_MULS: o << "        const double v = " << cur(a) << " * fl[" << SVB << " + " << b << "];\n"; break;
        case WK_SQ:
            if (k == 0)
                o << "        const double v = " << Tk(a, 0) << " * " << Tk(a, 0) << ";\n";
            else {
                o << "        const double acc = "
                  << (old ? "__builtin_fma(" + Tk(a, 0) + ", " + Tk(a, k) + ", " + O + ")" : Tk(a, 0) + " * " + Tk(a, k))
                  << ";\n";
                if (k % 2)
                    o << "        const double v = acc + acc;\n";
                else
                    o << "        const double v = __builtin_fma(" << Tk(a, k / 2) << ", " << Tk(a, k / 2) << ", acc + acc);\n";
            }
            break;
        case WK_POW:
            if (k == 0) {
                o << "        const double a0 = " << Tk(a, 0) << ";\n        const double v = pow(a0, " << dv(t.al) << ");\n";
                o << "        const double ra = 1.0 / a0;\n        " << rv(s, "raa") << " = ra * " << dv(t.al) << ";\n        "
                  << rv(s, "rab") << " = ra * " << dv(t.al1) << ";\n";
            } else {
                o << "        const double s1 = "
                  << (old ? "__builtin_fma(" + Tk(a, k) + ", " + Tk(c, 0) + ", " + O + ")" : Tk(a, k) + " * " + Tk(c, 0))
                  << ";\n";
                if (k == 1)
                    o << "        const double v = s1 * " << rv(s, "raa") << ";\n";
                else
                    o << "        const double v = __builtin_fma(s1, " << rv(s, "raa") << ", -(" << kq(s) << " * " << rv(s, "rab")
                      << ") * " << lit(1.0 / k) << ");\n";
            }
            break;
        case WK_DIV:
            if (k == 0)
                o << "        const double v = " << Tk(a, 0) << " / " << Tk(b, 0) << ";\n        " << rv(s, "rb") << " = 1.0 / "
                  << Tk(b, 0) << ";\n";
            else
                o << "        const double v = (" << Tk(a, k) << " - "
                  << (old ? "__builtin_fma(" + Tk(b, k) + ", " + Tk(c, 0) + ", " + O + ")" : Tk(b, k) + " * " + Tk(c, 0))
                  << ") * " << rv(s, "rb") << ";\n";
            break;
        case WK_SQRT:
            if (k == 0)
                o << "        const double v = sqrt(" << Tk(a, 0) << ");\n        " << rv(s, "rs") << " = 0.5 / v;\n";
            else
                o << "        const double v = (" << Tk(a, k) << (old ? " - (" + O + ")" : "") << ") * " << rv(s, "rs") << ";\n";
            break;
        case WK_EXP:
        case WK_TRIG: {
            const std::string other = sl.kind == WK_EXP ? c : b;
            if (k == 0) {
                if (sl.kind == WK_EXP)
                    o << "        const double v = exp(" << Tk(a, 0) << ");\n";
                else
                    o << "        double sn, cs;\n        sincos(" << Tk(a, 0) << ", &sn, &cs);\n        const double v = "
                      << wv(t.cosf) << " ? cs : sn;\n";
                break;
            }
            const std::string d = k == 1 ? Tk(a, 1) : "(" + lit((double)k) + " * " + Tk(a, k) + ")";
            o << "        const double acc = "
              << (old ? "__builtin_fma(" + d + ", " + Tk(other, 0) + ", " + O + ")" : d + " * " + Tk(other, 0)) << ";\n";
            if (sl.kind == WK_EXP)
                o << "        const double v = acc * " << lit(1.0 / k) << ";\n";
            else
                o << "        const double u = acc * " << lit(1.0 / k) << ";\n        const double v = " << wv(t.cosf)
                  << " ? -u : u;\n";
            break;
        }
        case WK_LOG:
            if (k == 0)
                o << "        const double v = log(" << Tk(a, 0) << ");\n        " << rv(s, "rl") << " = 1.0 / " << Tk(a, 0)
                  << ";\n";
            else
                o << "        const double v = (" << Tk(a, k) << (old ? " - " + O : "") << ") * " << rv(s, "rl") << ";\n";
            break;
        case WK_TIME: o << "        const double v = " << (k == 0 ? "t0" : k == 1 ? "1.0" : "0.0") << ";\n"; break;
        }
        if (t.cur_any) o << "        " << cur(wv(t.oc)) << " = v;\n";
        if (t.hist_any) o << "        " << Tk(wv(t.orow), k) << " = v;\n";
        // Derived steps: the recurrences of the nodes they replace (a product with a
        // constant at every order; a linear node's constant operands are zero above order 0).
        for (size_t d = 0; d < dst[s].size(); ++d) {
            const auto &ds = dst[s][d];
            const std::string dn = "dv" + std::to_string(d);
            const std::string src = ds.src == 0 ? "v" : "dv" + std::to_string(ds.src - 1);
            if (ds.kd == WK_MULS)
                o << "        const double " << dn << " = " << src << " * " << dv(ds.cval) << ";\n";
            else {
                for (size_t q = 0; q < ds.cterm.size(); ++q) {
                    const std::string y = ds.cterm[q] < 0 ? src : k == 0 ? cur(wv(ds.cterm[q])) : std::string("0.0");
                    const bool neg = ds.op == HYX_OP_NEG || (ds.op == HYX_OP_SUB && q == 1);
                    o << "        const double " << dn << "_" << q << " = " << (neg ? "-" : "") << y << ";\n";
                }
                o << "        double " << dn << " = " << dn << "_0;\n";
                for (size_t q = 1; q < ds.cterm.size(); ++q) o << "        " << dn << " = " << dn << " + " << dn << "_" << q << ";\n";
            }
            if (ds.cur_any) o << "        " << cur(wv(ds.oc)) << " = " << dn << ";\n";
            if (ds.hist_any) o << "        " << Tk(wv(ds.orow), k) << " = " << dn << ";\n";
        }
        o << "    }\n";
    };
    // Every lane runs every slot: lanes past a slot's width hold the first lane's table
    // entries, so they compute the same value and store it to the same address (no
    // branch per slot or per lane). A level is then one straight-line block whose LDS
    // loads the compiler can issue together behind one wait.
    // Order k: per level, the order-k-independent parts of the next level's slots, then
    // this level's slots, then one fence.
    auto emit_order_slots = [&](int k) {
        cur_k = k;
        size_t s0 = 0;
        int maxlv = 0;
        for (size_t s = 0; s < slots.size(); ++s) maxlv = std::max(maxlv, slvl[s]);
        for (int l = 1; l <= maxlv; ++l) {
            if (k > 0)
                for (size_t s = 0; s < slots.size(); ++s)
                    if (slvl[s] == l + 1) emit_old(s, k);
            for (size_t s = s0; s < slots.size() && slvl[s] == l; ++s) emit_new(s, k);
            while (s0 < slots.size() && slvl[s0] == l) ++s0;
            o << "    hyx_wsync();\n";
        }
    };

    // The per-lane table struct and its loader.
    const size_t NWP = (W.size() + tb_per - 1) / tb_per;
    h << "struct hyx_gtab {\n    int vo[HYX_OWN];\n    int vc[HYX_OWN];\n    int w[" << std::max<size_t>(NWP, 1) << "];\n    double d["
      << std::max<size_t>(D.size(), 1) << "];\n};\n";
    {
        std::vector<int> wf;
        for (size_t p2 = 0; p2 < NWP; ++p2)
            for (int g = 0; g < G; ++g) {
                unsigned word = 0;
                for (int h2 = 0; h2 < tb_per; ++h2) {
                    const size_t f = p2 * tb_per + h2;
                    if (f < W.size()) word |= ((unsigned)W[f][g]) << (tb_bits * h2);
                }
                wf.push_back((int)word);
            }
        if (wf.empty()) wf.push_back(0);
        h << "__constant__ int hyx_wtab_w[" << wf.size() << "] = {";
        for (size_t q = 0; q < wf.size(); ++q) h << (q ? "," : "") << wf[q];
        h << "};\n";
        std::vector<double> df;
        for (auto &f : D) df.insert(df.end(), f.begin(), f.end());
        if (df.empty()) df.push_back(0.0);
        h << "__constant__ double hyx_wtab_d[" << df.size() << "] = {";
        for (size_t q = 0; q < df.size(); ++q) h << (q ? "," : "") << lit(df[q]);
        h << "};\n";
        h << "__constant__ int hyx_wtab_vo[" << vo.size() << "] = {";
        for (size_t q = 0; q < vo.size(); ++q) h << (q ? "," : "") << vo[q];
        h << "};\n";
    }
    h << "__device__ __forceinline__ void hyx_gtab_load(hyx_gtab &t, int g)\n{\n"
         "    for (int q = 0; q < HYX_OWN; ++q) { t.vo[q] = hyx_wtab_vo[g * HYX_OWN + q]; t.vc[q] = t.vo[q] >= 0 ? t.vo[q] : HYX_N; }\n";
    if (NWP) h << "    for (int f = 0; f < " << NWP << "; ++f) t.w[f] = hyx_wtab_w[f * HYX_G + g];\n";
    if (!D.empty()) h << "    for (int f = 0; f < " << D.size() << "; ++f) t.d[f] = hyx_wtab_d[f * HYX_G + g];\n";
    h << "}\n";

    auto xv = [](int oo, int k) { return "x" + std::to_string(oo) + "_" + std::to_string(k); };
    auto vcol = [](int oo) { return "vc" + std::to_string(oo); };
    o << "template <bool WTC, class TAB>\n"
         "__device__ __forceinline__ void hyx_gstep(double (&xo)[HYX_OWN], const double (&pv)[HYX_NPAR], double t0, "
         "double max_dt, double &h_out, bool &nf_out, int g, double *fl, const TAB &tab, double *__restrict__ tc, "
         "hyx_u64 tcs"
      << (E ? ", unsigned long long &hyx_evm, double &hyx_mx" : "") << ")\n{\n";
    o << "    (void)pv; (void)t0; (void)tc; (void)tcs; (void)g;\n";
    for (size_t s = 0; s < slots.size(); ++s) {
        const int kd = slots[s].kind;
        if (kd == WK_POW) o << "    double " << rv(s, "raa") << " = 0.0, " << rv(s, "rab") << " = 0.0;\n";
        if (kd == WK_DIV) o << "    double " << rv(s, "rb") << " = 0.0;\n";
        if (kd == WK_SQRT) o << "    double " << rv(s, "rs") << " = 0.0;\n";
        if (kd == WK_LOG) o << "    double " << rv(s, "rl") << " = 0.0;\n";
    }
    for (int oo = 0; oo < own; ++oo)
        o << "    const int " << vcol(oo) << " = (tab.vo[" << oo << "] >= 0) ? tab.vo[" << oo << "] : HYX_N;\n";
    o << "    hyx_wsync();\n";
    // Order 0: state rows, constants and parameters (lane 0), then the slots.
    for (int oo = 0; oo < own; ++oo) {
        o << "    const double " << xv(oo, 0) << " = xo[" << oo << "];\n";
        o << "    fl[" << vcol(oo) << "] = " << xv(oo, 0) << ";\n";
        o << "    " << cur(vcol(oo)) << " = " << xv(oo, 0) << ";\n";
    }
    if (NS) {
        o << "    if (g == 0) {\n";
        for (int q = 0; q < NS; ++q) {
            const uint32_t w = scal[q];
            const std::string val = HYX_OPERAND_KIND(w) == HYX_KIND_CONST ? lit(S.consts[HYX_OPERAND_INDEX(w)])
                                                                           : "pv[" + std::to_string(HYX_OPERAND_INDEX(w)) + "]";
            if (srow[q] >= 0) o << "        fl[" << srow[q] << "] = " << val << ";\n";
            o << "        fl[" << (CB + NN + 1 + q) << "] = " << val << ";\n";
            o << "        fl[" << (SVB + q) << "] = " << val << ";\n";
        }
        o << "    }\n";
    }
    o << "    hyx_wsync();\n";
    emit_order_slots(0);
    for (int k = 1; k <= P; ++k) {
        o << "    // order " << k << "\n";
        cur_k = k;
        if (k < P)
            for (size_t s = 0; s < slots.size(); ++s)
                if (slvl[s] == 1) emit_old(s, k);
        for (int oo = 0; oo < own; ++oo) o << "    const double r" << oo << "_" << k << " = " << cur(wv(f_rhs[oo])) << ";\n";
        o << "    hyx_wsync();\n";
        for (int oo = 0; oo < own; ++oo) {
            o << "    const double " << xv(oo, k) << " = r" << oo << "_" << k << " * " << lit(1.0 / k) << ";\n";
            o << "    fl[" << vcol(oo) << " + " << k * R << "] = " << xv(oo, k) << ";\n";
            if (k < P) o << "    " << cur(vcol(oo)) << " = " << xv(oo, k) << ";\n";
        }
        if (k == 1 && NS) {
            o << "    if (g == 0) {\n";
            for (int q = 0; q < NS; ++q) o << "        fl[" << (CB + NN + 1 + q) << "] = 0.0;\n";
            o << "    }\n";
        }
        o << "    hyx_wsync();\n";
        if (k < P) emit_order_slots(k);
    }
    o << "    if constexpr (WTC) if (tc) {\n";
    for (int oo = 0; oo < own; ++oo) {
        o << "        if (tab.vo[" << oo << "] >= 0)\n";
        o << "            for (int k = 0; k <= HYX_P; ++k) tc[((hyx_u64)tab.vo[" << oo << "] * " << (P + 1)
          << " + k) * tcs] = fl[HYX_XB + k * HYX_NX + " << vcol(oo) << "];\n";
    }
    o << "    }\n";
    if (E) {
        Gen eg(*E, P, high_acc);
        eg.lds_xb = 0;
        eg.lds_nx = R;
        o << "    double hyx_evg[" << nev << "][" << (P + 1) << "];\n    {\n";
        for (uint32_t j = E->n; j < E->nnodes; ++j) eg.order0(j);
        for (int k = 1; k <= P; ++k)
            for (uint32_t j = E->n; j < E->nnodes; ++j) eg.orderk(j, k);
        for (uint32_t e = 0; e < nev; ++e)
            for (int k = 0; k <= P; ++k)
                eg.o << "    hyx_evg[" << e << "][" << k << "] = " << (eg.zero(E->rhs[e], k) ? "0.0" : eg.val(E->rhs[e], k))
                     << ";\n";
        o << eg.o.str() << "    }\n";
    }
    o << "    double mx = 0.0, mo = 0.0, mom1 = 0.0, ab;\n";
    for (int oo = 0; oo < own; ++oo) {
        o << "    ab = fabs(" << xv(oo, 0) << "); mx = (ab > mx) ? ab : mx;\n";
        o << "    ab = fabs(" << xv(oo, P) << "); mo = (ab > mo) ? ab : mo;\n";
        o << "    ab = fabs(" << xv(oo, P - 1) << "); mom1 = (ab > mom1) ? ab : mom1;\n";
    }
    for (uint32_t e = 0; E && e < nev; ++e) {
        o << "    ab = fabs(hyx_evg[" << e << "][" << P << "]); mo = (ab > mo) ? ab : mo;\n";
        o << "    ab = fabs(hyx_evg[" << e << "][" << (P - 1) << "]); mom1 = (ab > mom1) ? ab : mom1;\n";
    }
    o << "#pragma unroll\n    for (int off = 1; off < HYX_G; off <<= 1) {\n"
         "        ab = __shfl_xor(mx, off); mx = (ab > mx) ? ab : mx;\n"
         "        ab = __shfl_xor(mo, off); mo = (ab > mo) ? ab : mo;\n"
         "        ab = __shfl_xor(mom1, off); mom1 = (ab > mom1) ? ab : mom1;\n    }\n";
    o << "    const double num = (mx <= 1.0) ? 1.0 : mx;\n";
    o << "    const double rho_o = pow(num / mo, " << lit(1.0 / P) << ");\n";
    o << "    const double rho_om1 = pow(num / mom1, " << lit(1.0 / (P - 1)) << ");\n";
    o << "    const double rho_m = (rho_om1 < rho_o) ? rho_om1 : rho_o;\n";
    const double e1 = std::exp(1.0);
    const double rhofac = std::exp(-7.0 / 10.0 / (double)(P - 1)) / (e1 * e1);
    o << "    double h = rho_m * " << lit(rhofac) << ";\n";
    o << "    const double lim = fabs(max_dt);\n    h = (h < lim) ? h : lim;\n    if (max_dt < 0.0) h = -h;\n";
    if (E) {
        const uint64_t off = (uint64_t)n * (P + 1) + (uint64_t)(E->nnodes - E->n) * (P + 1);
        o << "    hyx_evm = 0ull;\n";
        o << "#pragma unroll\n    for (int e = 0; e < " << nev << "; ++e) {\n"
             "        double hk = 1.0, emag = 0.0, lo = 0.0, hi = 0.0;\n"
             "#pragma unroll\n        for (int k = 0; k <= " << P << "; ++k) {\n"
             "            const double c = hyx_evg[e][k] * hk;\n            hk *= h;\n            emag += fabs(c);\n"
             "            if (k > 0) { lo += fmin(c, 0.0); hi += fmax(c, 0.0); }\n        }\n"
             "        const double slack = " << lit(4.0 * (P + 1) * 1.1102230246251565e-16) << " * emag;\n"
             "        const double c0 = hyx_evg[e][0];\n"
             "        if (emag > 0.0 && isfinite(emag) && !(c0 + lo > slack || c0 + hi < -slack)) hyx_evm |= 1ull << e;\n"
             "    }\n";
        o << "    if (hyx_evm != 0ull && tc) {\n";
        for (uint32_t e = 0; e < nev; ++e)
            for (int k = 0; k <= P; ++k)
                o << "        tc[(hyx_u64)" << (off + (uint64_t)e * (P + 1) + k) << " * tcs] = hyx_evg[" << e << "][" << k
                  << "];\n";
        o << "        if constexpr (!WTC) {\n";
        for (int oo = 0; oo < own; ++oo) {
            o << "            if (tab.vo[" << oo << "] >= 0)\n";
            o << "                for (int k = 0; k <= HYX_P; ++k) tc[((hyx_u64)tab.vo[" << oo << "] * " << (P + 1)
              << " + k) * tcs] = fl[HYX_XB + k * HYX_NX + " << vcol(oo) << "];\n";
        }
        o << "        }\n    }\n";
        o << "    hyx_mx = mx;\n";
    }
    o << "    int nf = 0;\n";
    for (int oo = 0; oo < own; ++oo) {
        if (!high_acc) {
            o << "    { double r = " << xv(oo, P) << ";\n";
            o << "#pragma unroll\n      for (int k = HYX_P - 1; k >= 0; --k) r = __builtin_fma(r, h, fl[HYX_XB + k * HYX_NX + "
              << vcol(oo) << "]);\n";
        } else {
            o << "    { double r = " << xv(oo, 0) << ", cp = 0.0, hk = 1.0;\n";
            o << "#pragma unroll\n      for (int k = 1; k <= HYX_P; ++k) { hk *= h; const double y = fl[HYX_XB + k * HYX_NX + "
              << vcol(oo) << "] * hk - cp; const double tt = r + y; cp = (tt - r) - y; r = tt; }\n";
        }
        o << "      xo[" << oo << "] = r; nf |= (tab.vo[" << oo << "] >= 0 && !isfinite(r)) ? 1 : 0; }\n";
    }
    o << "#pragma unroll\n    for (int off = 1; off < HYX_G; off <<= 1) nf |= __shfl_xor(nf, off);\n";
    o << "    h_out = h;\n    nf_out = nf != 0;\n}\n";
    return h.str() + o.str();
}

} // namespace hyx
