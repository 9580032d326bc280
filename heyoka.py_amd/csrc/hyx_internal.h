// Internal declarations shared by the runtime's translation units.
#ifndef HYX_INTERNAL_H
#define HYX_INTERNAL_H

#include <cstdint>
#include <string>
#include <vector>

namespace hyx
{

// A parsed op-list (include/hyx_oplist.h).
struct Sys {
    uint32_t n = 0, npar = 0, nnodes = 0, nconsts = 0;
    std::vector<uint32_t> op, aux;               // per node (state variables: 0)
    std::vector<std::vector<uint32_t>> args;     // per node, packed operands
    std::vector<uint32_t> rhs;                   // per state variable
    std::vector<double> consts;
    std::vector<int> deg;                        // highest possibly non-zero order, per node
};

int parse_oplist(const uint32_t *w, uint32_t n_words, const double *consts, uint32_t n_consts, Sys &S,
                 std::string &err);

// HIP source of hyx_taylor_step<WTC>() for S at the given order, unrolled or compact.
std::string generate_prelude(const Sys &S, int order, bool high_acc, bool compact);
// Number of nodes whose coefficient history the recurrences read.
int history_nodes(const Sys &S);
// Default choice of code form: compact when the histories exceed the register file.
bool choose_compact(const Sys &S, int order);

// Exact hexadecimal double literal.
std::string lit(double v);

} // namespace hyx

#endif
