/* Op-list format: the compact lowering of a Taylor decomposition.
 *
 * This is the data format that crosses the C-ABI in hyx_sys_create() (hyx.h).
 * It encodes the same information as heyoka's Taylor decomposition
 * (the `decomposition` property, heyoka/expose_batch_integrators.cpp:231):
 * n state-variable definitions u_0..u_{n-1}, then the elementary operations
 * u_n..u_{N-1} in dependency order, then one right-hand-side operand per state
 * variable (x_i' = rhs_i).
 *
 * Layout (uint32 words):
 *   [0] HYX_OPLIST_MAGIC
 *   [1] n_state   [2] n_par   [3] n_nodes (state variables included)
 *   [4] n_consts  [5] word offset of the rhs section
 *   [6 ...]       one record per node j in [n_state, n_nodes):
 *                 { opcode, nargs, aux, arg_0, ..., arg_{nargs-1} }
 *   [rhs ...]     n_state operands
 * Constants are passed beside the words as a double array of n_consts values.
 *
 * An operand packs a kind in bits 30-31 and an index in bits 0-29:
 *   HYX_KIND_NODE  -> u_index (a state variable or an earlier node)
 *   HYX_KIND_CONST -> index into the constants array
 *   HYX_KIND_PAR   -> runtime parameter index (par[i])
 * `aux` is the exponent's constant index for HYX_OP_POW and the partner node
 * (cos for sin, sin for cos) for HYX_OP_SIN / HYX_OP_COS; 0 otherwise.
 */
#ifndef HYX_OPLIST_H
#define HYX_OPLIST_H

#define HYX_OPLIST_MAGIC 0x31585948u /* "HYX1" */

enum hyx_opcode {
    HYX_OP_ADD = 1,
    HYX_OP_SUB = 2,
    HYX_OP_NEG = 3,
    HYX_OP_MUL = 4,
    HYX_OP_DIV = 5,
    HYX_OP_SQUARE = 6,
    HYX_OP_SQRT = 7,
    HYX_OP_POW = 8,
    HYX_OP_EXP = 9,
    HYX_OP_LOG = 10,
    HYX_OP_SIN = 11,
    HYX_OP_COS = 12,
    HYX_OP_TIME = 13,
    HYX_OP_SUM = 14
};

enum hyx_operand_kind { HYX_KIND_NODE = 0, HYX_KIND_CONST = 1, HYX_KIND_PAR = 2 };

#define HYX_OPERAND_KIND(w) ((w) >> 30)
#define HYX_OPERAND_INDEX(w) ((w)&0x3fffffffu)
#define HYX_OPERAND(kind, idx) ((((unsigned)(kind)) << 30) | ((unsigned)(idx)))

#endif
