/* CPU oracle for the Taylor integrator: double and __float128 instantiations.
 *
 * TEST INFRASTRUCTURE ONLY (parity checker and CPU baseline). See
 * taylor_oracle_impl.h for what is restated from the reference and where.
 *
 * Outcome codes are heyoka's taylor_outcome values (heyoka/core.cpp:324-329):
 * success = -2^32-1 and time_limit = -2^32-3 are printed in
 * doc/notebooks/Batch mode overview.ipynb:242,320; step_limit = -2^32-2 and
 * err_nf_state = -2^32-4 follow the same enumeration (their values are not
 * printed anywhere in the reference tree).
 */
#include <math.h>
#include <pthread.h>
#include <quadmath.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/hyx_oplist.h"

#define ORA_SUCCESS (-4294967297LL)
#define ORA_STEP_LIMIT (-4294967298LL)
#define ORA_TIME_LIMIT (-4294967299LL)
#define ORA_ERR_NF_STATE (-4294967300LL)

/* ---- double ---- */
#define REAL double
#define SFX f64
#define R_SIN sin
#define R_COS cos
#define R_EXP exp
#define R_LOG log
#define R_POW pow
#define R_SQRT sqrt
#define R_FABS fabs
#define R_ISFINITE(x) isfinite(x)
#include "taylor_oracle_impl.h"
#undef REAL
#undef SFX
#undef R_SIN
#undef R_COS
#undef R_EXP
#undef R_LOG
#undef R_POW
#undef R_SQRT
#undef R_FABS
#undef R_ISFINITE

/* ---- __float128 (the real128 oracle) ---- */
#define REAL __float128
#define SFX f128
#define R_SIN sinq
#define R_COS cosq
#define R_EXP expq
#define R_LOG logq
#define R_POW powq
#define R_SQRT sqrtq
#define R_FABS fabsq
#define R_ISFINITE(x) finiteq(x)
#include "taylor_oracle_impl.h"
