"""Compiled CPU port of the integrator for one system (the CPU baseline).

TEST / BENCHMARK INFRASTRUCTURE ONLY: bench.py's cpu_baseline leg times it;
tests/test_oracle_cport.py checks it against the interpreter oracle.

heyoka runs a system through LLVM-compiled straight-line code over SIMD
batches. The oracle interpreter (taylor_oracle_impl.h) dispatches on the op
code per node, order and lane, which is far slower. This module generates C
for one system's Taylor-coefficient recurrences, with the orders as a loop,
the node list unrolled inside it and the batch lanes as the innermost
fixed-width loop (vectorised by gcc). Operands whose series stop at order 0 are
specialised as heyoka does (a product with a number costs one multiply). The
file then includes the oracle's step-size rule, Horner update and
propagate loop, and gcc builds it with -O3 -march=native -ffp-contract=off on
the machine that runs it.
"""

from __future__ import annotations

import ctypes as ct
import hashlib
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys_path_pkg = os.path.dirname(HERE)

KIND_NODE, KIND_CONST, KIND_PAR = 0, 1, 2
OP = dict(ADD=1, SUB=2, NEG=3, MUL=4, DIV=5, SQUARE=6, SQRT=7, POW=8, EXP=9, LOG=10, SIN=11, COS=12, TIME=13, SUM=14)


def _lit(v):
    return f"({float(v).hex()})"


def generate(dec, order, L):
    n = dec.n_state
    nn = dec.n_nodes
    deg = dec.degrees()
    out = []
    w = out.append
    w(f"#define L {L}\n#define P {order}\n")
    w("#define T(node, k) (tape + ((size_t)(node) * (P + 1) + (k)) * L)\n")
    w("static void gen_coeffs(double *restrict tape, const double *restrict pars, const double *restrict x, "
      "const double *restrict t0, uint32_t Lrt)\n{\n  (void)Lrt;\n")

    def kind(a):
        return a >> 30

    def idx(a):
        return a & 0x3FFFFFFF

    def scalar(a):
        return kind(a) != KIND_NODE or deg[idx(a)] == 0

    def v0(a, lane="l"):
        if kind(a) == KIND_NODE:
            return f"T({idx(a)}, 0)[{lane}]"
        if kind(a) == KIND_CONST:
            return _lit(dec.consts[idx(a)])
        return f"pars[{idx(a)} * L + {lane}]"

    def vk(a, k):
        return f"T({idx(a)}, {k})[l]" if kind(a) == KIND_NODE else "0.0"

    w(f"  for (int i = 0; i < {n}; ++i) for (int l = 0; l < L; ++l) T(i, 0)[l] = x[i * L + l];\n")
    for j, nd in enumerate(dec.nodes):
        u = n + j
        a = nd.args
        c = f"T({u}, 0)[l]"
        op = nd.op
        if op == OP["ADD"]:
            e = f"{v0(a[0])} + {v0(a[1])}"
        elif op == OP["SUB"]:
            e = f"{v0(a[0])} - {v0(a[1])}"
        elif op == OP["NEG"]:
            e = f"-{v0(a[0])}"
        elif op == OP["SUM"]:
            e = " + ".join(v0(q) for q in a)
        elif op == OP["MUL"]:
            e = f"{v0(a[0])} * {v0(a[1])}"
        elif op == OP["DIV"]:
            e = f"{v0(a[0])} / {v0(a[1])}"
        elif op == OP["SQUARE"]:
            e = f"{v0(a[0])} * {v0(a[0])}"
        elif op == OP["SQRT"]:
            e = f"sqrt({v0(a[0])})"
        elif op == OP["POW"]:
            e = f"pow({v0(a[0])}, {_lit(dec.consts[nd.aux])})"
        elif op == OP["EXP"]:
            e = f"exp({v0(a[0])})"
        elif op == OP["LOG"]:
            e = f"log({v0(a[0])})"
        elif op == OP["SIN"]:
            e = f"sin({v0(a[0])})"
        elif op == OP["COS"]:
            e = f"cos({v0(a[0])})"
        else:
            e = "t0[l]"
        w(f"  for (int l = 0; l < L; ++l) {c} = {e};\n")
    w("  for (int k = 1; k <= P; ++k) {\n")
    for i in range(n):
        r = dec.rhs[i]
        if kind(r) == KIND_NODE:
            w(f"    for (int l = 0; l < L; ++l) T({i}, k)[l] = T({idx(r)}, k - 1)[l] / (double)k;\n")
        else:
            w(f"    for (int l = 0; l < L; ++l) T({i}, k)[l] = (k == 1 ? {v0(r)} : 0.0) / (double)k;\n")
    w("    if (k == P) break;\n")
    for j, nd in enumerate(dec.nodes):
        u = n + j
        a = nd.args
        op = nd.op
        c = f"T({u}, k)[l]"
        if deg[u] == 0:
            w(f"    for (int l = 0; l < L; ++l) {c} = 0.0;\n")
            continue
        if op in (OP["ADD"], OP["SUB"], OP["NEG"], OP["SUM"]):
            terms = []
            for q, x_ in enumerate(a):
                if scalar(x_):
                    continue
                neg = op == OP["NEG"] or (op == OP["SUB"] and q == 1)
                terms.append(("-" if neg else "+", vk(x_, "k")))
            if not terms:
                e = "0.0"
            else:
                e = ("-" if terms[0][0] == "-" else "") + terms[0][1]
                for sg, t in terms[1:]:
                    e += f" {sg} {t}"
            w(f"    for (int l = 0; l < L; ++l) {c} = {e};\n")
        elif op == OP["MUL"]:
            if scalar(a[0]) and scalar(a[1]):
                w(f"    for (int l = 0; l < L; ++l) {c} = 0.0;\n")
            elif scalar(a[0]):
                w(f"    for (int l = 0; l < L; ++l) {c} = {v0(a[0])} * {vk(a[1], 'k')};\n")
            elif scalar(a[1]):
                w(f"    for (int l = 0; l < L; ++l) {c} = {vk(a[0], 'k')} * {v0(a[1])};\n")
            else:
                w(f"    {{ double acc[L] = {{0}}; for (int i = 0; i <= k; ++i) for (int l = 0; l < L; ++l) "
                  f"acc[l] += T({idx(a[0])}, i)[l] * T({idx(a[1])}, k - i)[l]; for (int l = 0; l < L; ++l) {c} = acc[l]; }}\n")
        elif op == OP["DIV"]:
            num = "0.0" if scalar(a[0]) else vk(a[0], "k")
            if scalar(a[1]):
                w(f"    for (int l = 0; l < L; ++l) {c} = {num} / {v0(a[1])};\n")
            else:
                w(f"    {{ double acc[L] = {{0}}; for (int i = 1; i <= k; ++i) for (int l = 0; l < L; ++l) "
                  f"acc[l] += T({idx(a[1])}, i)[l] * T({u}, k - i)[l]; for (int l = 0; l < L; ++l) "
                  f"{c} = ({num} - acc[l]) / T({idx(a[1])}, 0)[l]; }}\n")
        elif op == OP["SQUARE"]:
            b = idx(a[0])
            w(f"    {{ double acc[L] = {{0}}; for (int i = 0; 2 * i < k; ++i) for (int l = 0; l < L; ++l) "
              f"acc[l] += T({b}, i)[l] * T({b}, k - i)[l]; for (int l = 0; l < L; ++l) {{ acc[l] = acc[l] + acc[l]; "
              f"if (k % 2 == 0) {{ double m = T({b}, k / 2)[l]; acc[l] += m * m; }} {c} = acc[l]; }} }}\n")
        elif op == OP["SQRT"]:
            num = "0.0" if scalar(a[0]) else vk(a[0], "k")
            w(f"    {{ double acc[L] = {{0}}; for (int i = 1; 2 * i < k; ++i) for (int l = 0; l < L; ++l) "
              f"acc[l] += T({u}, i)[l] * T({u}, k - i)[l]; for (int l = 0; l < L; ++l) {{ acc[l] = acc[l] + acc[l]; "
              f"if (k % 2 == 0) {{ double m = T({u}, k / 2)[l]; acc[l] += m * m; }} {c} = ({num} - acc[l]) / (2 * T({u}, 0)[l]); }} }}\n")
        elif op == OP["POW"]:
            al = dec.consts[nd.aux]
            b = idx(a[0])
            w(f"    {{ double acc[L] = {{0}}; for (int i = 0; i < k; ++i) {{ const double cf = (double)k * {_lit(al)} - "
              f"(double)i * ({_lit(al)} + 1); for (int l = 0; l < L; ++l) acc[l] += cf * T({b}, k - i)[l] * T({u}, i)[l]; }} "
              f"for (int l = 0; l < L; ++l) {c} = acc[l] / ((double)k * T({b}, 0)[l]); }}\n")
        elif op == OP["EXP"]:
            b = idx(a[0])
            w(f"    {{ double acc[L] = {{0}}; for (int i = 1; i <= k; ++i) for (int l = 0; l < L; ++l) "
              f"acc[l] += (double)i * T({b}, i)[l] * T({u}, k - i)[l]; for (int l = 0; l < L; ++l) {c} = acc[l] / (double)k; }}\n")
        elif op == OP["LOG"]:
            b = idx(a[0])
            w(f"    {{ double acc[L] = {{0}}; for (int i = 1; i < k; ++i) for (int l = 0; l < L; ++l) "
              f"acc[l] += (double)i * T({u}, i)[l] * T({b}, k - i)[l]; for (int l = 0; l < L; ++l) "
              f"{c} = (T({b}, k)[l] - acc[l] / (double)k) / T({b}, 0)[l]; }}\n")
        elif op in (OP["SIN"], OP["COS"]):
            b = idx(a[0])
            sg = "" if op == OP["SIN"] else "-"
            w(f"    {{ double acc[L] = {{0}}; for (int i = 1; i <= k; ++i) for (int l = 0; l < L; ++l) "
              f"acc[l] += (double)i * T({b}, i)[l] * T({nd.aux}, k - i)[l]; for (int l = 0; l < L; ++l) "
              f"{{ double v = acc[l] / (double)k; {c} = {sg}v; }} }}\n")
        else:  # time
            w(f"    for (int l = 0; l < L; ++l) {c} = (k == 1) ? 1.0 : 0.0;\n")
    w("  }\n}\n#undef T\n#undef L\n#undef P\n")
    return "".join(out)


_cache = {}


def build(dec, order, L=8):
    words, consts = dec.to_words(), dec.consts_array()
    key = hashlib.sha1(words.tobytes() + consts.tobytes() + f"{order}-{L}".encode()).hexdigest()[:16]
    if key in _cache:
        return _cache[key]
    src = ('#include <math.h>\n#include <pthread.h>\n#include <stdint.h>\n#include <stdlib.h>\n#include <string.h>\n'
           '#include "../include/hyx_oplist.h"\n'
           "#define ORA_SUCCESS (-4294967297LL)\n#define ORA_STEP_LIMIT (-4294967298LL)\n"
           "#define ORA_TIME_LIMIT (-4294967299LL)\n#define ORA_ERR_NF_STATE (-4294967300LL)\n"
           + generate(dec, order, L)
           + "#define ORA_GENERATED_COEFFS(tape, pars, x, t0, Lrt) gen_coeffs(tape, pars, x, t0, Lrt)\n"
           "#define REAL double\n#define SFX f64\n#define R_SIN sin\n#define R_COS cos\n#define R_EXP exp\n"
           "#define R_LOG log\n#define R_POW pow\n#define R_SQRT sqrt\n#define R_FABS fabs\n"
           "#define R_ISFINITE(x) isfinite(x)\n#include \"taylor_oracle_impl.h\"\n")
    bdir = os.path.join(HERE, "build")
    os.makedirs(bdir, exist_ok=True)
    cpath = os.path.join(bdir, f"cport_{key}.c")
    so = os.path.join(bdir, f"cport_{key}.so")
    if not os.path.exists(so):
        with open(cpath, "w") as f:
            f.write(src)
        tmp = so + f".tmp{os.getpid()}"
        subprocess.run(["gcc", "-O3", "-march=native", "-ffp-contract=off", "-fno-fast-math", "-std=gnu11", "-fPIC",
                        "-shared", "-I", HERE, "-o", tmp, cpath, "-lpthread", "-lm"], check=True, capture_output=True)
        os.replace(tmp, so)
    lib = ct.CDLL(so)
    P = ct.POINTER
    d, u32, i64, u64 = ct.c_double, ct.c_uint32, ct.c_int64, ct.c_uint64
    f = lib.ora_f64_propagate_until
    f.restype = ct.c_int
    f.argtypes = [P(u32), P(d), u32, ct.c_int, u32, u32, P(d), P(d), P(d), P(d), P(d), P(d), P(d), P(d), u64, P(i64),
                  P(d), P(d), P(u64), ct.c_int]
    _cache[key] = (lib, L)
    return lib, L


def propagate_until(dec, order, state, tf, pars=None, t_hi=None, nthreads=1, L=8):
    """Same contract as pyoracle.propagate_until (double), with the compiled recurrences."""
    import pyoracle as O

    lib, L = build(dec, order, L)
    words = np.ascontiguousarray(dec.to_words())
    consts = dec.consts_array()
    state = np.ascontiguousarray(np.asarray(state, dtype=np.float64)).copy()
    n, B = state.shape
    pars = np.ascontiguousarray(np.asarray(pars if pars is not None else np.zeros((max(dec.n_par, 1), B)),
                                           dtype=np.float64))
    th = np.ascontiguousarray(np.broadcast_to(np.asarray(0.0 if t_hi is None else t_hi, dtype=float), (B,))).copy()
    tl = np.zeros(B)
    tfh = np.ascontiguousarray(np.broadcast_to(np.asarray(tf, dtype=float), (B,))).copy()
    tfl = np.zeros(B)
    mdt = np.full(B, np.inf)
    oc = np.zeros(B, dtype=np.int64)
    mn, mx = np.zeros(B), np.zeros(B)
    ns = np.zeros(B, dtype=np.uint64)
    p = O._p
    rc = lib.ora_f64_propagate_until(p(words, ct.c_uint32), p(consts, ct.c_double), order, 0, B, L,
                                     p(state, ct.c_double), None, p(pars, ct.c_double), p(th, ct.c_double),
                                     p(tl, ct.c_double), p(tfh, ct.c_double), p(tfl, ct.c_double),
                                     p(mdt, ct.c_double), 0, p(oc, ct.c_int64), p(mn, ct.c_double), p(mx, ct.c_double),
                                     p(ns, ct.c_uint64), int(nthreads))
    if rc != 0:
        raise RuntimeError(f"cport propagate_until failed with code {rc}")
    return dict(state=state, t_hi=th, outcome=oc, min_h=mn, max_h=mx, nsteps=ns)
