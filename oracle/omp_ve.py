"""ORACLE — test infrastructure only.  Nothing in the product imports this.

A CPU restatement of the reference OpenMP vector engine's semantics, used as
the parity checker for the hip engine and as the CPU baseline in bench.py.

It replays a bridge trace (bohrium_amd._bh_api.trace_start/stop) instruction
by instruction.  Each instruction becomes a C99 loop nest whose scalar body is
the reference's per-opcode C expression (core/jitk/instruction.cpp:58-578,
special forms of core/jitk/engines/engine.cpp:251-336), compiled like the
reference compiles its kernels (ve/openmp/CMakeLists.txt:58-93:
`-x c -fPIC -shared -std=gnu99 -O3 -march=native -fopenmp`, with <tgmath.h>
as in ve/openmp/engine_openmp.cpp:331-346) and loaded with ctypes.

Semantics pinned here (and checked against NumPy / closed forms in tests):
  * views: element address = start + sum(i_d * stride_d)            (view.cpp:31-88)
  * reductions start from sweep_identity (instruction.cpp:546-578), including
    its quirks: max-reduce of floats starts at the smallest positive normal,
    bool min/max start at true, bitwise-and starts at the 32-bit all-ones value;
    then accumulate sequentially along the axis
  * accumulate: out[k] = out[k-1] OP in[k] with out[0] = identity OP in[0] (engine.cpp:299-308)
  * random: out = random123(start, key, memory offset of the output element) (engine.cpp:263-277)
  * range: out = memory offset of the output element                        (engine.cpp:254-262)
  * gather: out[i] = in.base[in.start + idx[i]]; scatter: out.base[out.start + idx[i]] = in[i]
Parity is asserted on results, not on fusion partitions (this oracle does not fuse).
"""
import ctypes
import hashlib
import math
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")

# Reference compile flags of the OpenMP engine (ve/openmp/CMakeLists.txt:58-93); -w instead of -Werror
# because the generated checker code may carry unused variables.
CFLAGS = ["-x", "c", "-fPIC", "-shared", "-std=gnu99", "-O3", "-march=native", "-fopenmp", "-w"]

# opcode ids (reference core/codegen/opcodes.json)
OP = dict(ADD=1, SUBTRACT=2, MULTIPLY=3, DIVIDE=4, POWER=5, ABSOLUTE=6, GREATER=7, GREATER_EQUAL=8, LESS=9,
          LESS_EQUAL=10, EQUAL=11, NOT_EQUAL=12, LOGICAL_AND=13, LOGICAL_OR=14, LOGICAL_XOR=15, LOGICAL_NOT=16,
          MAXIMUM=17, MINIMUM=18, BITWISE_AND=19, BITWISE_OR=20, BITWISE_XOR=21, INVERT=22, LEFT_SHIFT=23,
          RIGHT_SHIFT=24, COS=25, SIN=26, TAN=27, COSH=28, SINH=29, TANH=30, ARCSIN=31, ARCCOS=32, ARCTAN=33,
          ARCSINH=34, ARCCOSH=35, ARCTANH=36, ARCTAN2=37, EXP=38, EXP2=39, EXPM1=40, LOG=41, LOG2=42, LOG10=43,
          LOG1P=44, SQRT=45, CEIL=46, TRUNC=47, FLOOR=48, RINT=49, MOD=50, ISNAN=51, ISINF=52, IDENTITY=53,
          FREE=55, NONE=57, TALLY=58, ADD_REDUCE=59, MULTIPLY_REDUCE=60, MINIMUM_REDUCE=61, MAXIMUM_REDUCE=62,
          LOGICAL_AND_REDUCE=63, BITWISE_AND_REDUCE=64, LOGICAL_OR_REDUCE=65, BITWISE_OR_REDUCE=66,
          LOGICAL_XOR_REDUCE=67, BITWISE_XOR_REDUCE=68, RANDOM=71, RANGE=72, REAL=73, IMAG=74,
          ADD_ACCUMULATE=75, MULTIPLY_ACCUMULATE=76, SIGN=77, GATHER=79, SCATTER=80, REMAINDER=81,
          COND_SCATTER=82, ISFINITE=83, CONJ=84)
NAME = {v: k for k, v in OP.items()}
REDUCE = {59, 60, 61, 62, 63, 64, 65, 66, 67, 68}
ACCUMULATE = {75, 76}

CTYPE = {"|b1": "bool", "|i1": "int8_t", "<i2": "int16_t", "<i4": "int32_t", "<i8": "int64_t", "|u1": "uint8_t",
         "<u2": "uint16_t", "<u4": "uint32_t", "<u8": "uint64_t", "<f4": "float", "<f8": "double",
         "<c8": "float complex", "<c16": "double complex"}


def _kind(dt):
    return np.dtype(dt).kind


def c_literal(dt, value):
    """An exact C99 literal of numpy dtype `dt`."""
    dt = np.dtype(dt)
    v = np.asarray(value).astype(dt)[()]
    if dt.kind == "b":
        return "((bool)%d)" % int(bool(v))
    if dt.kind in "iu":
        iv = int(v)
        if dt == np.int64 and iv == -2 ** 63:
            return "((int64_t)(-9223372036854775807LL - 1))"
        suffix = "ULL" if dt.kind == "u" else "LL"
        return "((%s)%d%s)" % (CTYPE[dt.str], iv, suffix)
    if dt.kind == "f":
        return _c_float(float(v), dt == np.float32)
    if dt.kind == "c":
        single = dt == np.complex64
        return "((%s)(%s + %s * I))" % (CTYPE[dt.str], _c_float(float(v.real), single), _c_float(float(v.imag), single))
    raise ValueError(dt)


def _c_float(x, single):
    if math.isnan(x):
        return "NAN"
    if math.isinf(x):
        return "(-INFINITY)" if x < 0 else "INFINITY"
    return "(%s%s)" % (float.hex(x), "f" if single else "")


def sweep_identity(opcode, dt):
    """Identity of a sweep (reference core/jitk/instruction.cpp:546-578 and core/bh_constant.cpp:32-104)."""
    dt = np.dtype(dt)
    if opcode in (OP["ADD_REDUCE"], OP["BITWISE_OR_REDUCE"], OP["BITWISE_XOR_REDUCE"], OP["LOGICAL_OR_REDUCE"],
                  OP["LOGICAL_XOR_REDUCE"], OP["ADD_ACCUMULATE"]):
        return np.zeros((), dt)[()]
    if opcode in (OP["MULTIPLY_REDUCE"], OP["MULTIPLY_ACCUMULATE"]):
        return np.ones((), dt)[()]
    if opcode in (OP["BITWISE_AND_REDUCE"], OP["LOGICAL_AND_REDUCE"]):
        # bh_constant(~0u, dtype): the unsigned 32-bit all-ones value converted with a C cast.
        return np.array(0xFFFFFFFF, np.uint64).astype(dt)[()] if dt.kind != "b" else np.bool_(True)
    if opcode == OP["MAXIMUM_REDUCE"]:
        if dt.kind == "b":
            return np.bool_(True)
        if dt.kind == "f":
            return np.finfo(dt).tiny  # numeric_limits<T>::min(): smallest positive normal
        if dt.kind == "i":
            return np.array(np.iinfo(dt).min + 1, dt)[()]
        return np.zeros((), dt)[()]
    if opcode == OP["MINIMUM_REDUCE"]:
        if dt.kind == "b":
            return np.bool_(True)
        if dt.kind == "f":
            return np.finfo(dt).max
        return np.array(np.iinfo(dt).max, dt)[()]
    raise ValueError("no identity for opcode %d" % opcode)


def c_operation(opcode, types, ops):
    """C99 statement for one element (reference write_operation, non-OpenCL branch).
    `types` are the numpy dtypes of the operands (constant included), `ops` their C spellings."""
    a = ops
    t0 = np.dtype(types[0])
    t1 = np.dtype(types[1]) if len(types) > 1 else None

    def un(fn):
        return "%s = %s(%s);" % (a[0], fn, a[1])

    def bin_(op):
        return "%s = %s %s %s;" % (a[0], a[1], op, a[2])

    n = NAME[opcode]
    simple = {"BITWISE_AND": "&", "BITWISE_OR": "|", "BITWISE_XOR": "^", "LOGICAL_OR": "||", "LOGICAL_AND": "&&",
              "LEFT_SHIFT": "<<", "RIGHT_SHIFT": ">>", "GREATER": ">", "GREATER_EQUAL": ">=", "LESS": "<",
              "LESS_EQUAL": "<=", "NOT_EQUAL": "!=", "EQUAL": "==", "ADD": "+", "SUBTRACT": "-", "MULTIPLY": "*",
              "ADD_ACCUMULATE": "+", "MULTIPLY_ACCUMULATE": "*"}
    if n in simple:
        return bin_(simple[n])
    if n == "BITWISE_AND_REDUCE":
        return "%s = %s & %s;" % (a[0], a[0], a[1])
    if n == "BITWISE_OR_REDUCE":
        return "%s = %s | %s;" % (a[0], a[0], a[1])
    if n == "BITWISE_XOR_REDUCE":
        return "%s = %s ^ %s;" % (a[0], a[0], a[1])
    if n == "LOGICAL_NOT":
        return "%s = !%s;" % (a[0], a[1])
    if n == "LOGICAL_OR_REDUCE":
        return "%s = %s || %s;" % (a[0], a[0], a[1])
    if n == "LOGICAL_AND_REDUCE":
        return "%s = %s && %s;" % (a[0], a[0], a[1])
    if n == "LOGICAL_XOR":
        return "%s = !%s != !%s;" % (a[0], a[1], a[2])
    if n == "LOGICAL_XOR_REDUCE":
        return "%s = !%s != !%s;" % (a[0], a[0], a[1])
    if n == "MAXIMUM":
        return "%s = %s > %s ? %s : %s;" % (a[0], a[1], a[2], a[1], a[2])
    if n == "MAXIMUM_REDUCE":
        return "%s = %s > %s ? %s : %s;" % (a[0], a[0], a[1], a[0], a[1])
    if n == "MINIMUM":
        return "%s = %s < %s ? %s : %s;" % (a[0], a[1], a[2], a[1], a[2])
    if n == "MINIMUM_REDUCE":
        return "%s = %s < %s ? %s : %s;" % (a[0], a[0], a[1], a[0], a[1])
    if n == "INVERT":
        return ("%s = !%s;" if t0.kind == "b" else "%s = ~%s;") % (a[0], a[1])
    if n == "MOD":
        if t0.kind == "f":
            return "%s = fmod(%s, %s);" % (a[0], a[1], a[2])
        return bin_("%")
    if n == "REMAINDER":
        if t0.kind == "f":
            return "%s = %s - floor(%s / %s) * %s;" % (a[0], a[1], a[1], a[2], a[2])
        if t0.kind == "u":
            return bin_("%")
        return ("%s = ((%s > 0) == (%s > 0) || (%s %% %s) == 0) ? (%s %% %s) : (%s %% %s) + %s;"
                % (a[0], a[1], a[2], a[1], a[2], a[1], a[2], a[1], a[2], a[2]))
    fns = {"RINT": "rint", "EXP2": "exp2", "EXPM1": "expm1", "LOG1P": "log1p", "ARCSIN": "asin", "ARCCOS": "acos",
           "ARCTAN": "atan", "ARCSINH": "asinh", "ARCCOSH": "acosh", "ARCTANH": "atanh", "FLOOR": "floor",
           "CEIL": "ceil", "TRUNC": "trunc", "LOG2": "log2", "SIN": "sin", "COS": "cos", "TAN": "tan",
           "SINH": "sinh", "COSH": "cosh", "TANH": "tanh", "EXP": "exp", "SQRT": "sqrt", "LOG": "log",
           "CONJ": "conj", "REAL": "creal", "IMAG": "cimag"}
    if n in fns:
        return un(fns[n])
    if n == "ARCTAN2":
        return "%s = atan2(%s, %s);" % (a[0], a[1], a[2])
    if n in ("ISNAN", "ISINF", "ISFINITE"):
        fn = n.lower()
        if t1.kind == "c":
            return "%s = %s(creal(%s));" % (a[0], fn, a[1])
        if t1.kind == "f":
            return "%s = %s(%s);" % (a[0], fn, a[1])
        return "%s = %s;" % (a[0], "true" if n == "ISFINITE" else "false")
    if n in ("RANGE", "RANDOM", "GATHER", "SCATTER"):
        return "%s = %s;" % (a[0], a[1])
    if n == "COND_SCATTER":
        return "if (%s) { %s = %s; }" % (a[2], a[0], a[1])
    if n == "ABSOLUTE":
        if t1.kind == "b" or t1.kind == "u":
            return "%s = %s;" % (a[0], a[1])
        if t1.kind == "f":
            return un("fabs")
        if t1 == np.int64:
            return un("llabs")
        return "%s = abs((int)%s);" % (a[0], a[1])
    if n == "POWER":
        if t0.kind in "iub":
            return "%s = pow((double) %s, (double)%s);" % (a[0], a[1], a[2])
        return "%s = pow(%s, %s);" % (a[0], a[1], a[2])
    if n == "ADD_REDUCE":
        return "%s += %s;" % (a[0], a[1])
    if n == "MULTIPLY_REDUCE":
        return "%s *= %s;" % (a[0], a[1])
    if n == "DIVIDE":
        if t0.kind == "i":
            return ("%s = ((%s > 0) != (%s > 0) && (%s %% %s) != 0)?(%s / %s - 1):(%s / %s);"
                    % (a[0], a[1], a[2], a[1], a[2], a[1], a[2], a[1], a[2]))
        return bin_("/")
    if n == "IDENTITY":
        return "%s = %s;" % (a[0], a[1])
    if n == "LOG10":
        if t0.kind == "c":
            return "%s = clog(%s) / log(10.0f);" % (a[0], a[1])
        return un("log10")
    if n == "SIGN":
        if t0.kind == "c":
            ct = "float" if t0 == np.complex64 else "double"
            return ("{ %s real = creal(%s); %s imag = cimag(%s); %s = (real == 0 ? ((imag > 0) - (0 > imag)) : "
                    "((real > 0) - (0 > real))); }" % (ct, a[1], ct, a[1], a[0]))
        return "%s = ((%s > 0) - (0 > %s));" % (a[0], a[1], a[1])
    raise NotImplementedError("oracle: opcode %s" % n)


def _index(view, idx_names):
    """start + sum(i_d * stride_d) with literal numbers (a sliding view's start is st[j])."""
    start, shape, stride = view[1], view[2], view[3]
    terms = [str(start) if len(view) < 5 else "st[%d]" % view[4]]
    for name, st in zip(idx_names, stride):
        if st != 0:
            terms.append("%s*(%d)" % (name, st))
    return "(" + " + ".join(terms) + ")"


class _Gen:
    def __init__(self, dtypes):
        self.dtypes = dtypes  # bid -> dtype str
        self.lines = []
        self.sliding = []     # [start, slide dims] of every sliding operand, indexed by j in st[j]

    def with_slides(self, ops, slides):
        """Tag sliding operands with their slot j (their start is read from st[j] at run time)."""
        if not slides:
            return ops
        out = []
        for o, dims in zip(ops, slides):
            if o is not None and dims:
                out.append(tuple(o) + (len(self.sliding),))
                self.sliding.append([o[1], dims, 0])
            else:
                out.append(o)
        return out

    def arr(self, bid):
        if getattr(self, "referenced", None) is not None:
            self.referenced.add(bid)
        return "((%s*)d[%d])" % (CTYPE[self.dtypes[bid]], bid)

    def ref(self, view, idx_names):
        return "%s[%s]" % (self.arr(view[0]), _index(view, idx_names))

    def loops(self, shape, body, names=None, indent="    "):
        names = names or ["i%d" % k for k in range(len(shape))]
        out = []
        for k, (n, s) in enumerate(zip(names, shape)):
            out.append("%sfor (int64_t %s = 0; %s < %d; ++%s) {" % (indent + "  " * k, n, n, s, n))
        out.extend(indent + "  " * len(shape) + line for line in body)
        for k in reversed(range(len(shape))):
            out.append(indent + "  " * k + "}")
        return out

    def instr(self, opcode, ops, const):
        L = []
        if opcode in (OP["FREE"], OP["NONE"], OP["TALLY"]):
            return L
        out = ops[0]
        types = []
        for o in ops:
            if o is None:
                types.append(np.dtype(const[0]) if const[0] != "r123" else np.dtype(np.uint64))
            else:
                types.append(np.dtype(self.dtypes[o[0]]))
        if opcode in REDUCE or opcode in ACCUMULATE:
            inp, axis = ops[1], int(const[1])
            shape = inp[2]
            t = np.dtype(self.dtypes[out[0]])
            ident = c_literal(t, sweep_identity(opcode, t))
            names = ["i%d" % k for k in range(len(shape))]
            outer = [n for k, n in enumerate(names) if k != axis]
            outer_shape = [s for k, s in enumerate(shape) if k != axis]
            in_ref = self.ref(inp, names)
            body = ["%s acc = %s;" % (CTYPE[t.str], ident)]
            inner = []
            if opcode in REDUCE:
                inner.append(c_operation(opcode, [t, t], ["acc", in_ref]))
                out_names = outer if len(shape) > 1 else []
                out_ref = self.ref(out, out_names) if out_names else "%s[%s]" % (self.arr(out[0]), _index(out, []))
                body += self.loops([shape[axis]], inner, [names[axis]], indent="")
                body.append("%s = acc;" % out_ref)
            else:
                inner.append(c_operation(opcode, [t, t, t], ["acc", "acc", in_ref]))
                inner.append("%s = acc;" % self.ref(out, names))
                body += self.loops([shape[axis]], inner, [names[axis]], indent="")
            L += ["  {"] + self.loops(outer_shape, body, outer) + ["  }"]
            return L
        shape = out[2]
        if opcode in (OP["GATHER"], OP["SCATTER"], OP["COND_SCATTER"]):
            shape = ops[2][2]
        names = ["i%d" % k for k in range(len(shape))]
        if opcode == OP["RANDOM"]:
            start, key = const[1]
            body = ["%s = oracle_random123(%dULL, %dULL, (uint64_t)%s);" % (
                self.ref(out, names), start, key, _index(out, names))]
        elif opcode == OP["RANGE"]:
            body = ["%s = %s;" % (self.ref(out, names), _index(out, names))]
        elif opcode == OP["GATHER"]:
            inp, idx = ops[1], ops[2]
            nb = self.nelem[inp[0]]
            body = ["{ int64_t g = %s + (int64_t)%s;" % (_index(inp, []), self.ref(idx, names)),
                    "  if (g < 0 || g >= %d) err = 1; else %s = %s[g]; }" % (nb, self.ref(out, names),
                                                                          self.arr(inp[0]))]
        elif opcode in (OP["SCATTER"], OP["COND_SCATTER"]):
            val, idx = ops[1], ops[2]
            nb = self.nelem[out[0]]
            cond = "" if opcode == OP["SCATTER"] else "if (%s) " % self.ref(ops[3], names)
            body = ["%s{ int64_t g = %s + (int64_t)%s;" % (cond, _index(out, []), self.ref(idx, names)),
                    "  if (g < 0 || g >= %d) err = 1; else %s[g] = %s; }" % (nb, self.arr(out[0]),
                                                                          self.ref(val, names))]
        else:
            cops = []
            for o in ops:
                if o is None:
                    cops.append(c_literal(const[0], const[1]))
                else:
                    cops.append(self.ref(o, names))
            body = [c_operation(opcode, types, cops)]
        return ["  {"] + self.loops(shape, body) + ["  }"]


def _compile(source):
    h = hashlib.sha1(source.encode()).hexdigest()[:20]
    os.makedirs(BUILD, exist_ok=True)
    so = os.path.join(BUILD, "oracle_%s.so" % h)
    if not os.path.exists(so):
        with tempfile.NamedTemporaryFile("w", suffix=".c", dir=BUILD, delete=False) as f:
            f.write(source)
            src = f.name
        tmp = so + ".tmp%d" % os.getpid()
        r = subprocess.run(["gcc"] + CFLAGS + ["-I" + CSRC, src, "-o", tmp, "-lm"],
                           stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        os.unlink(src)
        if r.returncode != 0:
            raise RuntimeError("oracle: gcc failed:\n" + r.stdout[-4000:])
        os.replace(tmp, so)
    return ctypes.CDLL(so)


def build_checker():
    """Compile the checker's fixed C piece (the philox restatement) once, as a build check."""
    src = HEADER + "\nuint64_t oracle_philox(uint64_t s, uint64_t k, uint64_t i) { return oracle_random123(s, k, i); }\n"
    lib = _compile(src)
    lib.oracle_philox.restype = ctypes.c_uint64
    lib.oracle_philox.argtypes = [ctypes.c_uint64] * 3
    return lib


HEADER = """#include <stdint.h>
#include <stdlib.h>
#include <stdbool.h>
#include <complex.h>
#include <tgmath.h>
#include <math.h>
#include "r123_philox.h"
"""


def _split_groups(trace):
    """Split a trace into flush groups: (records, nrepeats, cond_bid)."""
    groups, cur = [], []
    for rec in trace:
        if rec[0] == "flush":
            groups.append((cur, rec[1], rec[2]))
            cur = []
        else:
            cur.append(rec)
    if cur:
        groups.append((cur, 1, None))
    return groups


def _slide(entry, st, j):
    """Move one sliding operand to its next position (reference
    include/bohrium/jitk/engines/dyn_view.hpp:21-67, without shape changes or resets)."""
    dims, counter = entry[1], entry[2]
    start = st[j]
    for (rank, offset_change, shape_change, stride, shape, step_delay) in dims:
        if step_delay == 1 or counter % step_delay == step_delay - 1:
            if stride:
                change = offset_change * stride
                max_rel = stride * shape
                rel = start % (stride * shape) + change
                if rel < 0:
                    change += max_rel
                elif rel >= max_rel:
                    change -= max_rel
                start += change
    st[j] = start
    entry[2] = counter + 1


def replay(trace, keep=None):
    """Execute `trace` on the CPU.  Returns {base id: numpy array} for every base
    still alive at the end (or only those in `keep`)."""
    dtypes, nelem = {}, {}
    for rec in trace:
        if rec[0] == "new":
            dtypes[rec[1]] = np.dtype(rec[2]).str
            nelem[rec[1]] = rec[3]
    gen = _Gen(dtypes)
    gen.nelem = nelem
    groups = _split_groups(trace)
    src = [HEADER]
    slots = []
    for gi, (recs, _, _) in enumerate(groups):
        src.append("int g%d(void **d, int64_t *st) {\n  int err = 0;" % gi)
        first = len(gen.sliding)
        for rec in recs:
            if rec[0] == "op":
                ops = gen.with_slides(rec[2], rec[4] if len(rec) > 4 else None)
                src.extend(gen.instr(rec[1], ops, rec[3]))
        slots.append((first, len(gen.sliding)))
        src.append("  return err;\n}\n")
    lib = _compile("\n".join(src))
    st = (ctypes.c_int64 * max(1, len(gen.sliding)))()
    for j, (start, _, _) in enumerate(gen.sliding):
        st[j] = start
    data = {}
    freed = set()
    maxbid = max(dtypes) + 1 if dtypes else 1
    for gi, (recs, nrep, cond) in enumerate(groups):
        for rec in recs:
            if rec[0] == "new":
                data[rec[1]] = np.zeros(max(1, rec[3]), dtype=np.dtype(rec[2]))
            elif rec[0] == "data":
                data[rec[1]] = np.array(rec[2], dtype=np.dtype(dtypes[rec[1]]), copy=True).reshape(-1)
            elif rec[0] == "free":
                freed.add(rec[1])
            elif rec[0] == "ext":
                raise NotImplementedError("oracle: extension method %s" % rec[1])
        ptrs = (ctypes.c_void_p * maxbid)()
        for b, arr in data.items():
            ptrs[b] = arr.ctypes.data
        fn = getattr(lib, "g%d" % gi)
        fn.argtypes = [ctypes.c_void_p * maxbid, ctypes.c_int64 * max(1, len(gen.sliding))]
        for _ in range(int(nrep)):
            if fn(ptrs, st) != 0:
                raise IndexError("oracle: gather/scatter index out of bounds")
            if cond is not None and not bool(data[cond][0]):
                break
            for j in range(*slots[gi]):
                _slide(gen.sliding[j], st, j)
        for b in list(freed):
            data.pop(b, None)
        freed.clear()
    if keep is not None:
        return {b: data[b] for b in keep if b in data}
    return data


# ---------------------------------------------------------------------------------------------
# Fused OpenMP form (CPU baseline).  Runs of consecutive instructions over one iteration shape (any
# rank, any strides) become one loop nest, written the way the reference OpenMP engine writes a
# fused block (ve/openmp/engine_openmp.cpp:274-329 loopHeadWriter / writeHeader, openmp_util.hpp):
#   * `#pragma omp parallel for` on the outermost loop.  A sweep of that loop goes into a
#     `reduction(op:acc)` clause when its operator has one and its output is a single element
#     (scalar-replaced); otherwise each update is guarded by `#pragma omp atomic` (add, multiply,
#     bitwise) or `#pragma omp critical` (everything else), as openmp_reduce_compatible /
#     openmp_atomic_compatible decide.
#   * `#pragma omp simd` on the innermost loop when every sweep of that loop has a reduction
#     operator and nothing in the block is atomic or critical (simd_compatible); sweeps of the
#     innermost loop are scalar accumulators named in its simd reduction clause (for rank > 0).
#     A one-dimensional block is both: `parallel for simd`.
#   * Temporaries created and freed inside the run are scalars (array contraction); reductions over
#     the first or the last axis of the block.
# Everything else (gather, scatter, accumulate, reductions over middle axes, sliding views) runs
# per instruction as in `replay`.

_OMP_RED = {OP["ADD_REDUCE"]: "+", OP["MULTIPLY_REDUCE"]: "*", OP["MINIMUM_REDUCE"]: "min",
            OP["MAXIMUM_REDUCE"]: "max", OP["BITWISE_AND_REDUCE"]: "&", OP["BITWISE_OR_REDUCE"]: "|",
            OP["BITWISE_XOR_REDUCE"]: "^"}
_OMP_ATOMIC = {OP["ADD_REDUCE"]: "+=", OP["MULTIPLY_REDUCE"]: "*=", OP["BITWISE_AND_REDUCE"]: "&=",
               OP["BITWISE_OR_REDUCE"]: "|=", OP["BITWISE_XOR_REDUCE"]: "^="}
_NO_FUSE = {OP["GATHER"], OP["SCATTER"], OP["COND_SCATTER"], OP["ADD_ACCUMULATE"], OP["MULTIPLY_ACCUMULATE"],
            OP["FREE"], OP["NONE"], OP["TALLY"]}


def _iter_shape(rec):
    ops = rec[2]
    v = ops[1] if rec[1] in REDUCE else ops[0]
    return tuple(v[2])


def _fusible(rec):
    opcode, ops, const = rec[1], rec[2], rec[3]
    if opcode in _NO_FUSE or opcode not in NAME:
        return False
    if len(rec) > 4 and rec[4] and any(rec[4]):
        return False
    S = _iter_shape(rec)
    if len(S) == 0 or 0 in S:
        return False
    if opcode in REDUCE:
        axis = int(const[1])
        if axis not in (0, len(S) - 1):
            return False
        out_n = int(np.prod([s for k, s in enumerate(S) if k != axis])) if len(S) > 1 else 1
        return int(np.prod(ops[0][2])) == out_n and (len(S) == 1 or tuple(ops[0][2]) == tuple(
            s for k, s in enumerate(S) if k != axis))
    return all(o is None or tuple(o[2]) == S for o in ops)


def _view_key(o):
    return (o[0], o[1], tuple(o[2]), tuple(o[3]))


class FusedProgram:
    """Compile a trace in the fused OpenMP form; run its flush groups one at a time."""

    def __init__(self, trace):
        self.dtypes, self.nelem = {}, {}
        for rec in trace:
            if rec[0] == "new":
                self.dtypes[rec[1]] = np.dtype(rec[2]).str
                self.nelem[rec[1]] = rec[3]
        self.groups = _split_groups(trace)
        gen = _Gen(self.dtypes)
        gen.nelem = self.nelem
        src = [HEADER]
        self.runs = []  # (number of instructions, iteration shape) of every fused run, for tests
        self.arrays = []  # per group: the bases its code addresses as arrays (temporaries are scalars)
        for gi, (recs, _, _) in enumerate(self.groups):
            gen.referenced = set()
            src.append("int g%d(void **d) {\n  int err = 0;" % gi)
            src.extend(self._group_source(gen, recs))
            src.append("  return err;\n}\n")
            self.arrays.append(gen.referenced)
        gen.referenced = None
        self.source = "\n".join(src)
        self.lib = _compile(self.source)
        self.data = {}
        self._gi = 0

    def _group_source(self, gen, recs):
        ops_recs = [r for r in recs if r[0] == "op"]
        frees = set(r[1] for r in recs if r[0] == "free")
        uses = {}
        for k, r in enumerate(ops_recs):
            for o in r[2]:
                if o is not None:
                    uses.setdefault(o[0], set()).add(k)
        out, k = [], 0
        while k < len(ops_recs):
            r = ops_recs[k]
            if len(r) > 4 and r[4] and any(r[4]):
                raise NotImplementedError("oracle fused form: sliding views (use replay)")
            if not _fusible(r):
                out.extend(gen.instr(r[1], r[2], r[3]))
                k += 1
                continue
            S = _iter_shape(r)
            written, read, red_outs, read_first = {}, {}, set(), set()
            run, j = [], k
            while j < len(ops_recs):
                q = ops_recs[j]
                if not _fusible(q) or _iter_shape(q) != S:
                    break
                ins = [o for o in q[2][1:] if o is not None]
                outv = q[2][0]
                clash = any(o[0] in red_outs or (o[0] in written and written[o[0]] != _view_key(o)) for o in ins)
                clash |= outv[0] in red_outs
                clash |= outv[0] in written and written[outv[0]] != _view_key(outv)
                clash |= any(v != _view_key(outv) for v in read.get(outv[0], ()))
                if clash:
                    break
                for o in ins:
                    read.setdefault(o[0], set()).add(_view_key(o))
                    if o[0] not in written:
                        read_first.add(o[0])
                if q[1] in REDUCE:
                    red_outs.add(outv[0])
                else:
                    written[outv[0]] = _view_key(outv)
                run.append(q)
                j += 1
            idx = set(range(k, j))
            temps = set(b for b in written
                        if b in frees and uses[b] <= idx and b not in red_outs and b not in read_first)
            self.runs.append((len(run), S))
            out.extend(self._run_source(gen, run, S, temps))
            k = j
        return out

    def _run_source(self, gen, run, S, temps):
        n = len(S)
        names = ["i%d" % a for a in range(n)]
        inner = names[-1]

        def ref(o):
            if o[0] in temps:
                return "t%d" % o[0]
            return "%s[%s]" % (gen.arr(o[0]), _index(o, names))

        # sweeps: axis 0 (outermost loop) or the last axis (innermost loop)
        outer_sweeps, inner_sweeps = [], []
        for r in run:
            if r[1] in REDUCE:
                axis = int(r[3][1])
                (outer_sweeps if axis == 0 else inner_sweeps).append(r)
        clause, atomic, critical = [], set(), set()
        for r in outer_sweeps:
            scalar = n == 1
            if r[1] in _OMP_RED and scalar:
                clause.append(r)
        clause_ids = set(id(r) for r in clause)
        for r in outer_sweeps:
            if id(r) in clause_ids:
                continue
            if r[1] in _OMP_ATOMIC:
                atomic.add(id(r))
            else:
                critical.add(id(r))
        simd = all(r[1] in _OMP_RED for r in inner_sweeps) and not atomic and not critical
        if n == 1:
            simd = simd and all(r[1] in _OMP_RED for r in outer_sweeps)

        def out_ref(r, idx_names):
            o = r[2][0]
            return "%s[%s]" % (gen.arr(o[0]), _index(o, idx_names))

        L = ["  {"]
        # a sweep of the outermost loop starts from its identity (the reference's identity block)
        for r in outer_sweeps:
            t = np.dtype(self.dtypes[r[2][0][0]])
            ident = c_literal(t, sweep_identity(r[1], t))
            if id(r) in clause_ids:
                L.append("    %s acc%d = %s;" % (CTYPE[t.str], r[2][0][0], ident))
            else:
                L.extend(gen.loops(list(S[1:]), ["%s = %s;" % (out_ref(r, names[1:]), ident)], names[1:]))
        pragma = "parallel for" + (" simd" if n == 1 and simd else "")
        pragma += "".join(" reduction(%s:acc%d)" % (_OMP_RED[r[1]], r[2][0][0]) for r in clause)
        body = []
        for b in sorted(temps):
            body.append("%s t%d;" % (CTYPE[self.dtypes[b]], b))
        for r in run:
            opcode, ops, const = r[1], r[2], r[3]
            if opcode in REDUCE:
                t = np.dtype(self.dtypes[ops[0][0]])
                axis = int(const[1])
                if axis == n - 1 and n > 1:
                    body.append(c_operation(opcode, [t, t], ["r%d" % ops[0][0], ref(ops[1])]))
                elif id(r) in clause_ids:
                    body.append(c_operation(opcode, [t, t], ["acc%d" % ops[0][0], ref(ops[1])]))
                elif id(r) in atomic:
                    body.append("{ %s v_ = %s;" % (CTYPE[t.str], ref(ops[1])))
                    body.append("#pragma omp atomic")
                    body.append("  %s %s v_; }" % (out_ref(r, names[1:] if n > 1 else []), _OMP_ATOMIC[opcode]))
                else:
                    body.append("{ %s v_ = %s;" % (CTYPE[t.str], ref(ops[1])))
                    body.append("#pragma omp critical")
                    body.append("  { " + c_operation(opcode, [t, t], [out_ref(r, names[1:] if n > 1 else []), "v_"])
                                + " } }")
            elif opcode == OP["RANDOM"]:
                s0, key = const[1]
                body.append("%s = oracle_random123(%dULL, %dULL, (uint64_t)%s);" % (ref(ops[0]), s0, key,
                                                                                   _index(ops[0], names)))
            elif opcode == OP["RANGE"]:
                body.append("%s = %s;" % (ref(ops[0]), _index(ops[0], names)))
            else:
                types, cops = [], []
                for o in ops:
                    if o is None:
                        types.append(np.dtype(const[0]))
                        cops.append(c_literal(const[0], const[1]))
                    else:
                        types.append(np.dtype(self.dtypes[o[0]]))
                        cops.append(ref(o))
                body.append(c_operation(opcode, types, cops))
        # the loop nest: parallel outermost loop, sequential middle loops, simd innermost loop with
        # the scalar accumulators of its sweeps
        ind = "    "
        L.append("    #pragma omp " + pragma)
        lines = []
        for a in range(n):
            if a == n - 1 and n > 1:
                for r in inner_sweeps:
                    t = np.dtype(self.dtypes[r[2][0][0]])
                    lines.append(ind + "  " * a + "%s r%d = %s;" % (CTYPE[t.str], r[2][0][0],
                                                                   c_literal(t, sweep_identity(r[1], t))))
                if simd:
                    red = "".join(" reduction(%s:r%d)" % (_OMP_RED[r[1]], r[2][0][0]) for r in inner_sweeps)
                    lines.append(ind + "  " * a + "#pragma omp simd" + red)
            lines.append(ind + "  " * a + "for (int64_t %s = 0; %s < %d; ++%s) {" % (names[a], names[a], S[a],
                                                                                  names[a]))
        lines.extend(ind + "  " * n + b for b in body)
        for a in reversed(range(n)):
            lines.append(ind + "  " * a + "}")
            if a == n - 1 and n > 1:
                for r in inner_sweeps:
                    lines.append(ind + "  " * a + "%s = r%d;" % (out_ref(r, names[:-1]), r[2][0][0]))
        L.extend(lines)
        for r in clause:
            o = r[2][0]
            L.append("    %s[%s] = acc%d;" % (gen.arr(o[0]), _index(o, []), o[0]))
        L.append("  }")
        return L

    def run_next_group(self):
        """Execute the next flush group; returns its wall time in seconds (kernel execution only)."""
        import time
        recs, nrep, cond = self.groups[self._gi]
        for rec in recs:
            if rec[0] == "new" and rec[1] in self.arrays[self._gi]:
                # zero-filled here, outside the timed region (first-touch page faults included)
                a = np.empty(max(1, rec[3]), dtype=np.dtype(rec[2]))
                a.fill(0)
                self.data[rec[1]] = a
            elif rec[0] == "data":
                self.data[rec[1]] = np.array(rec[2], dtype=np.dtype(self.dtypes[rec[1]])).reshape(-1)
        maxbid = max(self.dtypes) + 1 if self.dtypes else 1
        ptrs = (ctypes.c_void_p * maxbid)()
        for b, arr in self.data.items():
            ptrs[b] = arr.ctypes.data
        fn = getattr(self.lib, "g%d" % self._gi)
        fn.argtypes = [ctypes.c_void_p * maxbid]
        t0 = time.perf_counter()
        for _ in range(int(nrep)):
            if fn(ptrs) != 0:
                raise IndexError("oracle: gather/scatter index out of bounds")
            if cond is not None and not bool(self.data[cond][0]):
                break
        dt = time.perf_counter() - t0
        self._gi += 1
        for rec in recs:
            if rec[0] == "free":
                self.data.pop(rec[1], None)
        return dt

    def run_all(self):
        """Execute every group; returns (total seconds, {base id: data} of the bases alive at the end)."""
        total = 0.0
        while self._gi < len(self.groups):
            total += self.run_next_group()
        return total, dict(self.data)
