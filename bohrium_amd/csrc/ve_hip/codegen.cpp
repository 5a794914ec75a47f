#include "codegen.hpp"

#include <bohrium/jitk/fuser.hpp>

#include <algorithm>
#include <functional>
#include <cstring>
#include <map>
#include <sstream>
#include <stdexcept>

namespace bohrium {
namespace vehip {

using jitk::Block;
using jitk::InstrPtr;
using jitk::LoopB;

const char *strategy_name(Strategy s) {
    switch (s) {
        case Strategy::MAP: return "map";
        case Strategy::FULL_REDUCE: return "reduce";
        case Strategy::SCAN: return "scan";
        case Strategy::SERIAL: return "serial";
        case Strategy::ROW_REDUCE: return "row";
    }
    return "?";
}

const char *hip_type(bh_type t) {
    switch (t) {
        case bh_type::BOOL: return "bool";
        case bh_type::INT8: return "int8_t";
        case bh_type::INT16: return "int16_t";
        case bh_type::INT32: return "int32_t";
        case bh_type::INT64: return "int64_t";
        case bh_type::UINT8: return "uint8_t";
        case bh_type::UINT16: return "uint16_t";
        case bh_type::UINT32: return "uint32_t";
        case bh_type::UINT64: return "uint64_t";
        case bh_type::FLOAT32: return "float";
        case bh_type::FLOAT64: return "double";
        case bh_type::COMPLEX64: return "bh_c64";
        case bh_type::COMPLEX128: return "bh_c128";
        case bh_type::R123: return "uint64_t";
    }
    throw std::runtime_error("hip_type(): unknown type");
}

namespace {

// ------------------------------------------------------------------ helpers

bool sweeping_innermost_axis(const bh_instruction &i) {
    return bh_opcode_is_sweep(i.opcode) && i.sweep_axis() == i.operand[1].ndim - 1;
}

bool is_complex_op(const bh_instruction &i, int idx) { return bh_type_is_complex(i.operand_type(idx)); }

const char *cplx_prefix(bh_type t) { return t == bh_type::COMPLEX64 ? "make_c64" : "make_c128"; }

// Conversion with C99 assignment semantics.
std::string cast_expr(bh_type to, bh_type from, const std::string &e) {
    const bool cto = bh_type_is_complex(to), cfrom = bh_type_is_complex(from);
    if (!cto && !cfrom) return "((" + std::string(hip_type(to)) + ")(" + e + "))";
    if (cto && !cfrom) {
        const char *f = to == bh_type::COMPLEX64 ? "float" : "double";
        return std::string(cplx_prefix(to)) + "((" + f + ")(" + e + "), (" + f + ")0)";
    }
    if (cto && cfrom) {
        const char *f = to == bh_type::COMPLEX64 ? "float" : "double";
        return std::string(cplx_prefix(to)) + "((" + f + ")(" + e + ").real, (" + f + ")(" + e + ").imag)";
    }
    if (to == bh_type::BOOL) return "((" + e + ").real != 0 || (" + e + ").imag != 0)";
    return "((" + std::string(hip_type(to)) + ")(" + e + ").real)";
}

// Scalar body of one instruction; `ops` are the rendered operands in the
// reference's order (out, in1, in2, ...).  Semantics follow the reference's
// write_operation() for C99 (core/jitk/instruction.cpp:58-544).
std::string hip_operation(const bh_instruction &instr, const std::vector<std::string> &ops, bool math64) {
    std::ostringstream o;
    const auto &a = ops;
    auto bin = [&](const char *op) { o << a[0] << " = " << a[1] << " " << op << " " << a[2] << ";"; };
    auto un = [&](const char *fn, const char *cfn) {
        if (bh_type_is_complex(instr.operand_type(1)) && cfn != nullptr)
            o << a[0] << " = " << cfn << "(" << a[1] << ");";
        else
            o << a[0] << " = " << fn << "(" << a[1] << ");";
    };
    switch (instr.opcode) {
        case BH_BITWISE_AND: bin("&"); break;
        case BH_BITWISE_AND_REDUCE: o << a[0] << " = " << a[0] << " & " << a[1] << ";"; break;
        case BH_BITWISE_OR: bin("|"); break;
        case BH_BITWISE_OR_REDUCE: o << a[0] << " = " << a[0] << " | " << a[1] << ";"; break;
        case BH_BITWISE_XOR: bin("^"); break;
        case BH_BITWISE_XOR_REDUCE: o << a[0] << " = " << a[0] << " ^ " << a[1] << ";"; break;
        case BH_LOGICAL_NOT: o << a[0] << " = !" << a[1] << ";"; break;
        case BH_LOGICAL_OR: bin("||"); break;
        case BH_LOGICAL_OR_REDUCE: o << a[0] << " = " << a[0] << " || " << a[1] << ";"; break;
        case BH_LOGICAL_AND: bin("&&"); break;
        case BH_LOGICAL_AND_REDUCE: o << a[0] << " = " << a[0] << " && " << a[1] << ";"; break;
        case BH_LOGICAL_XOR: o << a[0] << " = !" << a[1] << " != !" << a[2] << ";"; break;
        case BH_LOGICAL_XOR_REDUCE: o << a[0] << " = !" << a[0] << " != !" << a[1] << ";"; break;
        case BH_LEFT_SHIFT: bin("<<"); break;
        case BH_RIGHT_SHIFT: bin(">>"); break;
        case BH_GREATER: bin(">"); break;
        case BH_GREATER_EQUAL: bin(">="); break;
        case BH_LESS: bin("<"); break;
        case BH_LESS_EQUAL: bin("<="); break;
        case BH_MAXIMUM: o << a[0] << " = " << a[1] << " > " << a[2] << " ? " << a[1] << " : " << a[2] << ";"; break;
        case BH_MAXIMUM_REDUCE:
            o << a[0] << " = " << a[0] << " > " << a[1] << " ? " << a[0] << " : " << a[1] << ";";
            break;
        case BH_MINIMUM: o << a[0] << " = " << a[1] << " < " << a[2] << " ? " << a[1] << " : " << a[2] << ";"; break;
        case BH_MINIMUM_REDUCE:
            o << a[0] << " = " << a[0] << " < " << a[1] << " ? " << a[0] << " : " << a[1] << ";";
            break;
        case BH_INVERT:
            if (instr.operand[0].base->dtype() == bh_type::BOOL)
                o << a[0] << " = !" << a[1] << ";";
            else
                o << a[0] << " = ~" << a[1] << ";";
            break;
        case BH_MOD:
            if (bh_type_is_float(instr.operand[0].base->dtype()))
                o << a[0] << " = fmod(" << a[1] << ", " << a[2] << ");";
            else
                o << a[0] << " = " << a[1] << " % " << a[2] << ";";
            break;
        case BH_REMAINDER: {
            const bh_type t = instr.operand[0].base->dtype();
            if (bh_type_is_float(t)) {
                o << a[0] << " = " << a[1] << " - floor(" << a[1] << " / " << a[2] << ") * " << a[2] << ";";
            } else if (bh_type_is_unsigned_integer(t)) {
                o << a[0] << " = " << a[1] << " % " << a[2] << ";";
            } else {  // Python/NumPy sign convention
                o << a[0] << " = ((" << a[1] << " > 0) == (" << a[2] << " > 0) || (" << a[1] << " % " << a[2]
                  << ") == 0) ? (" << a[1] << " % " << a[2] << ") : (" << a[1] << " % " << a[2] << ") + " << a[2]
                  << ";";
            }
            break;
        }
        case BH_RINT: un("rint", nullptr); break;
        case BH_EXP2: un("exp2", nullptr); break;
        case BH_EXPM1: un("expm1", nullptr); break;
        case BH_LOG1P: un("log1p", nullptr); break;
        case BH_ARCSIN: un("asin", nullptr); break;
        case BH_ARCCOS: un("acos", nullptr); break;
        case BH_ARCTAN: un("atan", nullptr); break;
        case BH_ARCTAN2: o << a[0] << " = atan2(" << a[1] << ", " << a[2] << ");"; break;
        case BH_ARCSINH: un("asinh", nullptr); break;
        case BH_ARCCOSH: un("acosh", nullptr); break;
        case BH_ARCTANH: un("atanh", nullptr); break;
        case BH_FLOOR: un("floor", nullptr); break;
        case BH_CEIL: un("ceil", nullptr); break;
        case BH_TRUNC: un("trunc", nullptr); break;
        case BH_LOG2: un("log2", nullptr); break;
        case BH_ISNAN:
        case BH_ISINF:
        case BH_ISFINITE: {
            const char *fn = instr.opcode == BH_ISNAN ? "isnan" : (instr.opcode == BH_ISINF ? "isinf" : "isfinite");
            const bh_type t = instr.operand_type(1);
            if (bh_type_is_complex(t)) {
                o << a[0] << " = " << fn << "((" << a[1] << ").real);";
            } else if (bh_type_is_float(t)) {
                o << a[0] << " = " << fn << "(" << a[1] << ");";
            } else {
                o << a[0] << " = " << (instr.opcode == BH_ISFINITE ? "true" : "false") << ";";
            }
            break;
        }
        case BH_CONJ: o << a[0] << " = bh_conj(" << a[1] << ");"; break;
        case BH_RANGE:
        case BH_RANDOM:
        case BH_GATHER:
        case BH_SCATTER: o << a[0] << " = " << a[1] << ";"; break;
        case BH_COND_SCATTER: o << "if (" << a[2] << ") { " << a[0] << " = " << a[1] << "; }"; break;
        case BH_SIN: un("sin", "bh_csin"); break;
        case BH_COS: un("cos", "bh_ccos"); break;
        case BH_TAN: un("tan", "bh_ctan"); break;
        case BH_SINH: un("sinh", "bh_csinh"); break;
        case BH_COSH: un("cosh", "bh_ccosh"); break;
        case BH_TANH: un("tanh", "bh_ctanh"); break;
        case BH_EXP: un(math64 && instr.operand_type(1) == bh_type::FLOAT64 ? "bh_exp64" : "exp", "bh_cexp"); break;
        case BH_ABSOLUTE: {
            const bh_type t = instr.operand_type(1);
            if (t == bh_type::BOOL || bh_type_is_unsigned_integer(t)) {
                o << a[0] << " = " << a[1] << ";";
            } else if (bh_type_is_float(t)) {
                o << a[0] << " = fabs(" << a[1] << ");";
            } else if (t == bh_type::INT64) {
                o << a[0] << " = llabs(" << a[1] << ");";
            } else if (bh_type_is_complex(t)) {
                o << a[0] << " = abs((int)(" << a[1] << ").real);";  // C99 cast of a complex to int
            } else {
                o << a[0] << " = abs((int)" << a[1] << ");";
            }
            break;
        }
        case BH_SQRT: un("sqrt", "bh_csqrt"); break;
        case BH_LOG: un(math64 && instr.operand_type(1) == bh_type::FLOAT64 ? "bh_log64" : "log", "bh_clog"); break;
        case BH_NOT_EQUAL: bin("!="); break;
        case BH_EQUAL: bin("=="); break;
        case BH_POWER:
            if (bh_type_is_complex(instr.operand_type(0))) {
                o << a[0] << " = bh_cpow(" << a[1] << ", " << a[2] << ");";
            } else if (bh_type_is_integer(instr.operand_type(0)) || instr.operand_type(0) == bh_type::BOOL) {
                o << a[0] << " = bh_pow_int((double) " << a[1] << ", (double)" << a[2] << ");";
            } else {
                o << a[0] << " = pow(" << a[1] << ", " << a[2] << ");";
            }
            break;
        case BH_ADD: bin("+"); break;
        case BH_ADD_REDUCE: o << a[0] << " += " << a[1] << ";"; break;
        case BH_ADD_ACCUMULATE: bin("+"); break;
        case BH_SUBTRACT: bin("-"); break;
        case BH_MULTIPLY: bin("*"); break;
        case BH_MULTIPLY_REDUCE: o << a[0] << " *= " << a[1] << ";"; break;
        case BH_MULTIPLY_ACCUMULATE: bin("*"); break;
        case BH_DIVIDE:
            if (bh_type_is_signed_integer(instr.operand[0].base->dtype())) {  // NumPy floor division
                o << a[0] << " = ((" << a[1] << " > 0) != (" << a[2] << " > 0) && (" << a[1] << " % " << a[2]
                  << ") != 0)?(" << a[1] << " / " << a[2] << " - 1):(" << a[1] << " / " << a[2] << ");";
            } else {
                bin("/");
            }
            break;
        case BH_IDENTITY:
            o << a[0] << " = " << cast_expr(instr.operand_type(0), instr.operand_type(1), a[1]) << ";";
            break;
        case BH_LOG10:
            if (bh_type_is_complex(instr.operand_type(0))) {
                const bh_type t = instr.operand_type(0);
                o << a[0] << " = bh_clog(" << a[1] << ") / " << cplx_prefix(t) << "(log(10.0f), 0);";
            } else {
                o << a[0] << " = log10(" << a[1] << ");";
            }
            break;
        case BH_REAL: o << a[0] << " = bh_creal(" << a[1] << ");"; break;
        case BH_IMAG: o << a[0] << " = bh_cimag(" << a[1] << ");"; break;
        case BH_SIGN: {
            const bh_type t = instr.operand_type(0);
            if (bh_type_is_complex(t)) {
                const char *ft = t == bh_type::COMPLEX64 ? "float" : "double";
                o << "{ " << ft << " re_ = (" << a[1] << ").real, im_ = (" << a[1] << ").imag; " << a[0] << " = "
                  << cplx_prefix(t) << "(re_ == 0 ? ((im_ > 0) - (0 > im_)) : ((re_ > 0) - (0 > re_)), 0); }";
            } else {
                o << a[0] << " = ((" << a[1] << " > 0) - (0 > " << a[1] << "));";
            }
            break;
        }
        default:
            throw std::runtime_error(std::string("hip engine: instruction not supported: ") +
                                     bh_opcode_text(instr.opcode));
    }
    return o.str();
}

// The binary operator of a reduction/accumulate as an expression of (x, y).
std::string combine_expr(bh_opcode op, const std::string &x, const std::string &y) {
    switch (op) {
        case BH_ADD_REDUCE:
        case BH_ADD_ACCUMULATE: return "(" + x + " + " + y + ")";
        case BH_MULTIPLY_REDUCE:
        case BH_MULTIPLY_ACCUMULATE: return "(" + x + " * " + y + ")";
        case BH_MINIMUM_REDUCE: return "(" + x + " < " + y + " ? " + x + " : " + y + ")";
        case BH_MAXIMUM_REDUCE: return "(" + x + " > " + y + " ? " + x + " : " + y + ")";
        case BH_LOGICAL_AND_REDUCE: return "(" + x + " && " + y + ")";
        case BH_LOGICAL_OR_REDUCE: return "(" + x + " || " + y + ")";
        case BH_LOGICAL_XOR_REDUCE: return "(!" + x + " != !" + y + ")";
        case BH_BITWISE_AND_REDUCE: return "(" + x + " & " + y + ")";
        case BH_BITWISE_OR_REDUCE: return "(" + x + " | " + y + ")";
        case BH_BITWISE_XOR_REDUCE: return "(" + x + " ^ " + y + ")";
        default: throw std::runtime_error("combine_expr(): not a sweep");
    }
}

std::string literal(const bh_constant &c) {
    std::ostringstream s;
    c.write_literal(s);
    return s.str();
}

// A view key that ignores size-1 dimensions (which never change an address).
struct VKey {
    const bh_base *base;
    int64_t start;
    std::vector<std::pair<int64_t, int64_t> > dims;
    bool operator<(const VKey &o) const {
        if (base != o.base) return base < o.base;
        if (start != o.start) return start < o.start;
        return dims < o.dims;
    }
};
VKey vkey(const bh_view &v) {
    VKey k{v.base, v.start, {}};
    for (int64_t d = 0; d < v.ndim; ++d)
        if (v.shape[d] > 1) k.dims.emplace_back(v.shape[d], v.stride[d]);
    return k;
}

// Key of an offset-and-strides argument group: views with equal start and strides share it.
typedef std::pair<int64_t, std::vector<int64_t> > OSKey;
OSKey oskey(const bh_view &v) { return {v.start, std::vector<int64_t>(v.stride.begin(), v.stride.end())}; }

struct Scope {
    const Scope *parent = nullptr;
    std::set<const bh_base *> tmps;
    std::set<VKey> srep;
    // Scalar-replaced reduction outputs declared by this scope and stored when it closes.
    std::vector<std::pair<const bh_view *, int64_t> > wb;
    // Temporaries that stand for an accumulator of an enclosing skeleton (collapsed reductions).
    std::map<const bh_base *, std::string> alias;
    // Input views already loaded into a register of this name (unrolled grid-stride groups).
    std::map<VKey, std::string> preload;
    const std::string *preloadOf(const bh_view &v) const {
        if (!preload.empty()) {
            auto it = preload.find(vkey(v));
            if (it != preload.end()) return &it->second;
        }
        return parent ? parent->preloadOf(v) : nullptr;
    }
    bool isTmp(const bh_base *b) const { return tmps.count(b) || (parent && parent->isTmp(b)); }
    bool isSrep(const bh_view &v) const { return srep.count(vkey(v)) || (parent && parent->isSrep(v)); }
    const std::string *aliasOf(const bh_base *b) const {
        auto it = alias.find(b);
        if (it != alias.end()) return &it->second;
        return parent ? parent->aliasOf(b) : nullptr;
    }
    bool isArray(const bh_view &v) const { return !(isTmp(v.base) || isSrep(v) || aliasOf(v.base)); }
    bool isDeclared(const bh_view &v) const { return isTmp(v.base) || isSrep(v); }
};

void indent(std::ostringstream &o, int n) { o << std::string(n * 2, ' '); }

// ------------------------------------------------------------------ writer

struct KernelParam {
    std::string type;  // value type: "double *", "int64_t", ...
    std::string name;
    bool is_ptr;
};

class Writer {
public:
    Writer(const LoopB &unit, const WriterConfig &cfg) : unit(unit), cfg(cfg) { collect(); }

    KernelPlan plan(Strategy s);

private:
    const LoopB &unit;
    const WriterConfig &cfg;
    KernelPlan kp;

    std::map<const bh_base *, int> bid;
    std::vector<const bh_base *> base_order;
    std::set<const bh_base *> always_array;
    std::set<bh_base *> temps;
    std::map<VKey, int> vid;
    std::map<OSKey, int> osid;
    std::vector<const bh_view *> os_views;
    // Sliding views get an offset argument of their own, re-read from (origin, operand) at launch.
    std::map<const bh_view *, std::pair<int64_t, int> > sliding_src;
    std::map<const LoopB *, int> lid;
    std::vector<const LoopB *> loops;
    std::vector<InstrPtr> const_instrs;
    std::vector<InstrPtr> const_params;  // one instruction per distinct constant: parameter c<j>
    std::map<int64_t, int> const_of_origin;
    std::set<const bh_base *> index_checked;  // bases needing their element count (gather/scatter)
    // Instructions a skeleton implements itself (collapsed reductions and their identities).
    std::set<const bh_instruction *> skip;
    // ROW_REDUCE: the inner loop whose iterations are spread over the 64 lanes of a wave.
    std::set<const LoopB *> lane_loops;  // loops strided by the lanes of a wave (row skeleton)
    bool fence_lanes = false;            // order the lanes' memory accesses after each lane loop
    // Instruction replaced by "store its input into `scan_sink`" in the scan skeleton.
    const bh_instruction *scan_instr = nullptr;
    std::string scan_sink;

    std::ostringstream body;
    std::string main_body;  // FULL_REDUCE: the main kernel's body while `body` holds the finalize's

    void collect();
    int view_id(const bh_view &v) {
        auto it = vid.find(vkey(v));
        if (it != vid.end()) return it->second;
        const int id = static_cast<int>(vid.size());
        vid[vkey(v)] = id;
        return id;
    }
    int os_id(const bh_view &v) {
        if (v.hasSlide()) {
            OSKey k{INT64_MIN, {static_cast<int64_t>(reinterpret_cast<uintptr_t>(&v))}};
            auto it = osid.find(k);
            if (it != osid.end()) return it->second;
            const int id = static_cast<int>(os_views.size());
            osid[k] = id;
            os_views.push_back(&v);
            return id;
        }
        auto it = osid.find(oskey(v));
        if (it != osid.end()) return it->second;
        const int id = static_cast<int>(os_views.size());
        osid[oskey(v)] = id;
        os_views.push_back(&v);
        return id;
    }
    std::string name(const Scope &sc, const bh_view &v) {
        std::ostringstream s;
        if (const std::string *a = sc.aliasOf(v.base)) return *a;
        if (sc.isTmp(v.base)) {
            s << "t" << bid.at(v.base);
        } else if (sc.isSrep(v)) {
            s << "s" << bid.at(v.base) << "_" << view_id(v);
        } else {
            s << "a" << bid.at(v.base);
        }
        return s.str();
    }
    // Strides are part of the plan key, so views with equal strides share one parameter (and the
    // compiler one product i * stride for all of them, e.g. the five loads of a stencil).
    std::map<int64_t, int> stride_ids;
    std::vector<int64_t> stride_vals;
    int stride_id(int64_t stride) {
        auto it = stride_ids.find(stride);
        if (it != stride_ids.end()) return it->second;
        const int id = static_cast<int>(stride_vals.size());
        stride_ids[stride] = id;
        stride_vals.push_back(stride);
        return id;
    }
    std::string index(const bh_view &v, int hidden_axis = BH_MAXDIM, int off_axis = -1, int off = 0) {
        std::ostringstream s;
        const int k = os_id(v);
        s << "vo" << k;
        if (!v.is_scalar()) {
            // Strides are part of the plan key, so unit and zero strides are spelled out.
            for (int64_t d = 0; d < v.ndim; ++d) {
                int64_t t = d >= hidden_axis ? d + 1 : d;
                if (v.stride[d] == 0) continue;
                if (t == off_axis)
                    s << " + (i" << t << " + (i" << t << " == 0 ? 0 : " << off << "))";
                else
                    s << " + i" << t;
                if (v.stride[d] != 1) s << " * vs" << stride_id(v.stride[d]);
            }
        }
        return s.str();
    }
    std::string ref(const Scope &sc, const bh_view &v, int hidden_axis = BH_MAXDIM) {
        if (const std::string *p = sc.preloadOf(v)) return *p;
        std::string n = name(sc, v);
        if (sc.isArray(v)) n += "[" + index(v, hidden_axis) + "]";
        return n;
    }
    std::string const_ref(const bh_instruction &i) { return "c" + std::to_string(const_of_origin.at(i.origin_id)); }

    void declare_scope(const LoopB &l, Scope &sc, int ind);
    void write_back(const LoopB &l, Scope &sc, int ind);
    void emit_block(const Block &b, Scope &sc, int ind);
    void emit_blocks(const LoopB &l, Scope &sc, int ind);
    void emit_body(const LoopB &l, const Scope *parent, int ind);
    void emit_instr(const bh_instruction &i, Scope &sc, int ind);

    /* One element of a grid-stride loop: `decode(ind)` derives the loop indices from f_,
     * `compute(ind, outer)` emits the body in a scope whose parent is `outer`. */
    struct Element {
        std::function<void(int)> decode;
        std::function<void(int, const Scope *)> compute;
        const LoopB *body_loop = nullptr;  // innermost loop whose read-only inputs may be preloaded
        const Scope *outer = nullptr;
        /* A 2-D iteration (i0 outer, i1 inner of this extent): without unrolling, the two indices
         * are advanced by the grid stride instead of being divided out of f_ for every element. */
        std::string inner;
        /* decode is i0 = f_ (one rank): an unrolled group of consecutive f_ touches consecutive
         * elements of every unit-stride view. */
        bool unit = false;
    };
    /* Grid-stride over [0, tot).  With `parts`, the range is cut into min(8, bh_nblk) contiguous
     * parts, part x walked by the workgroups b with b mod 8 = x (the workgroups one XCD runs), so a
     * stencil's neighbouring rows are read by the workgroups of one L2. */
    void grid_stride(const std::string &tot, int unroll, const Element &e);
    bool parts = false;
    std::vector<const bh_view *> preloadable(const LoopB &l);
    void emit_map();
    void emit_row_reduce();
    void emit_full_reduce();
    void emit_scan();
    void emit_serial();
    std::vector<KernelParam> parameters();
};

void Writer::collect() {
    std::vector<InstrPtr> all = jitk::allInstr(unit);
    for (const InstrPtr &i : all) {
        for (const bh_view &v : i->operand) {
            if (v.isConstant()) continue;
            if (!bid.count(v.base)) {
                bid[v.base] = static_cast<int>(bid.size());
                base_order.push_back(v.base);
            }
        }
        if (bh_opcode_is_accumulate(i->opcode)) {
            always_array.insert(i->operand[0].base);
        } else if (i->opcode == BH_GATHER) {
            always_array.insert(i->operand[1].base);
            index_checked.insert(i->operand[1].base);
        } else if (i->opcode == BH_SCATTER || i->opcode == BH_COND_SCATTER) {
            always_array.insert(i->operand[0].base);
            index_checked.insert(i->operand[0].base);
        }
        if (i->has_constant()) const_instrs.push_back(i);
    }
    for (const bh_base *b : unit.getAllFrees()) {
        if (!bid.count(b)) {
            bid[b] = static_cast<int>(bid.size());
            base_order.push_back(b);
        }
    }
    std::sort(const_instrs.begin(), const_instrs.end(),
              [](const InstrPtr &a, const InstrPtr &b) { return a->origin_id < b->origin_id; });
    // Constants of equal type and value share one parameter (a loop body repeats 1.0, 0.5, its
    // coefficients ...; each parameter takes SGPRs, and past the SGPR budget the compiler spills them
    // to VGPR lanes and reads them back in every iteration).  The flush key records which constants
    // of a flush are equal, so a cached plan is reused only where the same ones are.
    for (size_t j = 0; j < const_instrs.size(); ++j) {
        const bh_constant &c = const_instrs[j]->constant;
        int id = -1;
        if (c.type != bh_type::R123) {
            for (size_t k = 0; k < const_params.size() && id < 0; ++k) {
                const bh_constant &d = const_params[k]->constant;
                if (d.type == c.type && std::memcmp(&d.value, &c.value, bh_type_size(c.type)) == 0)
                    id = static_cast<int>(k);
            }
        }
        if (id < 0) {
            id = static_cast<int>(const_params.size());
            const_params.push_back(const_instrs[j]);
        }
        const_of_origin[const_instrs[j]->origin_id] = id;
    }
    temps = unit.getAllTemps();
    std::set<bh_base *> non_temps = unit.getAllNonTemps();
    for (const bh_base *b : always_array) non_temps.insert(const_cast<bh_base *>(b));
    for (const InstrPtr &i : all)
        for (const bh_view &v : i->operand)
            if (!v.isConstant() && non_temps.count(v.base) &&
                std::find(kp.params.begin(), kp.params.end(), v.base) == kp.params.end())
                kp.params.push_back(v.base);
    for (const InstrPtr &i : all) {
        if (i->operand.empty() || i->operand[0].isConstant()) continue;
        if (non_temps.count(i->operand[0].base)) kp.written.insert(i->operand[0].base);
    }
    // Loop ids (pre-order) for the size arguments.
    std::vector<const LoopB *> stack{&unit};
    while (!stack.empty()) {
        const LoopB *l = stack.back();
        stack.pop_back();
        if (l->rank >= 0 && !lid.count(l)) {
            lid[l] = static_cast<int>(loops.size());
            loops.push_back(l);
        }
        for (auto it = l->_block_list.rbegin(); it != l->_block_list.rend(); ++it)
            if (!it->isInstr()) stack.push_back(&it->getLoop());
    }
    // Register every view's offset-and-stride group in instruction order.
    for (const InstrPtr &i : all) {
        for (size_t o = 0; o < i->operand.size(); ++o) {
            const bh_view &v = i->operand[o];
            if (v.isConstant()) continue;
            os_id(v);
            if (v.hasSlide()) sliding_src[&v] = {i->origin_id, static_cast<int>(o)};
        }
    }
}

void Writer::declare_scope(const LoopB &l, Scope &sc, int ind) {
    const std::set<bh_base *> local_tmps = l.getLocalTemps();
    for (const InstrPtr &i : jitk::allInstr(l)) {
        for (const bh_view &v : i->operand) {
            if (v.isConstant() || !local_tmps.count(v.base)) continue;
            if (sc.isDeclared(v) || always_array.count(v.base)) continue;
            sc.tmps.insert(v.base);
            indent(body, ind);
            body << hip_type(v.base->dtype()) << " t" << bid.at(v.base) << ";\n";
        }
    }
    for (const Block &b : l._block_list) {
        if (b.isInstr()) continue;
        for (const InstrPtr &i : jitk::allLocalInstr(b.getLoop())) {
            if (!(bh_opcode_is_reduction(i->opcode) && sweeping_innermost_axis(*i))) continue;
            const bh_view &v = i->operand[0];
            if (sc.isDeclared(v) || always_array.count(v.base)) continue;
            sc.srep.insert(vkey(v));
            sc.wb.emplace_back(&v, i->sweep_axis());
            indent(body, ind);
            body << hip_type(v.base->dtype()) << " s" << bid.at(v.base) << "_" << view_id(v) << ";\n";
        }
    }
}

// Store the scope's scalar-replaced reduction outputs once, after the loops that sweep them.
void Writer::write_back(const LoopB &, Scope &sc, int ind) {
    for (const auto &w : sc.wb) {
        const bh_view &v = *w.first;
        indent(body, ind);
        body << "a" << bid.at(v.base) << "[" << index(v, static_cast<int>(w.second)) << "] = s" << bid.at(v.base) << "_"
             << view_id(v) << ";\n";
    }
    sc.wb.clear();
}

void Writer::emit_block(const Block &b, Scope &sc, int ind) {
    if (b.isInstr()) {
        if (b.getInstr() != nullptr && !bh_opcode_is_system(b.getInstr()->opcode) && !skip.count(b.getInstr().get()))
            emit_instr(*b.getInstr(), sc, ind);
        return;
    }
    const LoopB &c = b.getLoop();
    const bool lanes = lane_loops.count(&c) != 0;
    if (lanes && fence_lanes) {
        indent(body, ind);
        body << "__builtin_amdgcn_fence(__ATOMIC_SEQ_CST, \"workgroup\");\n";
    }
    indent(body, ind);
    if (lanes) {
        body << "for (int64_t i" << c.rank << " = lane_; i" << c.rank << " < n" << lid.at(&c) << "; i" << c.rank
             << " += 64) {\n";
    } else {
        body << "for (int64_t i" << c.rank << " = 0; i" << c.rank << " < n" << lid.at(&c) << "; ++i" << c.rank << ") {\n";
    }
    emit_body(c, &sc, ind + 1);
    indent(body, ind);
    body << "}\n";
    if (!lanes) return;
    if (fence_lanes) {
        // Elements written by one lane are read or overwritten by another (before and after every
        // lane loop): the wave's memory accesses so far complete before it goes on (the lanes share
        // one CU's L1).
        indent(body, ind);
        body << "__builtin_amdgcn_fence(__ATOMIC_SEQ_CST, \"workgroup\");\n";
    }
    // Fold the lanes' partial reductions; every lane ends with the row's value.
    for (const InstrPtr &sw : c._sweeps) {
        const bh_view &v = sw->operand[0];
        const std::string var = ref(sc, v, static_cast<int>(sw->sweep_axis()));
        if (var.empty() || var[0] == 'a') throw std::logic_error("row skeleton: a reduction output is not in a register");
        const std::string T = hip_type(v.base->dtype());
        indent(body, ind);
        body << var << " = bh_wave_reduce(" << var << ", [](" << T << " x, " << T << " y) -> " << T << " { return "
             << combine_expr(sw->opcode, "x", "y") << "; });\n";
    }
}

void Writer::emit_blocks(const LoopB &l, Scope &sc, int ind) {
    for (const Block &b : l._block_list) emit_block(b, sc, ind);
}

void Writer::emit_body(const LoopB &l, const Scope *parent, int ind) {
    if (l.isSystemOnly()) return;
    Scope sc;
    sc.parent = parent;
    declare_scope(l, sc, ind);
    emit_blocks(l, sc, ind);
    write_back(l, sc, ind);
}

void Writer::emit_instr(const bh_instruction &instr, Scope &sc, int ind) {
    std::vector<std::string> ops;
    indent(body, ind);
    if (&instr == scan_instr) {
        body << scan_sink << " = " << ref(sc, instr.operand[1]) << ";\n";
        return;
    }
    if (instr.opcode == BH_RANGE) {
        ops.push_back(ref(sc, instr.operand[0]));
        ops.push_back("(" + index(instr.operand[0]) + ")");
    } else if (instr.opcode == BH_RANDOM) {
        ops.push_back(ref(sc, instr.operand[0]));
        const std::string c = const_ref(instr);
        ops.push_back("bh_random123(" + c + "_start, " + c + "_key, " + index(instr.operand[0]) + ")");
    } else if (instr.opcode == BH_GATHER) {
        // out[...] = in[in.start + idx[...]], with the index checked against the base.  No branch: an
        // index outside the base loads element 0 and marks bh_bad_, which raises the error flag when
        // the function ends; a body without branches keeps the loads of several elements in flight.
        const bh_view &in = instr.operand[1];
        const int k = os_id(in);
        const std::string out = ref(sc, instr.operand[0]);
        body << "{ const int64_t g_ = vo" << k << " + (int64_t)(" << ref(sc, instr.operand[2]) << "); "
             << "const bool ok_ = (uint64_t)g_ < (uint64_t)nb" << bid.at(in.base) << "; bh_bad_ |= !ok_; "
             << out << " = a" << bid.at(in.base) << "[ok_ ? g_ : (int64_t)0]; }\n";
        kp.uses_gather_flag = true;
        kp.uses_index_check = true;
        kp.uses_error_flag = true;
        return;
    } else if (instr.opcode == BH_SCATTER || instr.opcode == BH_COND_SCATTER) {
        const bh_view &out = instr.operand[0];
        const int k = os_id(out);
        body << "{ const int64_t g_ = vo" << k << " + (int64_t)(" << ref(sc, instr.operand[2]) << "); ";
        if (instr.opcode == BH_COND_SCATTER) body << "if (" << ref(sc, instr.operand[3]) << ") ";
        body << "{ if (g_ < 0 || g_ >= nb" << bid.at(out.base) << ") { atomicOr(bh_err, BH_ERR_INDEX); } else { a"
             << bid.at(out.base) << "[g_] = " << ref(sc, instr.operand[1]) << "; } } }\n";
        kp.uses_index_check = true;
        kp.uses_error_flag = true;
        return;
    } else if (bh_opcode_is_accumulate(instr.opcode)) {
        ops.push_back(ref(sc, instr.operand[0]));
        ops.push_back(name(sc, instr.operand[0]) + "[" +
                      index(instr.operand[0], BH_MAXDIM, instr.sweep_axis(), -1) + "]");
        ops.push_back(ref(sc, instr.operand[1]));
    } else {
        for (size_t o = 0; o < instr.operand.size(); ++o) {
            const bh_view &v = instr.operand[o];
            if (v.isConstant()) {
                ops.push_back(const_ref(instr));
            } else if (o == 0 && bh_opcode_is_reduction(instr.opcode) && instr.operand[1].ndim > 1) {
                ops.push_back(ref(sc, v, instr.sweep_axis()));
            } else {
                ops.push_back(ref(sc, v));
            }
        }
    }
    body << hip_operation(instr, ops, cfg.math64) << "\n";
}

std::vector<KernelParam> Writer::parameters() {
    std::vector<KernelParam> params;
    auto lit = [&](const std::string &type, const std::string &name, const void *p, int size, int align) {
        params.push_back({type, name, false});
        ArgSpec a;
        a.kind = ArgSpec::LITERAL;
        a.bytes.assign(static_cast<const uint8_t *>(p), static_cast<const uint8_t *>(p) + size);
        a.size = size;
        a.align = align;
        kp.args.push_back(a);
    };
    for (bh_base *b : kp.params) {
        params.push_back({std::string(hip_type(b->dtype())) + " *", "a" + std::to_string(bid.at(b)), true});
        ArgSpec a;
        a.kind = ArgSpec::BASE_PTR;
        a.base = b;
        kp.args.push_back(a);
    }
    for (size_t k = 0; k < os_views.size(); ++k) {
        const bh_view &v = *os_views[k];
        lit("int64_t", "vo" + std::to_string(k), &v.start, 8, 8);
        auto src = sliding_src.find(&v);
        if (src != sliding_src.end()) {
            kp.args.back().view_origin = src->second.first;
            kp.args.back().view_operand = src->second.second;
        } else if (v.hasSlide()) {
            kp.reusable = false;
        }
    }
    for (size_t j = 0; j < stride_vals.size(); ++j) lit("int64_t", "vs" + std::to_string(j), &stride_vals[j], 8, 8);
    for (size_t j = 0; j < loops.size(); ++j) lit("int64_t", "n" + std::to_string(j), &loops[j]->size, 8, 8);
    for (const bh_base *b : index_checked) {
        const int64_t n = b->nelem();
        lit("int64_t", "nb" + std::to_string(bid.at(b)), &n, 8, 8);
    }
    for (size_t j = 0; j < const_params.size(); ++j) {
        const bh_constant &c = const_params[j]->constant;
        const std::string cn = "c" + std::to_string(j);
        if (c.type == bh_type::R123) {
            lit("uint64_t", cn + "_start", &c.value.r123.start, 8, 8);
            kp.args.back().const_origin = const_params[j]->origin_id;
            kp.args.back().const_part = 0;
            lit("uint64_t", cn + "_key", &c.value.r123.key, 8, 8);
            kp.args.back().const_origin = const_params[j]->origin_id;
            kp.args.back().const_part = 1;
        } else {
            const int size = bh_type_size(c.type);
            const int align = bh_type_is_complex(c.type) ? size / 2 : size;
            lit(hip_type(c.type), cn, &c.value, size, align);
            kp.args.back().const_origin = const_params[j]->origin_id;
        }
    }
    if (kp.uses_error_flag) {
        params.push_back({"unsigned int *", "bh_err", true});
        ArgSpec a;
        a.kind = ArgSpec::ERRFLAG_PTR;
        kp.args.push_back(a);
    }
    return params;
}

// ---- MAP: grid-stride over the flattened leading parallel ranks
// Input views of an innermost loop that can be loaded before any of its instructions run:
// reads of parameter arrays that no instruction of the loop writes (not gathered, scattered
// or accumulated).  Distinct views in first-use order.
std::vector<const bh_view *> Writer::preloadable(const LoopB &l) {
    std::set<const bh_base *> written;
    for (const InstrPtr &i : jitk::allInstr(l))
        if (!i->operand.empty() && !i->operand[0].isConstant()) written.insert(i->operand[0].base);
    std::vector<const bh_view *> out;
    std::set<VKey> seen;
    for (const InstrPtr &i : jitk::allInstr(l)) {
        if (bh_opcode_is_system(i->opcode) || i->opcode == BH_GATHER) continue;
        for (size_t o = 1; o < i->operand.size(); ++o) {
            const bh_view &v = i->operand[o];
            if (v.isConstant() || written.count(v.base) || always_array.count(v.base)) continue;
            if (std::find(kp.params.begin(), kp.params.end(), v.base) == kp.params.end()) continue;
            if (seen.insert(vkey(v)).second) out.push_back(&v);
        }
    }
    return out;
}

// Grid-stride loop over the flat index f_ in [0, tot).  With unroll > 1, full groups of
// `unroll` elements first load every read-only input of all of them into registers and then
// compute them, so a wave has all of its loads in flight at once; the remaining elements run one
// at a time.  A group is `unroll` consecutive elements on a one-rank walk without XCD parts (a
// unit-stride view's loads merge into wide loads), else `unroll` elements strided by the grid.

// Full reductions with at most this many instructions in their loop count as light (half the grid).
constexpr size_t kLightReduceInstrs = 16;

namespace {
// A base read through two or more different views in `l` (a stencil: neighbouring elements of one
// array read at once).
bool shifted_reads(const LoopB &l) {
    std::map<const bh_base *, std::set<std::vector<int64_t> > > views;
    for (const InstrPtr &i : jitk::allInstr(l)) {
        for (size_t o = 1; o < i->operand.size(); ++o) {
            const bh_view &v = i->operand[o];
            if (v.isConstant()) continue;
            std::vector<int64_t> key{v.start, v.ndim};
            for (int64_t d = 0; d < v.ndim; ++d) {
                key.push_back(v.shape[d]);
                key.push_back(v.stride[d]);
            }
            auto &set = views[v.base];
            set.insert(key);
            if (set.size() > 1) return true;
        }
    }
    return false;
}
} // namespace

void Writer::grid_stride(const std::string &tot_in, int unroll, const Element &e) {
    std::string tot = tot_in;
    if (parts) {
        // Part x of X = min(8, bh_nblk) holds [x * pt_, (x + 1) * pt_); its workgroups are
        // b = x, x + X, ... (q_ or q_ + 1 of them), each taking every whole workgroup's worth in turn.
        body << "  const unsigned px_ = bh_nblk < BH_XCDS ? bh_nblk : BH_XCDS;\n";
        body << "  const unsigned pw_ = bh_nblk / px_ + (bh_bid % px_ < bh_nblk % px_ ? 1u : 0u);\n";
        body << "  const int64_t pt_ = ((" << tot << " + px_ - 1) / px_ + BH_WG - 1) / BH_WG * BH_WG;\n";
        body << "  const int64_t p0_ = (int64_t)(bh_bid % px_) * pt_;\n";
        body << "  const int64_t pe_ = p0_ + pt_ < " << tot << " ? p0_ + pt_ : " << tot << ";\n";
        body << "  const int64_t st_ = (int64_t)pw_ * BH_WG;\n";
        body << "  int64_t g_ = p0_ + (int64_t)(bh_bid / px_) * BH_WG + threadIdx.x;\n";
        tot = "pe_";
    } else {
        body << "  const int64_t st_ = (int64_t)bh_nblk * BH_WG;\n";
        body << "  int64_t g_ = (int64_t)bh_bid * BH_WG + threadIdx.x;\n";
    }
    if (unroll > 1 && e.unit && !parts) {
        // Groups of `unroll` consecutive elements: group q_ is [q_ * unroll, q_ * unroll + unroll), the
        // groups are walked grid-stride, so a lane's loads of one unit-stride view are adjacent (the
        // compiler merges them into one wide load, e.g. 16 bytes for 4 floats) and a wave's lanes read
        // consecutive 4 * unroll-byte pieces.  The last tot % unroll elements go to the first threads.
        const std::vector<const bh_view *> reads =
            e.body_loop != nullptr ? preloadable(*e.body_loop) : std::vector<const bh_view *>();
        const std::string U = std::to_string(unroll);
        body << "  const int64_t g0_ = g_, ng_ = " << tot << " / " << U << ";\n";
        body << "  for (; g_ < ng_; g_ += st_) {\n";
        for (int u = 0; u < unroll; ++u)
            for (size_t k = 0; k < reads.size(); ++k)
                body << "    " << hip_type(reads[k]->base->dtype()) << " ld" << k << "_" << u << ";\n";
        for (int u = 0; u < unroll; ++u) {
            if (reads.empty()) break;
            body << "    {\n      const int64_t f_ = g_ * " << U << " + " << u << ";\n";
            e.decode(3);
            for (size_t k = 0; k < reads.size(); ++k)
                body << "      ld" << k << "_" << u << " = " << ref(*e.outer, *reads[k]) << ";\n";
            body << "    }\n";
        }
        for (int u = 0; u < unroll; ++u) {
            Scope pre;
            pre.parent = e.outer;
            for (size_t k = 0; k < reads.size(); ++k)
                pre.preload[vkey(*reads[k])] = "ld" + std::to_string(k) + "_" + std::to_string(u);
            body << "    {\n      const int64_t f_ = g_ * " << U << " + " << u << ";\n";
            e.decode(3);
            e.compute(3, &pre);
            body << "    }\n";
        }
        body << "  }\n";
        body << "  g_ = ng_ * " << U << " + g0_;\n";
        body << "  if (g_ < " << tot << ") {\n";
        body << "    const int64_t f_ = g_;\n";
        e.decode(2);
        e.compute(2, e.outer);
        body << "  }\n";
        return;
    }
    if (unroll > 1) {
        const std::vector<const bh_view *> reads =
            e.body_loop != nullptr ? preloadable(*e.body_loop) : std::vector<const bh_view *>();
        body << "  for (; g_ + " << unroll - 1 << " * st_ < " << tot << "; g_ += " << unroll << " * st_) {\n";
        for (int u = 0; u < unroll; ++u)
            for (size_t k = 0; k < reads.size(); ++k)
                body << "    " << hip_type(reads[k]->base->dtype()) << " ld" << k << "_" << u << ";\n";
        for (int u = 0; u < unroll; ++u) {
            if (reads.empty()) break;
            body << "    {\n      const int64_t f_ = g_ + " << u << " * st_;\n";
            e.decode(3);
            for (size_t k = 0; k < reads.size(); ++k)
                body << "      ld" << k << "_" << u << " = " << ref(*e.outer, *reads[k]) << ";\n";
            body << "    }\n";
        }
        for (int u = 0; u < unroll; ++u) {
            Scope pre;
            pre.parent = e.outer;
            for (size_t k = 0; k < reads.size(); ++k)
                pre.preload[vkey(*reads[k])] = "ld" + std::to_string(k) + "_" + std::to_string(u);
            body << "    {\n      const int64_t f_ = g_ + " << u << " * st_;\n";
            e.decode(3);
            e.compute(3, &pre);
            body << "    }\n";
        }
        body << "  }\n";
    }
    if (unroll <= 1 && !e.inner.empty()) {
        const std::string &n = e.inner;
        body << "  int64_t i0_ = g_ / " << n << ", i1_ = g_ - i0_ * " << n << ";\n";
        body << "  const int64_t sq_ = st_ / " << n << ", sr_ = st_ - sq_ * " << n << ";\n";
        body << "  for (; g_ < " << tot << "; g_ += st_) {\n";
        body << "    const int64_t i0 = i0_, i1 = i1_;\n";
        e.compute(2, e.outer);
        body << "    i1_ += sr_;\n    i0_ += sq_;\n    if (i1_ >= " << n << ") { i1_ -= " << n << "; ++i0_; }\n";
        body << "  }\n";
        return;
    }
    body << "  for (; g_ < " << tot << "; g_ += st_) {\n";
    body << "    const int64_t f_ = g_;\n";
    e.decode(2);
    e.compute(2, e.outer);
    body << "  }\n";
}

// The inner loop of a row reduction: the innermost parallel loop of `l0` holds only constant
// IDENTITY seeds of the reductions of one innermost sub-loop (and frees), and that sub-loop's
// sweeps are all reductions over its own rank.  Null when `l0` has another form.
// The loops a wave's 64 lanes stride when a block runs one wave per row: every loop directly
// inside the innermost leading parallel loop P (at least one with 32 or more iterations), whose own
// sweeps are reductions over their own, last axis; loops nested in them run inside one lane.  A base written by one of
// them is accessed by the others only through the identical view (so each element stays with the
// lane that wrote it), no reduction output is read inside its own loop or held in memory by an
// accumulate, gather or scatter, and P's own instructions are element-wise.  Empty: another form.
std::vector<const LoopB *> row_lane_loops(const LoopB &l0, bool *cross = nullptr) {
    if (cross != nullptr) *cross = false;
    const uint64_t p = std::max<uint64_t>(1, jitk::parallel_ranks(l0).first);
    const std::vector<const LoopB *> chain = jitk::get_first_loop_blocks(l0);
    if (chain.size() < p) return {};
    const LoopB &P = *chain[p - 1];
    std::vector<const LoopB *> loops;
    int64_t longest = 0;
    for (const Block &b : P._block_list) {
        if (b.isInstr()) {
            const InstrPtr &i = b.getInstr();
            if (bh_opcode_is_system(i->opcode)) continue;
            if (bh_opcode_is_sweep(i->opcode) || i->opcode == BH_GATHER || i->opcode == BH_SCATTER ||
                i->opcode == BH_COND_SCATTER)
                return {};
            continue;
        }
        const LoopB &C = b.getLoop();
        for (const InstrPtr &sw : C._sweeps)
            if (!bh_opcode_is_reduction(sw->opcode) || sw->sweep_axis() != C.rank ||
                sw->sweep_axis() != sw->operand[1].ndim - 1)
                return {};
        // (loops nested in C run in the lane that owns C's index, sweeps over their axes included)
        for (const InstrPtr &i : jitk::allInstr(C))
            if (i->opcode == BH_SCATTER || i->opcode == BH_COND_SCATTER) return {};
        loops.push_back(&C);
        longest = std::max(longest, C.size);
    }
    if (loops.empty() || longest < 32) return {};
    // Reduction outputs stay in registers: not read inside their loop, not an array of another kind.
    std::set<const bh_base *> red_outs;
    for (const LoopB *C : loops) {
        std::set<const bh_base *> outs;
        for (const InstrPtr &sw : C->_sweeps) outs.insert(sw->operand[0].base);
        for (const InstrPtr &i : jitk::allInstr(*C))
            for (size_t o = 1; o < i->operand.size(); ++o)
                if (!i->operand[o].isConstant() && outs.count(i->operand[o].base)) return {};
        red_outs.insert(outs.begin(), outs.end());
    }
    for (const InstrPtr &i : jitk::allInstr(l0))
        if (i->opcode == BH_GATHER && red_outs.count(i->operand[1].base)) return {};
    // ... and are read only through the view they are reduced into (the register), since the array
    // is written once, after the row.
    std::map<const bh_base *, const bh_view *> red_view;
    for (const LoopB *C : loops)
        for (const InstrPtr &sw : C->_sweeps) red_view[sw->operand[0].base] = &sw->operand[0];
    for (const InstrPtr &i : jitk::allInstr(l0))
        for (size_t o = 0; o < i->operand.size(); ++o) {
            const bh_view &v = i->operand[o];
            if (v.isConstant()) continue;
            auto r = red_view.find(v.base);
            if (r != red_view.end() && !(v == *r->second)) return {};
        }
    // Element ownership: when what one loop (or the row's own instructions) writes is touched by
    // another loop through another view, an element moves between lanes, and the lanes' memory
    // accesses are ordered by a fence after each lane loop (`cross`).
    // Outputs written in a part, except the wave-folded reductions of a lane loop itself (their
    // registers are stored after the row, and read only through their own view).
    std::set<const bh_instruction *> folded;
    for (const LoopB *C : loops)
        for (const InstrPtr &sw : C->_sweeps) folded.insert(sw.get());
    bool several = false;
    auto writes = [&](const std::vector<InstrPtr> &instrs, std::map<const bh_base *, const bh_view *> &w) {
        for (const InstrPtr &i : instrs) {
            if (i->operand.empty() || i->operand[0].isConstant() || folded.count(i.get())) continue;
            auto it = w.find(i->operand[0].base);
            if (it != w.end() && !(*it->second == i->operand[0])) several = true;
            w[i->operand[0].base] = &i->operand[0];
        }
    };
    std::vector<std::vector<InstrPtr> > parts;
    for (const LoopB *C : loops) parts.push_back(jitk::allInstr(*C));
    parts.push_back(jitk::allLocalInstr(P));
    // Inside one lane loop (its nested loops included) nothing orders the lanes: a base written
    // there is touched there only through the written view.
    for (size_t a = 0; a + 1 < parts.size(); ++a) {
        std::map<const bh_base *, const bh_view *> written;
        writes(parts[a], written);
        if (several) return {};  // one base written through two views
        for (const InstrPtr &i : parts[a])
            for (const bh_view &v : i->operand) {
                if (v.isConstant()) continue;
                auto w = written.find(v.base);
                if (w != written.end() && !(v == *w->second)) return {};
            }
    }
    for (size_t a = 0; a < parts.size(); ++a) {
        std::map<const bh_base *, const bh_view *> written;
        writes(parts[a], written);
        for (size_t b = 0; b < parts.size(); ++b) {
            if (a == b) continue;
            for (const InstrPtr &i : parts[b])
                for (const bh_view &v : i->operand) {
                    if (v.isConstant()) continue;
                    auto w = written.find(v.base);
                    if (w != written.end() && !(v == *w->second) && cross != nullptr) *cross = true;
                }
        }
    }
    return loops;
}

void Writer::emit_row_reduce() {
    const LoopB &l0 = unit._block_list.back().getLoop();
    const uint64_t p = std::max<uint64_t>(1, jitk::parallel_ranks(l0).first);
    const std::vector<const LoopB *> chain = jitk::get_first_loop_blocks(l0);
    bool cross = false;
    const std::vector<const LoopB *> loops = row_lane_loops(l0, &cross);
    if (loops.empty()) throw std::logic_error("row skeleton: block has another form");
    fence_lanes = cross;
    lane_loops.clear();
    lane_loops.insert(loops.begin(), loops.end());
    uint64_t rows = 1;
    for (uint64_t r = 0; r < p; ++r) rows *= static_cast<uint64_t>(chain[r]->size);
    int64_t longest = 0;
    for (const LoopB *C : loops) longest = std::max(longest, C->size);
    kp.work_items = rows * static_cast<uint64_t>(longest);
    Scope ksc;
    declare_scope(unit, ksc, 1);
    body << "  const int64_t rows_ = 1";
    for (uint64_t r = 0; r < p; ++r) body << " * n" << lid.at(chain[r]);
    body << ";\n";
    body << "  const int lane_ = threadIdx.x & 63;\n";
    body << "  for (int64_t f_ = (int64_t)bh_bid * BH_WAVES + (threadIdx.x >> 6); f_ < rows_; f_ += (int64_t)bh_nblk * BH_WAVES) {\n";
    body << "    int64_t r_ = f_;\n";
    for (int64_t r = static_cast<int64_t>(p) - 1; r >= 1; --r)
        body << "    const int64_t i" << r << " = r_ % n" << lid.at(chain[r]) << "; r_ /= n" << lid.at(chain[r]) << ";\n";
    body << "    const int64_t i0 = r_;\n";
    std::vector<Scope> scopes(p);
    const Scope *parent = &ksc;
    for (uint64_t r = 0; r < p; ++r) {
        scopes[r].parent = parent;
        body << std::string(4 + 2 * r, ' ') << "{\n";
        declare_scope(*chain[r], scopes[r], 3 + static_cast<int>(r));
        parent = &scopes[r];
    }
    Scope &ps = scopes[p - 1];
    const int ind = 2 + static_cast<int>(p);
    // The row's own instructions run in every lane (same values); each lane loop is strided by the
    // lanes and its reductions are folded across the wave right after it (emit_block).
    emit_blocks(*chain[p - 1], ps, ind);
    indent(body, ind);
    body << "if (lane_ == 0) {\n";
    write_back(*chain[p - 1], ps, ind + 1);
    indent(body, ind);
    body << "}\n";
    for (int64_t r = static_cast<int64_t>(p) - 1; r >= 0; --r) {
        if (r < static_cast<int64_t>(p) - 1) write_back(*chain[r], scopes[r], 3 + static_cast<int>(r));
        body << std::string(4 + 2 * r, ' ') << "}\n";
    }
    body << "  }\n";
    write_back(unit, ksc, 1);
    kp.grid = std::max<uint64_t>(1, std::min<uint64_t>((rows + 3) / 4, cfg.max_grid));
}

// ---- MAP: grid-stride over the flattened leading parallel ranks
void Writer::emit_map() {
    const LoopB &l0 = unit._block_list.back().getLoop();
    const uint64_t p = std::max<uint64_t>(1, jitk::parallel_ranks(l0).first);
    const std::vector<const LoopB *> chain = jitk::get_first_loop_blocks(l0);
    uint64_t total = 1;
    for (uint64_t r = 0; r < p; ++r) total *= static_cast<uint64_t>(chain[r]->size);
    kp.work_items = total;
    const bool idx32 = total < (1ull << 31);
    Scope ksc;
    declare_scope(unit, ksc, 1);
    body << "  const int64_t tot_ = 1";
    for (uint64_t r = 0; r < p; ++r) body << " * n" << lid.at(chain[r]);
    body << ";\n";
    // Only loops without inner sequential work are unrolled (their body is straight-line code).
    const bool flat = p == 1 ? l0.isInnermost() : chain[p - 1]->isInnermost();
    Element e;
    e.outer = &ksc;
    e.body_loop = flat ? chain[p - 1] : nullptr;
    if (p == 2) e.inner = "n" + std::to_string(lid.at(chain[1]));
    e.unit = p == 1;
    e.decode = [&](int ind) {
        const std::string pad(2 * ind, ' ');
        if (p == 1) {
            body << pad << "const int64_t i0 = f_;\n";
        } else if (idx32) {
            body << pad << "uint32_t r_ = (uint32_t)f_;\n";
            for (int64_t r = static_cast<int64_t>(p) - 1; r >= 1; --r)
                body << pad << "const int64_t i" << r << " = (int64_t)(r_ % (uint32_t)n" << lid.at(chain[r])
                     << "); r_ /= (uint32_t)n" << lid.at(chain[r]) << ";\n";
            body << pad << "const int64_t i0 = (int64_t)r_;\n";
        } else {
            body << pad << "int64_t r_ = f_;\n";
            for (int64_t r = static_cast<int64_t>(p) - 1; r >= 1; --r)
                body << pad << "const int64_t i" << r << " = r_ % n" << lid.at(chain[r]) << "; r_ /= n"
                     << lid.at(chain[r]) << ";\n";
            body << pad << "const int64_t i0 = r_;\n";
        }
    };
    e.compute = [&](int ind, const Scope *outer) {
        // Open the scopes of the parallel loops; the innermost one carries the body.
        std::vector<Scope> scopes(p);
        const Scope *parent = outer;
        for (uint64_t r = 0; r < p; ++r) {
            scopes[r].parent = parent;
            body << std::string(2 * ind + 2 * r, ' ') << "{\n";
            declare_scope(*chain[r], scopes[r], ind + 1 + static_cast<int>(r));
            if (r == p - 1) emit_blocks(*chain[r], scopes[r], ind + 1 + static_cast<int>(r));
            parent = &scopes[r];
        }
        for (int64_t r = static_cast<int64_t>(p) - 1; r >= 0; --r) {
            write_back(*chain[r], scopes[r], ind + 1 + static_cast<int>(r));
            body << std::string(2 * ind + 2 * r, ' ') << "}\n";
        }
    };
    parts = cfg.xcd_parts && shifted_reads(l0);
    grid_stride("tot_", flat ? cfg.map_unroll : 1, e);
    parts = false;
    write_back(unit, ksc, 1);
    kp.grid = std::max<uint64_t>(1, std::min<uint64_t>((total + kp.block - 1) / kp.block, cfg.max_grid));
}

// A rank-0 reduction loop whose body is one innermost rank-1 loop that reduces rows into
// temporaries, which rank 0 reduces again with the same operator (max_i(max_j x[i][j])),
// is the full reduction of the 2-D iteration space.  Returns, per temporary, the rank-0
// reduction it feeds, and the instructions the flat skeleton implements itself; empty
// when the loop does not have that form.
std::map<const bh_base *, InstrPtr> collapsible_pairs(const LoopB &l0, const std::set<bh_base *> &temps,
                                                      std::set<const bh_instruction *> &skip) {
    std::map<const bh_base *, InstrPtr> none, pairs, inner;
    const LoopB *l1 = nullptr;
    for (const Block &b : l0._block_list) {
        if (b.isInstr()) continue;
        if (l1 != nullptr) return none;
        l1 = &b.getLoop();
    }
    if (l1 == nullptr || l1->rank != 1 || !l1->isInnermost()) return none;
    for (const InstrPtr &s : l1->_sweeps) {
        if (!bh_opcode_is_reduction(s->opcode) || s->sweep_axis() != 1 || !temps.count(s->operand[0].base)) return none;
        inner[s->operand[0].base] = s;
    }
    if (inner.empty()) return none;
    std::set<const bh_instruction *> sk;
    for (const Block &b : l0._block_list) {
        if (!b.isInstr()) continue;
        const InstrPtr &i = b.getInstr();
        if (bh_opcode_is_system(i->opcode)) continue;
        if (i->opcode == BH_IDENTITY && i->operand[1].isConstant() && inner.count(i->operand[0].base)) {
            sk.insert(i.get());
            continue;
        }
        if (bh_opcode_is_reduction(i->opcode) && i->sweep_axis() == 0 && inner.count(i->operand[1].base) &&
            inner.at(i->operand[1].base)->opcode == i->opcode && !pairs.count(i->operand[1].base)) {
            pairs[i->operand[1].base] = i;
            sk.insert(i.get());
            continue;
        }
        return none;
    }
    if (pairs.size() != inner.size()) return none;
    // The row temporaries may not be read anywhere else.
    for (const InstrPtr &i : jitk::allInstr(l0)) {
        if (sk.count(i.get()) || bh_opcode_is_system(i->opcode)) continue;
        for (size_t o = 1; o < i->operand.size(); ++o)
            if (!i->operand[o].isConstant() && inner.count(i->operand[o].base)) return none;
    }
    skip.insert(sk.begin(), sk.end());
    return pairs;
}

// ---- FULL_REDUCE: rank-0 reductions as two launches on one stream.  The main kernel
// leaves one partial per workgroup and reduction in scratch; the finalize kernel (one
// workgroup) folds the partials in workgroup order, stores the results and then runs
// the small scalar blocks that follow the sweeping loop in the unit (the tail) in
// thread 0.  The kernel boundary orders the two, so there is no fence, atomic ticket
// or counter to reset, the sum order is fixed, and the pair can be graph-captured.
void Writer::emit_full_reduce() {
    // The unit is [identity writers, the sweeping loop, tail blocks]; tail blocks may themselves be
    // small loops with sweeps (a 256-element sum or scan), so the sweeping loop is the first one.
    size_t pos = unit._block_list.size();
    for (size_t k = 0; k < unit._block_list.size() && pos == unit._block_list.size(); ++k)
        if (!unit._block_list[k].isInstr() && !unit._block_list[k].getLoop()._sweeps.empty()) pos = k;
    if (pos == unit._block_list.size()) throw std::logic_error("FULL_REDUCE unit without a sweeping loop");
    const LoopB &l0 = unit._block_list[pos].getLoop();
    std::vector<InstrPtr> sweeps;
    for (const InstrPtr &i : jitk::allLocalInstr(l0))
        if (bh_opcode_is_reduction(i->opcode) && i->sweep_axis() == 0) sweeps.push_back(i);
    std::sort(sweeps.begin(), sweeps.end(), [](const InstrPtr &a, const InstrPtr &b) { return a->origin_id < b->origin_id; });

    std::set<bh_base *> all_temps = temps;
    {
        const std::set<bh_base *> lt = unit.getLocalTemps();
        all_temps.insert(lt.begin(), lt.end());
    }
    const std::map<const bh_base *, InstrPtr> pairs = collapsible_pairs(l0, all_temps, skip);
    const LoopB *l1 = nullptr;
    uint64_t total = static_cast<uint64_t>(l0.size);
    if (!pairs.empty()) {
        for (const Block &b : l0._block_list)
            if (!b.isInstr()) l1 = &b.getLoop();
        total *= static_cast<uint64_t>(l1->size);
    }
    kp.work_items = total;
    // Each thread folds at least reduce_min_items elements before another workgroup (and its
    // partial) is added; at most reduce_max_grid workgroups.
    // A light loop body (a few operations per element, a stencil) streams better from half as many
    // workgroups, each folding twice the elements; a heavy one (Black-Scholes' ~70 operations) from
    // the full number (MI355X: Jacobi 8190^2 stencil+max 5 % faster at 2048 than 4096 workgroups,
    // Black-Scholes 2-4 % slower).
    const uint64_t per_wg = static_cast<uint64_t>(kp.block) * static_cast<uint64_t>(std::max(1, cfg.reduce_min_items));
    const bool light = jitk::allInstr(l0).size() <= kLightReduceInstrs;
    const uint64_t cap = std::max<uint64_t>(1, light ? cfg.reduce_max_grid / 2 : cfg.reduce_max_grid);
    const uint64_t grid = std::max<uint64_t>(1, std::min<uint64_t>((total + per_wg - 1) / per_wg, cap));
    kp.grid = grid;

    // Kernel scope: every rank-0 reduction output is a per-thread accumulator that starts at the identity.
    Scope ksc;
    struct Acc {
        InstrPtr instr;
        std::string var;
        bh_type t;
        bool is_tmp;
    };
    std::vector<Acc> accs;
    const std::set<bh_base *> ktmps = unit.getLocalTemps();
    for (const InstrPtr &s : sweeps) {
        const bh_view &v = s->operand[0];
        const bh_type t = v.base->dtype();
        Acc a;
        a.instr = s;
        a.t = t;
        a.is_tmp = ktmps.count(v.base) > 0 || temps.count(v.base) > 0;
        const std::string ident = literal(jitk::sweep_identity(s->opcode, t));
        if (a.is_tmp) {
            a.var = "t" + std::to_string(bid.at(v.base));
            if (!ksc.tmps.count(v.base)) {
                ksc.tmps.insert(v.base);
                body << "  " << hip_type(t) << " " << a.var << " = " << ident << ";\n";
            }
        } else {
            a.var = "s" + std::to_string(bid.at(v.base)) + "_" + std::to_string(view_id(v));
            if (!ksc.srep.count(vkey(v))) {
                ksc.srep.insert(vkey(v));
                body << "  " << hip_type(t) << " " << a.var << " = " << ident << ";\n";
            }
        }
        accs.push_back(a);
    }
    // Partials: one 16-byte slot per workgroup and distinct accumulator, workgroup-major, so the
    // layout (and the kernel source) does not depend on the grid size.
    std::map<std::string, int> slot;
    for (const Acc &a : accs)
        if (!slot.count(a.var)) slot.emplace(a.var, static_cast<int>(slot.size()));
    const int nslots = static_cast<int>(slot.size());
    kp.scratch_bytes = static_cast<int64_t>(grid) * nslots * 16;
    auto slot_addr = [&](const Acc &a, const std::string &wg) {
        return "(bh_scratch + ((size_t)" + wg + " * " + std::to_string(nslots) + " + " +
               std::to_string(slot.at(a.var)) + ") * 16)";
    };
    declare_scope(unit, ksc, 1);
    if (!pairs.empty()) {
        // One flat grid-stride loop over (i0, i1); each row temporary is the accumulator
        // of the rank-0 reduction it feeds.
        for (const auto &p : pairs)
            for (const Acc &a : accs)
                if (a.instr == p.second) ksc.alias[p.first] = a.var;
        const std::string n0 = "n" + std::to_string(lid.at(&l0)), n1 = "n" + std::to_string(lid.at(l1));
        body << "  const int64_t tot_ = " << n0 << " * " << n1 << ";\n";
        Element e;
        e.outer = &ksc;
        e.body_loop = l1;
        e.inner = n1;
        e.decode = [&](int ind) {
            const std::string pad(2 * ind, ' ');
            if (total < (1ull << 31)) {
                body << pad << "const int64_t i0 = (int64_t)((uint32_t)f_ / (uint32_t)" << n1 << ");\n";
                body << pad << "const int64_t i1 = (int64_t)((uint32_t)f_ - (uint32_t)i0 * (uint32_t)" << n1 << ");\n";
            } else {
                body << pad << "const int64_t i0 = f_ / " << n1 << ";\n" << pad << "const int64_t i1 = f_ - i0 * " << n1
                     << ";\n";
            }
        };
        e.compute = [&](int ind, const Scope *outer) {
            Scope sc0;
            sc0.parent = outer;
            Scope sc;
            sc.parent = &sc0;
            declare_scope(*l1, sc, ind);
            emit_blocks(*l1, sc, ind);
            write_back(*l1, sc, ind);
        };
        parts = cfg.xcd_parts && shifted_reads(l0);
        grid_stride("tot_", cfg.reduce_unroll, e);
        parts = false;
    } else {
        Element e;
        e.outer = &ksc;
        e.body_loop = l0.isInnermost() ? &l0 : nullptr;
        e.unit = true;
        e.decode = [&](int ind) { body << std::string(2 * ind, ' ') << "const int64_t i0 = f_;\n"; };
        e.compute = [&](int ind, const Scope *outer) {
            Scope sc;
            sc.parent = outer;
            declare_scope(l0, sc, ind);
            emit_blocks(l0, sc, ind);
            write_back(l0, sc, ind);
        };
        parts = cfg.xcd_parts && shifted_reads(l0);
        grid_stride("n" + std::to_string(lid.at(&l0)), l0.isInnermost() ? cfg.reduce_unroll : 1, e);
        parts = false;
    }
    body << "  __shared__ __attribute__((aligned(16))) unsigned char bh_lds_[BH_WAVES * 16];\n";
    auto op_lambda = [&](const Acc &a) {
        const std::string T = hip_type(a.t);
        return "auto op_ = [](" + T + " x, " + T + " y) -> " + T + " { return " +
               combine_expr(a.instr->opcode, "x", "y") + "; };";
    };
    std::set<std::string> done;
    for (const Acc &a : accs) {
        if (!done.insert(a.var).second) continue;
        const std::string T = hip_type(a.t);
        body << "  { " << op_lambda(a) << "\n";
        body << "    " << a.var << " = bh_block_reduce(" << a.var << ", op_, (" << T << " *)bh_lds_);\n";
        body << "    if (threadIdx.x == 0) *(" << T << " *)" << slot_addr(a, "bh_bid") << " = " << a.var << ";\n";
        body << "    __syncthreads(); }\n";
    }
    main_body = body.str();
    body.str("");
    body.clear();

    // Finalize kernel: fold the partials, store, run the tail in thread 0.
    Scope fsc;
    done.clear();
    body << "  __shared__ __attribute__((aligned(16))) unsigned char bh_lds_[BH_WAVES * 16];\n";
    for (const Acc &a : accs) {
        if (!done.insert(a.var).second) continue;
        const std::string T = hip_type(a.t);
        const bh_view &v = a.instr->operand[0];
        if (a.is_tmp) {
            fsc.tmps.insert(v.base);
        } else {
            fsc.srep.insert(vkey(v));
        }
        body << "  " << T << " " << a.var << ";\n";
        body << "  { " << op_lambda(a) << "\n";
        body << "    " << T << " v_ = " << literal(jitk::sweep_identity(a.instr->opcode, a.t)) << ";\n";
        // Sixteen partials per thread are loaded before any is folded, so their memory latencies overlap
        // (one load in flight at a time made this one-workgroup kernel latency-bound: up to 4096
        // partials take one round trip); each thread still folds j_ = tid, tid + BH_WG, ... in order.
        body << "    for (unsigned j_ = threadIdx.x; j_ < bh_nparts; j_ += 16u * BH_WG) {\n";
        body << "      " << T << " x_[16];\n";
        body << "#pragma unroll\n";
        body << "      for (unsigned u_ = 0; u_ < 16u; ++u_)\n";
        body << "        if (j_ + u_ * BH_WG < bh_nparts) x_[u_] = *(const " << T << " *)"
             << slot_addr(a, "(j_ + u_ * BH_WG)") << ";\n";
        body << "#pragma unroll\n";
        body << "      for (unsigned u_ = 0; u_ < 16u; ++u_)\n";
        body << "        if (j_ + u_ * BH_WG < bh_nparts) v_ = op_(v_, x_[u_]);\n";
        body << "    }\n";
        body << "    " << a.var << " = bh_block_reduce(v_, op_, (" << T << " *)bh_lds_);\n";
        body << "    __syncthreads(); }\n";
    }
    body << "  if (threadIdx.x != 0) return;\n";
    done.clear();
    for (const Acc &a : accs) {
        if (a.is_tmp || !done.insert(a.var).second) continue;
        const bh_view &v = a.instr->operand[0];
        body << "  a" << bid.at(v.base) << "[" << index(v) << "] = " << a.var << ";\n";
    }
    declare_scope(unit, fsc, 1);
    for (size_t k = pos + 1; k < unit._block_list.size(); ++k) emit_block(unit._block_list[k], fsc, 1);
    write_back(unit, fsc, 1);
    kp.fin_name = "bh_kernel_fin";
}

// ---- SCAN: decoupled look-back over tiles of BH_WG * items elements
void Writer::emit_scan() {
    const LoopB &l0 = unit._block_list.back().getLoop();
    kp.work_items = static_cast<uint64_t>(l0.size);
    InstrPtr acc;
    for (const InstrPtr &i : jitk::allLocalInstr(l0))
        if (bh_opcode_is_accumulate(i->opcode)) acc = i;
    const bh_type t = acc->operand[0].base->dtype();
    const std::string T = hip_type(t);
    const int items = cfg.scan_items;
    const int64_t tile_elems = static_cast<int64_t>(kp.block) * items;
    const int64_t ntiles = (l0.size + tile_elems - 1) / tile_elems;
    kp.grid = static_cast<uint64_t>(std::max<int64_t>(1, ntiles));
    // Scratch layout: [tile counter u64 = epoch << 32 | tiles taken, pad to 256], then one record
    // of 8 bytes per 32-bit chunk of T per tile; no offset depends on the size.
    const int64_t nchunks = std::max<int64_t>(1, bh_type_size(t) / 4);
    kp.scratch_bytes = 256 + ntiles * 8 * nchunks;
    const std::string ident = literal(jitk::sweep_identity(acc->opcode, t));
    scan_instr = acc.get();
    scan_sink = "x_";
    kp.uses_error_flag = true;

    body << "  auto op_ = [](" << T << " x, " << T << " y) -> " << T << " { return "
         << combine_expr(acc->opcode, "x", "y") << "; };\n";
    body << "  unsigned char *bh_rec = bh_scratch + 256;\n";
    body << "  __shared__ unsigned int bh_tile_s, bh_ep_s;\n";
    body << "  __shared__ " << T << " bh_wtot_[BH_WAVES];\n";
    body << "  __shared__ " << T << " bh_excl_s;\n";
    // Tiles are numbered in the order workgroups start, so a tile only waits for tiles whose
    // workgroups already run.  The workgroup taking the last tile starts the next epoch.
    body << "  if (threadIdx.x == 0) {\n";
    body << "    const unsigned long long t_ = atomicAdd((unsigned long long *)bh_scratch, 1ull);\n";
    body << "    bh_tile_s = (unsigned int)t_;\n    bh_ep_s = (unsigned int)(t_ >> 32);\n";
    body << "    if ((unsigned int)t_ == bh_nblk - 1u)\n";
    body << "      __hip_atomic_store((unsigned long long *)bh_scratch, (unsigned long long)(bh_ep_s + 1u) << 32, "
            "__ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);\n";
    body << "  }\n";
    body << "  __syncthreads();\n";
    body << "  const int64_t tile_ = (int64_t)__builtin_amdgcn_readfirstlane(bh_tile_s);\n"
            "  const unsigned int fb_ = bh_ep_s << 2;\n";
    body << "  const int lane_ = threadIdx.x & 63, wave_ = threadIdx.x >> 6;\n";
    body << "  const int64_t tb_ = tile_ * " << tile_elems << ";\n";
    body << "  " << T << " v_[" << items << "];\n";
    Scope ksc;
    declare_scope(unit, ksc, 1);
    // The elements first (every load of the tile issued before the first wave scan waits for one):
    // a full tile without bounds checks, unrolled; the last tile with them.
    for (int full = 1; full >= 0; --full) {
        if (full)
            body << "  if (tb_ + " << tile_elems << " <= n" << lid.at(&l0) << ") {\n#pragma unroll\n";
        else
            body << "  } else {\n";
        body << "  for (int j_ = 0; j_ < " << items << "; ++j_) {\n";
        body << "    const int64_t i0 = tb_ + ((int64_t)wave_ * " << items << " + j_) * 64 + lane_;\n";
        body << "    " << T << " x_ = " << ident << ";\n";
        body << (full ? "    {\n" : "    if (i0 < n" + std::to_string(lid.at(&l0)) + ") {\n");
        {
            Scope sc;
            sc.parent = &ksc;
            declare_scope(l0, sc, 3);
            emit_blocks(l0, sc, 3);
            write_back(l0, sc, 3);
        }
        body << "    }\n    v_[j_] = x_;\n  }\n";
    }
    body << "  }\n";
    body << "  " << T << " carry_ = " << ident << ";\n";
    body << "  for (int j_ = 0; j_ < " << items << "; ++j_) {\n";
    body << "    const " << T << " x_ = op_(carry_, bh_wave_inclusive_scan(v_[j_], op_));\n";
    body << "    carry_ = bh_lane(x_, 63);\n";
    body << "    v_[j_] = x_;\n  }\n";
    body << "  if (lane_ == 63) bh_wtot_[wave_] = carry_;\n  __syncthreads();\n";
    body << "  " << T << " wpre_ = " << ident << ", ttot_ = " << ident << ";\n";
    body << "  for (int w_ = 0; w_ < BH_WAVES; ++w_) { if (w_ < wave_) wpre_ = op_(wpre_, bh_wtot_[w_]); ttot_ = op_(ttot_, bh_wtot_[w_]); }\n";
    // Look-back by wave 0: lane l reads the record of tile (hi - l); the window ends at the
    // nearest inclusive prefix, and the aggregates after it are folded in.
    const std::string LB = "bh_lookback<" + T + ">";
    body << "  if (wave_ == 0) {\n";
    body << "    " << T << " ex_ = " << ident << ";\n";
    body << "    if (tile_ == 0) {\n";
    body << "      if (lane_ == 0) " << LB << "::put(bh_rec, 0, ttot_, fb_ | 2u);\n";
    body << "    } else {\n";
    body << "      if (lane_ == 0) " << LB << "::put(bh_rec, tile_, ttot_, fb_ | 1u);\n";
    body << "      for (int64_t hi_ = tile_ - 1;; hi_ -= 64) {\n";
    body << "        const int64_t j_ = hi_ - lane_;\n";
    body << "        " << T << " w_ = " << ident << ";\n";
    body << "        unsigned int st_ = 2u;\n";
    // A record that never arrives (lost or of another epoch) must not hang the device: after
    // BH_SCAN_SPIN polls the wait gives up, raises BH_ERR_SCAN and the host reports it.
    body << "        if (j_ >= 0) {\n";
    body << "          unsigned int sp_ = 0u;\n";
    body << "          while ((st_ = " << LB << "::get(bh_rec, j_, fb_, &w_)) == 0u) {\n";
    body << "            if (++sp_ >= BH_SCAN_SPIN) { atomicOr(bh_err, BH_ERR_SCAN); st_ = 2u; break; }\n";
    body << "          }\n";
    body << "        }\n";
    body << "        const unsigned long long pm_ = __ballot(st_ == 2u);\n";
    body << "        const int first_ = pm_ ? __ffsll((long long)pm_) - 1 : 64;\n";
    body << "        ex_ = op_(bh_wave_reduce(lane_ <= first_ ? w_ : " << ident << ", op_), ex_);\n";
    body << "        if (pm_) break;\n";
    body << "      }\n";
    body << "      if (lane_ == 0) " << LB << "::put(bh_rec, tile_, op_(ex_, ttot_), fb_ | 2u);\n";
    body << "    }\n";
    body << "    if (lane_ == 0) bh_excl_s = ex_;\n";
    body << "  }\n";
    body << "  __syncthreads();\n";
    body << "  const " << T << " pre_ = op_(bh_excl_s, wpre_);\n";
    const bh_view &out = acc->operand[0];
    body << "  for (int j_ = 0; j_ < " << items << "; ++j_) {\n";
    body << "    const int64_t i0 = tb_ + ((int64_t)wave_ * " << items << " + j_) * 64 + lane_;\n";
    body << "    if (i0 < n" << lid.at(&l0) << ") a" << bid.at(out.base) << "[" << index(out) << "] = op_(pre_, v_[j_]);\n";
    body << "  }\n";
    write_back(unit, ksc, 1);
}

// ---- SERIAL: one thread runs the whole loop nest
void Writer::emit_serial() {
    kp.block = 64;
    kp.grid = 1;
    body << "  if (bh_bid != 0 || threadIdx.x != 0) return;\n";
    emit_body(unit, nullptr, 1);
    kp.work_items = 1;
}

KernelPlan Writer::plan(Strategy s) {
    kp.strategy = s;
    kp.name = "bh_kernel";
    if (s == Strategy::SERIAL) kp.block = 64;
    switch (s) {
        case Strategy::MAP: emit_map(); break;
        case Strategy::ROW_REDUCE: emit_row_reduce(); break;
        case Strategy::FULL_REDUCE: emit_full_reduce(); break;
        case Strategy::SCAN: emit_scan(); break;
        case Strategy::SERIAL: emit_serial(); break;
    }
    std::vector<KernelParam> params = parameters();
    if (s == Strategy::FULL_REDUCE || s == Strategy::SCAN) {
        params.push_back({"unsigned char *", "bh_scratch", true});
        ArgSpec a;
        a.kind = ArgSpec::SCRATCH_PTR;
        kp.args.push_back(a);
    }
    if (s == Strategy::FULL_REDUCE) {
        // The finalize kernel folds this many partials (the main kernel's grid).
        params.push_back({"uint32_t", "bh_nparts", false});
        ArgSpec a;
        a.kind = ArgSpec::GRID_U32;
        a.size = 4;
        a.align = 4;
        kp.args.push_back(a);
    }
    kp.param_decls.clear();
    kp.param_types.clear();
    kp.param_names.clear();
    for (const KernelParam &p : params) {
        kp.param_decls.push_back(p.is_ptr ? p.type + "__restrict__ " + p.name : "const " + p.type + " " + p.name);
        kp.param_types.push_back(p.type);
        kp.param_names.push_back(p.name);
    }
    // The body is a device function of the parameters and of the workgroup's place in its
    // grid; the kernels below (and the batched kernels of batch_source) only call it.
    std::string plist, nlist;
    for (size_t i = 0; i < params.size(); ++i) {
        plist += (i ? ",\n    " : "") + kp.param_decls[i];
        nlist += (i ? ", " : "") + params[i].name;
    }
    std::ostringstream dev;
    const std::string main_src = s == Strategy::FULL_REDUCE ? main_body : body.str();
    // Gathers mark an index outside its base in bh_bad_; the flag is raised once per thread at the end
    // (every early return in a body comes before its first gather).
    const std::string bad0 = kp.uses_gather_flag ? "  bool bh_bad_ = false;\n" : "";
    const std::string bad1 = kp.uses_gather_flag ? "  if (bh_bad_) atomicOr(bh_err, BH_ERR_INDEX);\n" : "";
    dev << "__device__ __forceinline__ void bh_main_(" << plist << (plist.empty() ? "" : ",\n    ")
        << "const unsigned bh_bid, const unsigned bh_nblk) {\n" << bad0 << main_src << bad1 << "}\n";
    if (s == Strategy::FULL_REDUCE)
        dev << "\n__device__ __forceinline__ void bh_fin_(" << plist << ") {\n" << bad0 << body.str() << bad1 << "}\n";
    kp.device_code = dev.str();
    std::ostringstream src;
    // Occupancy hint for the main kernel (waves per SIMD the register allocator must allow).
    kp.attributes = cfg.waves_per_eu > 0 && s != Strategy::SERIAL
                        ? " __attribute__((amdgpu_waves_per_eu(" + std::to_string(cfg.waves_per_eu) + ")))"
                        : "";
    // With xcd_remap the grid-stride skeletons number their workgroups per XCD (bh_xcd_block); the
    // scan skeleton takes its tiles from a ticket counter and the serial one has one workgroup.
    const bool remap = cfg.xcd_remap && (s == Strategy::MAP || s == Strategy::FULL_REDUCE || s == Strategy::ROW_REDUCE);
    src << kp.device_code << "\nextern \"C\" __global__ void __launch_bounds__(" << kp.block << ")" << kp.attributes
        << " " << kp.name << "("
        << plist << ") {\n  bh_main_(" << nlist << (nlist.empty() ? "" : ", ")
        << (remap ? "bh_xcd_block(blockIdx.x, gridDim.x)" : "blockIdx.x") << ", gridDim.x);\n}\n";
    if (s == Strategy::FULL_REDUCE)
        src << "\nextern \"C\" __global__ void __launch_bounds__(" << kp.block << ") " << kp.fin_name << "(" << plist
            << ") {\n  bh_fin_(" << nlist << ");\n}\n";
    kp.source = src.str();
    return kp;
}

bool scan_compatible(const LoopB &l0, const InstrPtr &acc) {
    if (!(acc->opcode == BH_ADD_ACCUMULATE || acc->opcode == BH_MULTIPLY_ACCUMULATE)) return false;
    if (!l0.isInnermost()) return false;
    const bh_base *out = acc->operand[0].base;
    for (const InstrPtr &i : jitk::allInstr(l0)) {
        if (i == acc) continue;
        for (const bh_view &v : i->operand)
            if (!v.isConstant() && v.base == out) return false;
    }
    for (const bh_view &v : acc->operand)
        if (&v != &acc->operand[0] && !v.isConstant() && v.base == out) return false;
    return true;
}

} // namespace

Strategy classify_root_block(const LoopB &l0) {
    if (l0._sweeps.empty()) return Strategy::MAP;
    const auto local = jitk::allLocalInstr(l0);
    auto is_local = [&](const InstrPtr &s) { return std::find(local.begin(), local.end(), s) != local.end(); };
    bool all_red_local = true;
    for (const InstrPtr &s : l0._sweeps)
        if (!bh_opcode_is_reduction(s->opcode) || !is_local(s)) all_red_local = false;
    if (all_red_local) return Strategy::FULL_REDUCE;
    if (l0._sweeps.size() == 1) {
        const InstrPtr &s = *l0._sweeps.begin();
        if (bh_opcode_is_accumulate(s->opcode) && is_local(s) && scan_compatible(l0, s)) return Strategy::SCAN;
    }
    return Strategy::SERIAL;
}

namespace {
// Sweep axis per input rank of every sweep in `l` (nested loops included).
void sweep_axes(const LoopB &l, std::map<int64_t, std::set<int> > &out) {
    for (const InstrPtr &i : jitk::allInstr(l))
        if (bh_opcode_is_sweep(i->opcode) && i->operand.size() > 1 && !i->operand[1].isConstant())
            out[i->operand[1].ndim].insert(i->sweep_axis());
}
} // namespace

bool root_merge_allowed(const LoopB &l1, const LoopB &l2) {
    // Sweeps along two axes of inputs of one rank (a column sum and a row sum of one shape) cannot
    // both be inner reductions of one loop nest, at any depth of the nest; the merged block would
    // run on one thread.  (A full reduction through a row temporary reduces inputs of two ranks.)
    std::map<int64_t, std::set<int> > axes;
    sweep_axes(l1, axes);
    sweep_axes(l2, axes);
    for (const auto &kv : axes)
        if (kv.first > 1 && kv.second.size() > 1) return false;
    if (l1.rank != 0 || l2.rank != 0) return true;
    size_t nsweeps = l1._sweeps.size() + l2._sweeps.size();
    bool acc = false;
    for (const InstrPtr &s : l1._sweeps) acc |= bh_opcode_is_accumulate(s->opcode) != 0;
    for (const InstrPtr &s : l2._sweeps) acc |= bh_opcode_is_accumulate(s->opcode) != 0;
    if (acc && nsweeps > 1) return false;
    if (acc && (!l1.isInnermost() || !l2.isInnermost())) return false;
    if (acc) {
        // Work that shares no array with the scan gains nothing from its loop and would wait on its
        // tiles' look-back: it keeps a map of its own (measured on 4096 uneven arrays: a gather beside
        // an independent cumsum took 2.12 ms fused, 0.82 + 1.18 ms apart).
        const std::set<const bh_base *> b1 = l1.getAllBases();
        bool shared = false;
        for (const bh_base *b : l2.getAllBases()) shared |= b1.count(b) != 0;
        if (!shared) return false;
    }
    return true;
}

namespace {
// A block that only writes sweep identities (BH_IDENTITY of a constant).
bool identity_only(const LoopB &l) {
    const auto all = jitk::allInstr(l);
    if (all.empty()) return false;
    for (const InstrPtr &i : all)
        if (i->opcode != BH_IDENTITY || !i->operand[1].isConstant()) return false;
    return true;
}

// True when every identity in `ids` initialises the output of a sweep of `l0`.
bool identities_of(const std::vector<Block> &ids, const LoopB &l0) {
    std::set<const bh_base *> outs;
    for (const InstrPtr &s : l0._sweeps) outs.insert(s->operand[0].base);
    for (const Block &b : ids)
        for (const InstrPtr &i : jitk::allInstr(b))
            if (!outs.count(i->operand[0].base)) return false;
    return true;
}
} // namespace

namespace {
// Iterations one thread runs for a block: the product of its nested loop sizes.
uint64_t serial_cost(const Block &b) {
    if (b.isInstr()) return 1;
    const LoopB &l = b.getLoop();
    uint64_t inner = 1;
    for (const Block &c : l._block_list) inner = std::max(inner, serial_cost(c));
    return static_cast<uint64_t>(std::max<int64_t>(l.size, 0)) * inner;
}

// Blocks this small run in thread 0 of a FULL_REDUCE finalize kernel instead of a launch of their own.
constexpr uint64_t kTailLimit = 256;
bool tail_block(const Block &b) { return serial_cost(b) <= kTailLimit; }

bool is_full_reduce_loop(const Block &b) {
    return !b.isInstr() && !b.getLoop()._sweeps.empty() && classify_root_block(b.getLoop()) == Strategy::FULL_REDUCE;
}

// True when the kernel's last sizeable block is a FULL_REDUCE loop (only tail blocks after it).
bool open_tail(const LoopB &kernel) {
    for (auto it = kernel._block_list.rbegin(); it != kernel._block_list.rend(); ++it) {
        if (is_full_reduce_loop(*it)) return true;
        if (!(it->isSystemOnly() || tail_block(*it))) return false;
    }
    return false;
}
} // namespace

namespace {
// Bases a kernel reads and writes (a BH_FREE counts as a write).
struct Access {
    std::set<const bh_base *> reads, writes;
    explicit Access(const LoopB &k) {
        for (const InstrPtr &i : jitk::allInstr(k)) {
            for (size_t o = 0; o < i->operand.size(); ++o) {
                const bh_view &v = i->operand[o];
                if (v.isConstant()) continue;
                if (o == 0 || i->opcode == BH_FREE) {
                    writes.insert(v.base);
                } else {
                    reads.insert(v.base);
                }
            }
        }
    }
    // True when running `this` before `m` instead of after it could change a result.
    bool conflicts(const Access &m) const {
        for (const bh_base *b : m.writes)
            if (reads.count(b) || writes.count(b)) return true;
        for (const bh_base *b : writes)
            if (m.reads.count(b)) return true;
        return false;
    }
};
} // namespace

std::vector<LoopB> fold_serial_tails(std::vector<LoopB> kernels) {
    std::vector<LoopB> out;
    for (LoopB &k : kernels) {
        bool small = !k._block_list.empty();
        for (const Block &b : k._block_list) small &= b.isSystemOnly() || tail_block(b);
        size_t target = out.size();
        if (small) {
            // The nearest earlier kernel with an open tail, if `k` can move before the kernels after it.
            const Access ak(k);
            for (size_t m = out.size(); m-- > 0;) {
                if (open_tail(out[m])) {
                    target = m;
                    break;
                }
                if (ak.conflicts(Access(out[m]))) break;
            }
        }
        if (target < out.size()) {
            for (Block &b : k._block_list) out[target]._block_list.push_back(std::move(b));
            out[target].metadataUpdate();
        } else {
            out.push_back(std::move(k));
        }
    }
    return out;
}

namespace {
// Every operand dimension of every instruction must be iterated by the enclosing loop of the
// matching rank, and that loop may not run past the dimension (the generated index is
// start + sum(i_rank * stride), in bounds only then).  `sizes` holds the enclosing loop sizes
// by rank.
void verify_loop_ranks(const LoopB &l, std::vector<int64_t> &sizes, std::string &err) {
    if (l.rank >= 0) {
        if (static_cast<int64_t>(sizes.size()) != l.rank) {
            err += "loop rank " + std::to_string(l.rank) + " nested at depth " + std::to_string(sizes.size()) + "; ";
            return;
        }
        sizes.push_back(l.size);
    }
    for (const Block &b : l._block_list) {
        if (!b.isInstr()) {
            verify_loop_ranks(b.getLoop(), sizes, err);
            continue;
        }
        const InstrPtr &i = b.getInstr();
        if (bh_opcode_is_system(i->opcode)) continue;
        for (size_t o = 0; o < i->operand.size(); ++o) {
            const bh_view &v = i->operand[o];
            if (v.isConstant()) continue;
            if (i->opcode == BH_GATHER && o == 1) continue;            // indexed by value, checked at run time
            if ((i->opcode == BH_SCATTER || i->opcode == BH_COND_SCATTER) && o == 0) continue;
            int hidden = BH_MAXDIM;
            if (o == 0 && bh_opcode_is_reduction(i->opcode) && i->operand[1].ndim > 1) hidden = i->sweep_axis();
            if (o == 0 && bh_opcode_is_reduction(i->opcode) && i->operand[1].ndim <= 1) continue;  // scalar output
            for (int64_t d = 0; d < v.ndim; ++d) {
                const int64_t t = d >= hidden ? d + 1 : d;
                // an extent-1 axis only ever takes index 0 unless a loop of several iterations walks it
                if (v.shape[d] <= 1 && (v.stride[d] == 0 || t >= static_cast<int64_t>(sizes.size()) || sizes[t] <= 1))
                    continue;
                if (t >= static_cast<int64_t>(sizes.size()) || sizes[t] > v.shape[d]) {
                    err += std::string(bh_opcode_text(i->opcode)) + " operand " + std::to_string(o) + " dim " +
                           std::to_string(d) + " (extent " + std::to_string(v.shape[d]) + ") iterated by rank " +
                           std::to_string(t) + (t < static_cast<int64_t>(sizes.size()) ? " of size " + std::to_string(sizes[t]) : " (none)") + "; ";
                }
            }
        }
    }
    if (l.rank >= 0) sizes.pop_back();
}
} // namespace

std::string verify_kernel(const LoopB &kernel) {
    std::vector<int64_t> sizes;
    std::string err;
    verify_loop_ranks(kernel, sizes, err);
    return err;
}

std::vector<KernelPlan> plan_kernel(const LoopB &kernel, const WriterConfig &cfg) {
    std::vector<KernelPlan> ret;
    auto emit = [&](LoopB unit, Strategy s) {
        unit.metadataUpdate();
        Writer w(unit, cfg);
        ret.push_back(w.plan(s));
    };
    // Each top-level loop is one launch.  Identity writers that precede a sweeping
    // loop are folded into it when its skeleton seeds the sweep with the identity
    // itself (FULL_REDUCE, SCAN); otherwise they run first.
    std::vector<Block> pending;
    const std::vector<Block> &blocks = kernel._block_list;
    for (size_t bi = 0; bi < blocks.size(); ++bi) {
        const Block &b = blocks[bi];
        if (b.isInstr()) {
            pending.push_back(b);
            continue;
        }
        if (b.isSystemOnly()) continue;
        if (identity_only(b.getLoop())) {
            pending.push_back(b);
            continue;
        }
        const LoopB &l0 = b.getLoop();
        Strategy s = classify_root_block(l0);
        const bool seeds_itself = s == Strategy::FULL_REDUCE || s == Strategy::SCAN;
        std::vector<Block> scalar_ids;
        if (!(seeds_itself && identities_of(pending, l0))) {
            for (const Block &p : pending) {
                if (p.isInstr()) {
                    scalar_ids.push_back(p);
                } else {
                    emit(LoopB(-1, 1, {p}), Strategy::MAP);
                }
            }
        }
        pending.clear();
        if (s == Strategy::MAP && l0.size == 0) continue;
        // Scalar writers in front of a map run first as their own one-thread launch, so the
        // map keeps its grid instead of running serially behind them.
        if (s == Strategy::MAP && !scalar_ids.empty()) {
            emit(LoopB(-1, 1, scalar_ids), Strategy::SERIAL);
            scalar_ids.clear();
        }
        LoopB unit(-1, 1, scalar_ids);
        unit._block_list.push_back(b);
        if (s == Strategy::MAP && cfg.row_reduce && !row_lane_loops(l0).empty()) s = Strategy::ROW_REDUCE;
        if (s == Strategy::FULL_REDUCE) {
            while (bi + 1 < blocks.size() && (blocks[bi + 1].isSystemOnly() || tail_block(blocks[bi + 1])))
                unit._block_list.push_back(blocks[++bi]);
        }
        emit(unit, s);
    }
    for (const Block &p : pending) {
        if (p.isInstr()) {
            emit(LoopB(-1, 1, {p}), Strategy::SERIAL);
        } else {
            emit(LoopB(-1, 1, {p}), Strategy::MAP);
        }
    }
    return ret;
}

std::vector<size_t> arg_offsets(const std::vector<ArgSpec> &args, size_t *total) {
    std::vector<size_t> off;
    size_t n = 0;
    for (const ArgSpec &a : args) {
        while (n % static_cast<size_t>(a.align)) ++n;
        off.push_back(n);
        n += static_cast<size_t>(a.size);
    }
    while (n % 8) ++n;
    if (total != nullptr) *total = n;
    return off;
}

std::string persist_source(const std::vector<const KernelPlan *> &codes) {
    std::ostringstream s;
    std::ostringstream cases;
    for (size_t c = 0; c < codes.size(); ++c) {
        const KernelPlan &p = *codes[c];
        const std::string tag = "_c" + std::to_string(c) + "_";
        std::string code = p.device_code;
        for (const char *fn : {"bh_main_", "bh_fin_"}) {
            const std::string from = std::string("void ") + fn + "(";
            const std::string to = std::string("void ") + fn + tag.substr(1) + "(";
            for (size_t pos = code.find(from); pos != std::string::npos; pos = code.find(from, pos + to.size()))
                code.replace(pos, from.size(), to);
        }
        s << code << "\n";
        const std::vector<size_t> off = arg_offsets(p.args, nullptr);
        std::string unpack;
        for (size_t i = 0; i < p.param_types.size(); ++i) {
            const std::string &t = p.param_types[i];
            const bool ptr = !t.empty() && t.back() == '*';
            unpack += std::string(i ? ", " : "") + "*(" + t + (ptr ? "const *" : " const *") + ")(A_ + " +
                      std::to_string(off[i]) + ")";
        }
        const std::string sep = unpack.empty() ? "" : ", ";
        cases << "        case " << 2 * c << "u: bh_main" << tag << "(" << unpack << sep << "b_, nb_); break;\n";
        if (!p.fin_name.empty()) cases << "        case " << 2 * c + 1 << "u: bh_fin" << tag << "(" << unpack << "); break;\n";
    }
    // Workgroup w of the cooperative grid runs tasks w, w + grid, ... of each level, the levels in
    // order with a grid barrier between them.  A task is (case, workgroup number within its launch,
    // that launch's workgroup count, arguments offset / 16); case 2c runs code c's main body, 2c + 1
    // its finalize; any other case (padding) nothing.  All control flow is wave-uniform.
    s << "\nextern \"C\" __global__ void __launch_bounds__(256) bh_kernel_persist("
      << "const unsigned char *__restrict__ bh_args, const uint4 *__restrict__ bh_tasks, "
         "const unsigned *__restrict__ bh_levels, const unsigned bh_nlevels, unsigned *__restrict__ bh_bar, "
         "unsigned *__restrict__ bh_gerr, const unsigned bh_spin) {\n"
      << "  for (unsigned L_ = 0; L_ < bh_nlevels; ++L_) {\n"
      << "    const unsigned t1_ = __builtin_amdgcn_readfirstlane(bh_levels[L_ + 1u]);\n"
      << "    for (unsigned t_ = __builtin_amdgcn_readfirstlane(bh_levels[L_]) + blockIdx.x; t_ < t1_; t_ += gridDim.x) {\n"
      << "      const uint4 k_ = bh_tasks[t_];\n"
      << "      const unsigned c_ = __builtin_amdgcn_readfirstlane(k_.x);\n"
      << "      const unsigned b_ = __builtin_amdgcn_readfirstlane(k_.y);\n"
      << "      const unsigned nb_ = __builtin_amdgcn_readfirstlane(k_.z);\n"
      << "      const unsigned char *A_ = bh_args + (size_t)__builtin_amdgcn_readfirstlane(k_.w) * 16u;\n"
      << "      switch (c_) {\n" << cases.str() << "        default: break;\n      }\n"
      << "      __syncthreads();\n"
      << "    }\n"
      << "    if (L_ + 1u < bh_nlevels) bh_grid_sync(bh_bar, (L_ + 1u) * gridDim.x, bh_gerr, bh_spin);\n"
      << "  }\n"
      << "  bh_grid_exit(bh_bar);\n"
      << "}\n";
    return s.str();
}

std::string batch_source(const KernelPlan &p) {
    const std::vector<size_t> off = arg_offsets(p.args, nullptr);
    std::string unpack;
    for (size_t i = 0; i < p.param_types.size(); ++i) {
        const std::string &t = p.param_types[i];
        const bool ptr = !t.empty() && t.back() == '*';
        unpack += std::string(i ? ",\n      " : "") + "*(" + t + (ptr ? "const *" : " const *") + ")(A_ + " +
                  std::to_string(off[i]) + ")";
    }
    const std::string sep = unpack.empty() ? "" : ", ";
    std::ostringstream s;
    s << p.device_code;
    if (p.strategy == Strategy::SCAN) {
        // A scan tile takes its tile number from its item's ticket counter, so any workgroup may run
        // any item: bh_first holds the item of each workgroup (the order is the host's; all ones: no
        // item, the workgroup ends at once), then from position bh_items (the workgroup count) the tile
        // count of each item.
        s << "\nextern \"C\" __global__ void __launch_bounds__(" << p.block << ")" << p.attributes
          << " bh_kernel_batch("
          << "const unsigned char *__restrict__ bh_args, const unsigned *__restrict__ bh_first, const unsigned bh_items, "
             "const unsigned bh_stride) {\n"
          << "  const unsigned it_ = bh_first[blockIdx.x];\n"
          << "  if (it_ == 0xffffffffu) return;\n"
          << "  const unsigned char *A_ = bh_args + (size_t)it_ * bh_stride;\n"
          << "  bh_main_(" << unpack << sep << "blockIdx.x, bh_first[bh_items + it_]);\n}\n";
        return s.str();
    }
    // Workgroup b of the batched grid runs workgroup b - first[i] of item i, where
    // first[i] <= b < first[i + 1]; each item's packed arguments sit at args + i * stride.
    s << "\nextern \"C\" __global__ void __launch_bounds__(" << p.block << ")" << p.attributes << " bh_kernel_batch("
      << "const unsigned char *__restrict__ bh_args, const unsigned *__restrict__ bh_first, const unsigned bh_items, "
         "const unsigned bh_stride) {\n"
      << "  const unsigned b_ = blockIdx.x;\n"
      << "  unsigned lo_ = 0, hi_ = bh_items;\n"
      << "  while (hi_ - lo_ > 1u) { const unsigned mid_ = (lo_ + hi_) >> 1; if (bh_first[mid_] <= b_) lo_ = mid_; "
         "else hi_ = mid_; }\n"
      << "  const unsigned char *A_ = bh_args + (size_t)lo_ * bh_stride;\n"
      << "  bh_main_(" << unpack << sep << "b_ - bh_first[lo_], bh_first[lo_ + 1] - bh_first[lo_]);\n}\n";
    if (!p.fin_name.empty()) {
        s << "\nextern \"C\" __global__ void __launch_bounds__(" << p.block << ") bh_kernel_fin_batch("
          << "const unsigned char *__restrict__ bh_args, const unsigned bh_stride) {\n"
          << "  const unsigned char *A_ = bh_args + (size_t)blockIdx.x * bh_stride;\n"
          << "  bh_fin_(" << unpack << ");\n}\n";
    }
    return s.str();
}

} // namespace vehip
} // namespace bohrium
