/* Opcode metadata: name, operand count and class flags per opcode id.
 * The ids, arities and classes are the reference's (core/codegen/opcodes.json). */
#include <bohrium/bh_opcode.h>

#include <cstring>

namespace {
enum Flags { ELEM = 1, RED = 2, ACC = 4, SYS = 8 };
struct OpInfo {
    int64_t id;
    const char *name;
    int nop;
    int flags;
};
const OpInfo kOps[] = {
    {BH_ADD, "BH_ADD", 3, ELEM},
    {BH_SUBTRACT, "BH_SUBTRACT", 3, ELEM},
    {BH_MULTIPLY, "BH_MULTIPLY", 3, ELEM},
    {BH_DIVIDE, "BH_DIVIDE", 3, ELEM},
    {BH_POWER, "BH_POWER", 3, ELEM},
    {BH_ABSOLUTE, "BH_ABSOLUTE", 2, ELEM},
    {BH_GREATER, "BH_GREATER", 3, ELEM},
    {BH_GREATER_EQUAL, "BH_GREATER_EQUAL", 3, ELEM},
    {BH_LESS, "BH_LESS", 3, ELEM},
    {BH_LESS_EQUAL, "BH_LESS_EQUAL", 3, ELEM},
    {BH_EQUAL, "BH_EQUAL", 3, ELEM},
    {BH_NOT_EQUAL, "BH_NOT_EQUAL", 3, ELEM},
    {BH_LOGICAL_AND, "BH_LOGICAL_AND", 3, ELEM},
    {BH_LOGICAL_OR, "BH_LOGICAL_OR", 3, ELEM},
    {BH_LOGICAL_XOR, "BH_LOGICAL_XOR", 3, ELEM},
    {BH_LOGICAL_NOT, "BH_LOGICAL_NOT", 2, ELEM},
    {BH_MAXIMUM, "BH_MAXIMUM", 3, ELEM},
    {BH_MINIMUM, "BH_MINIMUM", 3, ELEM},
    {BH_BITWISE_AND, "BH_BITWISE_AND", 3, ELEM},
    {BH_BITWISE_OR, "BH_BITWISE_OR", 3, ELEM},
    {BH_BITWISE_XOR, "BH_BITWISE_XOR", 3, ELEM},
    {BH_INVERT, "BH_INVERT", 2, ELEM},
    {BH_LEFT_SHIFT, "BH_LEFT_SHIFT", 3, ELEM},
    {BH_RIGHT_SHIFT, "BH_RIGHT_SHIFT", 3, ELEM},
    {BH_COS, "BH_COS", 2, ELEM},
    {BH_SIN, "BH_SIN", 2, ELEM},
    {BH_TAN, "BH_TAN", 2, ELEM},
    {BH_COSH, "BH_COSH", 2, ELEM},
    {BH_SINH, "BH_SINH", 2, ELEM},
    {BH_TANH, "BH_TANH", 2, ELEM},
    {BH_ARCSIN, "BH_ARCSIN", 2, ELEM},
    {BH_ARCCOS, "BH_ARCCOS", 2, ELEM},
    {BH_ARCTAN, "BH_ARCTAN", 2, ELEM},
    {BH_ARCSINH, "BH_ARCSINH", 2, ELEM},
    {BH_ARCCOSH, "BH_ARCCOSH", 2, ELEM},
    {BH_ARCTANH, "BH_ARCTANH", 2, ELEM},
    {BH_ARCTAN2, "BH_ARCTAN2", 3, ELEM},
    {BH_EXP, "BH_EXP", 2, ELEM},
    {BH_EXP2, "BH_EXP2", 2, ELEM},
    {BH_EXPM1, "BH_EXPM1", 2, ELEM},
    {BH_LOG, "BH_LOG", 2, ELEM},
    {BH_LOG2, "BH_LOG2", 2, ELEM},
    {BH_LOG10, "BH_LOG10", 2, ELEM},
    {BH_LOG1P, "BH_LOG1P", 2, ELEM},
    {BH_SQRT, "BH_SQRT", 2, ELEM},
    {BH_CEIL, "BH_CEIL", 2, ELEM},
    {BH_TRUNC, "BH_TRUNC", 2, ELEM},
    {BH_FLOOR, "BH_FLOOR", 2, ELEM},
    {BH_RINT, "BH_RINT", 2, ELEM},
    {BH_MOD, "BH_MOD", 3, ELEM},
    {BH_ISNAN, "BH_ISNAN", 2, ELEM},
    {BH_ISINF, "BH_ISINF", 2, ELEM},
    {BH_IDENTITY, "BH_IDENTITY", 2, ELEM},
    {BH_FREE, "BH_FREE", 1, SYS},
    {BH_NONE, "BH_NONE", 0, ELEM | SYS},
    {BH_TALLY, "BH_TALLY", 0, SYS},
    {BH_ADD_REDUCE, "BH_ADD_REDUCE", 3, RED},
    {BH_MULTIPLY_REDUCE, "BH_MULTIPLY_REDUCE", 3, RED},
    {BH_MINIMUM_REDUCE, "BH_MINIMUM_REDUCE", 3, RED},
    {BH_MAXIMUM_REDUCE, "BH_MAXIMUM_REDUCE", 3, RED},
    {BH_LOGICAL_AND_REDUCE, "BH_LOGICAL_AND_REDUCE", 3, RED},
    {BH_BITWISE_AND_REDUCE, "BH_BITWISE_AND_REDUCE", 3, RED},
    {BH_LOGICAL_OR_REDUCE, "BH_LOGICAL_OR_REDUCE", 3, RED},
    {BH_BITWISE_OR_REDUCE, "BH_BITWISE_OR_REDUCE", 3, RED},
    {BH_LOGICAL_XOR_REDUCE, "BH_LOGICAL_XOR_REDUCE", 3, RED},
    {BH_BITWISE_XOR_REDUCE, "BH_BITWISE_XOR_REDUCE", 3, RED},
    {BH_RANDOM, "BH_RANDOM", 2, 0},
    {BH_RANGE, "BH_RANGE", 1, 0},
    {BH_REAL, "BH_REAL", 2, ELEM},
    {BH_IMAG, "BH_IMAG", 2, ELEM},
    {BH_ADD_ACCUMULATE, "BH_ADD_ACCUMULATE", 3, ACC},
    {BH_MULTIPLY_ACCUMULATE, "BH_MULTIPLY_ACCUMULATE", 3, ACC},
    {BH_SIGN, "BH_SIGN", 2, ELEM},
    {BH_GATHER, "BH_GATHER", 3, 0},
    {BH_SCATTER, "BH_SCATTER", 3, 0},
    {BH_REMAINDER, "BH_REMAINDER", 3, ELEM},
    {BH_COND_SCATTER, "BH_COND_SCATTER", 4, 0},
    {BH_ISFINITE, "BH_ISFINITE", 2, ELEM},
    {BH_CONJ, "BH_CONJ", 2, ELEM},
};
constexpr int kNumOps = sizeof(kOps) / sizeof(kOps[0]);

const OpInfo *find(int64_t opcode) {
    static const OpInfo *table[BH_MAX_OPCODE_ID + 1] = {nullptr};
    static bool init = false;
    if (!init) {
        for (int i = 0; i < kNumOps; ++i) table[kOps[i].id] = &kOps[i];
        init = true;
    }
    if (opcode < 0 || opcode > BH_MAX_OPCODE_ID) return nullptr;
    return table[opcode];
}
} // namespace

extern "C" {
const char *bh_opcode_text(int64_t opcode) {
    const OpInfo *o = find(opcode);
    return o ? o->name : "BH_UNKNOWN";
}
int bh_opcode_exists(int64_t opcode) { return find(opcode) != nullptr; }
int bh_noperands(int64_t opcode) {
    const OpInfo *o = find(opcode);
    return o ? o->nop : -1;
}
int bh_opcode_is_system(int64_t opcode) {
    const OpInfo *o = find(opcode);
    return o && (o->flags & SYS);
}
int bh_opcode_is_elementwise(int64_t opcode) {
    const OpInfo *o = find(opcode);
    return o && (o->flags & ELEM);
}
int bh_opcode_is_reduction(int64_t opcode) {
    const OpInfo *o = find(opcode);
    return o && (o->flags & RED);
}
int bh_opcode_is_accumulate(int64_t opcode) {
    const OpInfo *o = find(opcode);
    return o && (o->flags & ACC);
}
int bh_opcode_is_sweep(int64_t opcode) { return bh_opcode_is_reduction(opcode) || bh_opcode_is_accumulate(opcode); }
int64_t bh_opcode_from_text(const char *name) {
    for (int i = 0; i < kNumOps; ++i) {
        if (std::strcmp(kOps[i].name, name) == 0 || std::strcmp(kOps[i].name + 3, name) == 0) return kOps[i].id;
    }
    return -1;
}
}
