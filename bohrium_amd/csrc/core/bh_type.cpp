#include <bohrium/bh_type.hpp>

#include <cstring>
#include <limits>

namespace {
struct TypeInfo {
    const char *text;
    int size;
};
// Sizes and names of the element types, indexed by bh_type (reference core/codegen/types.json).
const TypeInfo kTypes[BH_NUM_TYPES] = {
    {"BH_BOOL", 1},    {"BH_INT8", 1},      {"BH_INT16", 2},      {"BH_INT32", 4},   {"BH_INT64", 8},
    {"BH_UINT8", 1},   {"BH_UINT16", 2},    {"BH_UINT32", 4},     {"BH_UINT64", 8},  {"BH_FLOAT32", 4},
    {"BH_FLOAT64", 8}, {"BH_COMPLEX64", 8}, {"BH_COMPLEX128", 16}, {"BH_R123", 16},
};
} // namespace

int bh_type_size(bh_type type) {
    const int t = static_cast<int>(type);
    if (t < 0 || t >= BH_NUM_TYPES) throw std::runtime_error("bh_type_size(): unknown type");
    return kTypes[t].size;
}

const char *bh_type_text(bh_type type) {
    const int t = static_cast<int>(type);
    if (t < 0 || t >= BH_NUM_TYPES) return "BH_UNKNOWN";
    return kTypes[t].text;
}

bh_type bh_type_from_text(const std::string &name) {
    for (int t = 0; t < BH_NUM_TYPES; ++t) {
        if (name == kTypes[t].text || name == kTypes[t].text + 3) return static_cast<bh_type>(t);
    }
    throw std::runtime_error("bh_type_from_text(): unknown type " + name);
}

bool bh_type_is_integer(bh_type t) {
    return t == bh_type::INT8 || t == bh_type::INT16 || t == bh_type::INT32 || t == bh_type::INT64 ||
           t == bh_type::UINT8 || t == bh_type::UINT16 || t == bh_type::UINT32 || t == bh_type::UINT64;
}

bool bh_type_is_signed_integer(bh_type t) {
    return t == bh_type::INT8 || t == bh_type::INT16 || t == bh_type::INT32 || t == bh_type::INT64;
}

bool bh_type_is_unsigned_integer(bh_type t) {
    return t == bh_type::UINT8 || t == bh_type::UINT16 || t == bh_type::UINT32 || t == bh_type::UINT64;
}

bool bh_type_is_float(bh_type t) { return t == bh_type::FLOAT32 || t == bh_type::FLOAT64; }

bool bh_type_is_complex(bh_type t) { return t == bh_type::COMPLEX64 || t == bh_type::COMPLEX128; }

uint64_t bh_type_limit_max_integer(bh_type t) {
    switch (t) {
        case bh_type::BOOL: return 1;
        case bh_type::INT8: return std::numeric_limits<int8_t>::max();
        case bh_type::INT16: return std::numeric_limits<int16_t>::max();
        case bh_type::INT32: return std::numeric_limits<int32_t>::max();
        case bh_type::INT64: return std::numeric_limits<int64_t>::max();
        case bh_type::UINT8: return std::numeric_limits<uint8_t>::max();
        case bh_type::UINT16: return std::numeric_limits<uint16_t>::max();
        case bh_type::UINT32: return std::numeric_limits<uint32_t>::max();
        case bh_type::UINT64: return std::numeric_limits<uint64_t>::max();
        default: throw std::runtime_error("bh_type_limit_max_integer(): not an integer type");
    }
}

int64_t bh_type_limit_min_integer(bh_type t) {
    switch (t) {
        case bh_type::BOOL: return 1;  // same as the reference (core/bh_type.cpp:123-126)
        case bh_type::INT8: return std::numeric_limits<int8_t>::min();
        case bh_type::INT16: return std::numeric_limits<int16_t>::min();
        case bh_type::INT32: return std::numeric_limits<int32_t>::min();
        case bh_type::INT64: return std::numeric_limits<int64_t>::min();
        case bh_type::UINT8:
        case bh_type::UINT16:
        case bh_type::UINT32:
        case bh_type::UINT64: return 0;
        default: throw std::runtime_error("bh_type_limit_min_integer(): not an integer type");
    }
}

double bh_type_limit_max_float(bh_type t) {
    if (t == bh_type::FLOAT32) return std::numeric_limits<float>::max();
    if (t == bh_type::FLOAT64) return std::numeric_limits<double>::max();
    throw std::runtime_error("bh_type_limit_max_float(): not a float type");
}

double bh_type_limit_min_float(bh_type t) {
    // numeric_limits::min(), i.e. the smallest positive normal value, as in the reference.
    if (t == bh_type::FLOAT32) return std::numeric_limits<float>::min();
    if (t == bh_type::FLOAT64) return std::numeric_limits<double>::min();
    throw std::runtime_error("bh_type_limit_min_float(): not a float type");
}
