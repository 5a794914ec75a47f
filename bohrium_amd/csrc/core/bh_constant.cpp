#include <bohrium/bh_constant.hpp>

#include <cmath>
#include <cstring>
#include <iomanip>
#include <limits>
#include <ostream>
#include <sstream>

/* Extreme values used as reduction identities.  They reproduce the reference's
 * choices exactly (core/bh_constant.cpp:32-104): signed minimum is `min()+1`,
 * the float "minimum" is numeric_limits::min() (smallest positive normal), and
 * BOOL's minimum is 0 while its maximum is 1. */
bh_constant bh_constant::get_min(bh_type type) {
    bh_constant c;
    c.type = type;
    switch (type) {
        case bh_type::BOOL: c.value.bool8 = 0; break;
        case bh_type::INT8: c.value.int8 = std::numeric_limits<int8_t>::min() + 1; break;
        case bh_type::INT16: c.value.int16 = std::numeric_limits<int16_t>::min() + 1; break;
        case bh_type::INT32: c.value.int32 = std::numeric_limits<int32_t>::min() + 1; break;
        case bh_type::INT64: c.value.int64 = std::numeric_limits<int64_t>::min() + 1; break;
        case bh_type::UINT8: c.value.uint8 = 0; break;
        case bh_type::UINT16: c.value.uint16 = 0; break;
        case bh_type::UINT32: c.value.uint32 = 0; break;
        case bh_type::UINT64: c.value.uint64 = 0; break;
        case bh_type::FLOAT32: c.value.float32 = std::numeric_limits<float>::min(); break;
        case bh_type::FLOAT64: c.value.float64 = std::numeric_limits<double>::min(); break;
        case bh_type::COMPLEX64:
            c.value.complex64 = {std::numeric_limits<float>::min(), std::numeric_limits<float>::min()};
            break;
        case bh_type::COMPLEX128:
            c.value.complex128 = {std::numeric_limits<double>::min(), std::numeric_limits<double>::min()};
            break;
        case bh_type::R123: c.value.r123 = {0, 0}; break;
    }
    return c;
}

bh_constant bh_constant::get_max(bh_type type) {
    bh_constant c;
    c.type = type;
    switch (type) {
        case bh_type::BOOL: c.value.bool8 = 1; break;
        case bh_type::INT8: c.value.int8 = std::numeric_limits<int8_t>::max(); break;
        case bh_type::INT16: c.value.int16 = std::numeric_limits<int16_t>::max(); break;
        case bh_type::INT32: c.value.int32 = std::numeric_limits<int32_t>::max(); break;
        case bh_type::INT64: c.value.int64 = std::numeric_limits<int64_t>::max(); break;
        case bh_type::UINT8: c.value.uint8 = std::numeric_limits<uint8_t>::max(); break;
        case bh_type::UINT16: c.value.uint16 = std::numeric_limits<uint16_t>::max(); break;
        case bh_type::UINT32: c.value.uint32 = std::numeric_limits<uint32_t>::max(); break;
        case bh_type::UINT64: c.value.uint64 = std::numeric_limits<uint64_t>::max(); break;
        case bh_type::FLOAT32: c.value.float32 = std::numeric_limits<float>::max(); break;
        case bh_type::FLOAT64: c.value.float64 = std::numeric_limits<double>::max(); break;
        case bh_type::COMPLEX64:
            c.value.complex64 = {std::numeric_limits<float>::max(), std::numeric_limits<float>::max()};
            break;
        case bh_type::COMPLEX128:
            c.value.complex128 = {std::numeric_limits<double>::max(), std::numeric_limits<double>::max()};
            break;
        case bh_type::R123:
            c.value.r123 = {std::numeric_limits<uint64_t>::max(), std::numeric_limits<uint64_t>::max()};
            break;
    }
    return c;
}

int64_t bh_constant::get_int64() const {
    switch (type) {
        case bh_type::BOOL: return value.bool8;
        case bh_type::INT8: return value.int8;
        case bh_type::INT16: return value.int16;
        case bh_type::INT32: return value.int32;
        case bh_type::INT64: return value.int64;
        case bh_type::UINT8: return value.uint8;
        case bh_type::UINT16: return value.uint16;
        case bh_type::UINT32: return value.uint32;
        case bh_type::UINT64:
            if (value.uint64 < static_cast<uint64_t>(std::numeric_limits<int64_t>::max()))
                return static_cast<int64_t>(value.uint64);
            break;
        case bh_type::FLOAT32: return static_cast<int64_t>(value.float32);
        case bh_type::FLOAT64: return static_cast<int64_t>(value.float64);
        default: break;
    }
    throw std::overflow_error("Constant cannot be converted to int64_t");
}

uint64_t bh_constant::get_uint64() const {
    if (type == bh_type::UINT64) return value.uint64;
    const int64_t v = get_int64();
    if (v < 0) throw std::overflow_error("Constant cannot be converted to uint64_t");
    return static_cast<uint64_t>(v);
}

double bh_constant::get_double() const {
    switch (type) {
        case bh_type::BOOL: return value.bool8;
        case bh_type::INT8: return value.int8;
        case bh_type::INT16: return value.int16;
        case bh_type::INT32: return value.int32;
        case bh_type::INT64: return static_cast<double>(value.int64);
        case bh_type::UINT8: return value.uint8;
        case bh_type::UINT16: return value.uint16;
        case bh_type::UINT32: return value.uint32;
        case bh_type::UINT64: return static_cast<double>(value.uint64);
        case bh_type::FLOAT32: return value.float32;
        case bh_type::FLOAT64: return value.float64;
        default: throw std::runtime_error("bh_constant::get_double(): not a real type");
    }
}

void bh_constant::set_double(double v) { *this = make(v, type); }

bool bh_constant::operator==(const bh_constant &other) const {
    if (type != other.type) return false;
    return std::memcmp(&value, &other.value, bh_type_size(type)) == 0;
}

namespace {
template<typename F>
void write_float(std::ostream &out, F v, const char *suffix) {
    if (std::isnan(v)) {
        out << "NAN";
    } else if (std::isinf(v)) {
        out << (v < 0 ? "(-INFINITY)" : "INFINITY");
    } else {
        // Hex-float literals keep every bit of the constant.
        std::ostringstream ss;
        ss << std::hexfloat << static_cast<double>(v);
        out << "(" << ss.str() << suffix << ")";
    }
}
} // namespace

void bh_constant::write_literal(std::ostream &out) const {
    switch (type) {
        case bh_type::BOOL: out << (value.bool8 ? "true" : "false"); break;
        case bh_type::INT8: out << "((int8_t)" << static_cast<int>(value.int8) << ")"; break;
        case bh_type::INT16: out << "((int16_t)" << value.int16 << ")"; break;
        case bh_type::INT32: out << "((int32_t)" << value.int32 << ")"; break;
        case bh_type::INT64:
            if (value.int64 == std::numeric_limits<int64_t>::min())
                out << "((int64_t)(-9223372036854775807LL - 1))";
            else
                out << "((int64_t)" << value.int64 << "LL)";
            break;
        case bh_type::UINT8: out << "((uint8_t)" << static_cast<unsigned>(value.uint8) << "u)"; break;
        case bh_type::UINT16: out << "((uint16_t)" << value.uint16 << "u)"; break;
        case bh_type::UINT32: out << "((uint32_t)" << value.uint32 << "u)"; break;
        case bh_type::UINT64: out << "((uint64_t)" << value.uint64 << "ull)"; break;
        case bh_type::FLOAT32: write_float(out, value.float32, "f"); break;
        case bh_type::FLOAT64: write_float(out, value.float64, ""); break;
        case bh_type::COMPLEX64:
            out << "make_c64(";
            write_float(out, value.complex64.real, "f");
            out << ", ";
            write_float(out, value.complex64.imag, "f");
            out << ")";
            break;
        case bh_type::COMPLEX128:
            out << "make_c128(";
            write_float(out, value.complex128.real, "");
            out << ", ";
            write_float(out, value.complex128.imag, "");
            out << ")";
            break;
        case bh_type::R123:
            out << "make_r123(" << value.r123.start << "ull, " << value.r123.key << "ull)";
            break;
    }
}

std::ostream &operator<<(std::ostream &out, const bh_constant &c) {
    switch (c.type) {
        case bh_type::BOOL: out << (c.value.bool8 ? "true" : "false"); break;
        case bh_type::INT8: out << static_cast<int>(c.value.int8); break;
        case bh_type::INT16: out << c.value.int16; break;
        case bh_type::INT32: out << c.value.int32; break;
        case bh_type::INT64: out << c.value.int64; break;
        case bh_type::UINT8: out << static_cast<unsigned>(c.value.uint8); break;
        case bh_type::UINT16: out << c.value.uint16; break;
        case bh_type::UINT32: out << c.value.uint32; break;
        case bh_type::UINT64: out << c.value.uint64; break;
        case bh_type::FLOAT32: out << std::setprecision(9) << c.value.float32; break;
        case bh_type::FLOAT64: out << std::setprecision(17) << c.value.float64; break;
        case bh_type::COMPLEX64: out << "(" << c.value.complex64.real << "+" << c.value.complex64.imag << "j)"; break;
        case bh_type::COMPLEX128:
            out << "(" << c.value.complex128.real << "+" << c.value.complex128.imag << "j)";
            break;
        case bh_type::R123: out << "R123(" << c.value.r123.start << ", " << c.value.r123.key << ")"; break;
    }
    return out;
}
