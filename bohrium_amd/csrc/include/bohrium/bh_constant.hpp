/* A typed scalar constant embedded in an instruction.
 *
 * Layout and meaning follow include/bohrium/bh_constant.hpp:25-196 of the
 * reference (a union of every element type plus the type tag) so constants
 * round-trip through the C bridge bit for bit.
 */
#pragma once

#include <complex>
#include <iosfwd>
#include <bohrium/bh_type.hpp>

union bh_constant_value {
    bh_bool bool8;
    bh_int8 int8;
    bh_int16 int16;
    bh_int32 int32;
    bh_int64 int64;
    bh_uint8 uint8;
    bh_uint16 uint16;
    bh_uint32 uint32;
    bh_uint64 uint64;
    bh_float32 float32;
    bh_float64 float64;
    bh_complex64 complex64;
    bh_complex128 complex128;
    bh_r123 r123;
};

class bh_constant {
public:
    bh_constant_value value{};
    bh_type type = bh_type::BOOL;

    bh_constant() { value.r123 = {0, 0}; }

    /* Construct a constant of `type` whose value is `v` converted with a C cast. */
    template<typename T>
    static bh_constant make(T v, bh_type type) {
        bh_constant c;
        c.type = type;
        c.value.r123 = {0, 0};
        switch (type) {
            case bh_type::BOOL: c.value.bool8 = static_cast<bool>(v); break;
            case bh_type::INT8: c.value.int8 = static_cast<int8_t>(v); break;
            case bh_type::INT16: c.value.int16 = static_cast<int16_t>(v); break;
            case bh_type::INT32: c.value.int32 = static_cast<int32_t>(v); break;
            case bh_type::INT64: c.value.int64 = static_cast<int64_t>(v); break;
            case bh_type::UINT8: c.value.uint8 = static_cast<uint8_t>(v); break;
            case bh_type::UINT16: c.value.uint16 = static_cast<uint16_t>(v); break;
            case bh_type::UINT32: c.value.uint32 = static_cast<uint32_t>(v); break;
            case bh_type::UINT64: c.value.uint64 = static_cast<uint64_t>(v); break;
            case bh_type::FLOAT32: c.value.float32 = static_cast<float>(v); break;
            case bh_type::FLOAT64: c.value.float64 = static_cast<double>(v); break;
            case bh_type::COMPLEX64: c.value.complex64 = {static_cast<float>(v), 0.0f}; break;
            case bh_type::COMPLEX128: c.value.complex128 = {static_cast<double>(v), 0.0}; break;
            case bh_type::R123: c.value.r123 = {static_cast<uint64_t>(v), 0}; break;
        }
        return c;
    }

    static bh_constant get_min(bh_type type);
    static bh_constant get_max(bh_type type);

    int64_t get_int64() const;
    uint64_t get_uint64() const;
    double get_double() const;
    void set_double(double v);

    bool operator==(const bh_constant &other) const;
    bool operator!=(const bh_constant &other) const { return !(*this == other); }

    /* Write the constant as a literal of its own type in C/HIP source. */
    void write_literal(std::ostream &out) const;
};

std::ostream &operator<<(std::ostream &out, const bh_constant &c);
