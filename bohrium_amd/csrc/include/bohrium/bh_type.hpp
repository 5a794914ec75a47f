/* Element types of the Bohrium bytecode.
 *
 * The enumerator order is the wire/ABI order of the reference type table
 * (core/codegen/types.json ids 0..13, include/bohrium/bh_type.hpp:43-58), so a
 * bytecode stream produced by any bridge decodes to the same types here.
 */
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>

typedef unsigned char bh_bool;
typedef int8_t bh_int8;
typedef int16_t bh_int16;
typedef int32_t bh_int32;
typedef int64_t bh_int64;
typedef uint8_t bh_uint8;
typedef uint16_t bh_uint16;
typedef uint32_t bh_uint32;
typedef uint64_t bh_uint64;
typedef float bh_float32;
typedef double bh_float64;
typedef struct { float real, imag; } bh_complex64;
typedef struct { double real, imag; } bh_complex128;
/* Random123 counter/key pair carried by BH_RANDOM's constant operand. */
typedef struct { bh_uint64 start, key; } bh_r123;

enum class bh_type : int32_t {
    BOOL = 0,
    INT8,
    INT16,
    INT32,
    INT64,
    UINT8,
    UINT16,
    UINT32,
    UINT64,
    FLOAT32,
    FLOAT64,
    COMPLEX64,
    COMPLEX128,
    R123
};

constexpr int BH_NUM_TYPES = 14;

typedef int64_t bh_opcode;

int bh_type_size(bh_type type);
const char *bh_type_text(bh_type type);
bool bh_type_is_integer(bh_type type);
bool bh_type_is_signed_integer(bh_type type);
bool bh_type_is_unsigned_integer(bh_type type);
bool bh_type_is_float(bh_type type);
bool bh_type_is_complex(bh_type type);
uint64_t bh_type_limit_max_integer(bh_type type);
int64_t bh_type_limit_min_integer(bh_type type);
double bh_type_limit_max_float(bh_type type);
double bh_type_limit_min_float(bh_type type);
bh_type bh_type_from_text(const std::string &name);
