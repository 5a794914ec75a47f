/* ORACLE (test infrastructure only): Random123 philox2x32 with 10 rounds,
 * restated from the published algorithm (Salmon, Moraes, Dror, Shaw, "Parallel
 * random numbers: as easy as 1, 2, 3", SC'11) and the constants of the
 * reference's vendored header (thirdparty/Random123-1.09/include/Random123/
 * philox.h:228-255, 271-283).  random123(start, key, index) follows
 * include/bohrium/jitk/kernel_dependencies/random123_openmp.h:28-37: the
 * 64-bit counter start+index is split little-endian into two 32-bit words and
 * the key is the low 32 bits of `key`. */
#ifndef ORACLE_R123_PHILOX_H
#define ORACLE_R123_PHILOX_H
#include <stdint.h>

static inline uint64_t oracle_random123(uint64_t start, uint64_t key, uint64_t index) {
    uint64_t ctr = start + index;
    uint32_t c0 = (uint32_t)ctr, c1 = (uint32_t)(ctr >> 32), k = (uint32_t)key;
    for (int r = 0; r < 10; ++r) {
        uint64_t prod = (uint64_t)0xD256D193u * (uint64_t)c0;
        uint32_t hi = (uint32_t)(prod >> 32), lo = (uint32_t)prod;
        uint32_t n0 = hi ^ k ^ c1;
        c1 = lo;
        c0 = n0;
        k += 0x9E3779B9u;
    }
    return (uint64_t)c0 | ((uint64_t)c1 << 32);
}
#endif
