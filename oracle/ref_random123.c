/* ORACLE (test infrastructure only).  Builds the reference's own random123()
 * (include/bohrium/jitk/kernel_dependencies/random123_openmp.h:28-37, which
 * includes the vendored thirdparty/Random123-1.09 philox.h) straight from the
 * reference tree and exports it under a stable name for the checker. */
#include <stdint.h>
#include "kernel_dependencies/random123_openmp.h"

uint64_t ref_random123(uint64_t start, uint64_t key, uint64_t index) { return random123(start, key, index); }

void ref_random123_fill(uint64_t start, uint64_t key, uint64_t n, uint64_t *out) {
    for (uint64_t i = 0; i < n; ++i) out[i] = random123(start, key, i);
}
