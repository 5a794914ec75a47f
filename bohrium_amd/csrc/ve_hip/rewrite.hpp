/* Bytecode rewrites the hip engine applies to a flush before fusion.
 *
 * unflatten_full_reductions: bh107 reduces a contiguous N-D array over every
 * axis by flattening it and reducing the 1-D view (bridge/bh107/bh107/ufuncs.py:383-386).
 * When that array is an N-D temporary produced in the same flush (Jacobi's
 * |tmp - c|), the 1-D reduction cannot fuse with its producer and the array is
 * written to and read back from HBM.  The rewrite restates the reduction on the
 * producer's shape as a reduction over the last axis into a new temporary
 * followed by a 1-D reduction of that temporary (the same two-step form bh107
 * emits for non-contiguous arrays, ufuncs.py:388-391), so the fuser can put it in
 * the producer's block.  Max/min/logical/bitwise results are unchanged; sums and
 * products of floats are reassociated (within the reduction-order tolerance).
 */
#pragma once

#include <deque>
#include <memory>
#include <vector>

#include <bohrium/bh_instruction.hpp>

namespace bohrium {
namespace vehip {

/* Storage for the instructions and temporaries a rewrite adds; it must outlive the flush. */
struct RewriteStore {
    std::deque<bh_instruction> instrs;
    std::vector<std::unique_ptr<bh_base> > bases;
    void clear() {
        instrs.clear();
        bases.clear();
    }
};

/* Returns the number of reductions rewritten. */
int unflatten_full_reductions(std::vector<bh_instruction *> &list, RewriteStore &store);

} // namespace vehip
} // namespace bohrium
