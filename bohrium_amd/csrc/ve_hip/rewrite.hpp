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
#include <functional>
#include <set>
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

/* In-place updates through a temporary, as the front ends emit them for `c[...] = f(views of G)`
 * with c a view of G: T = f(...) (T a temporary of c's shape), IDENTITY(c, T), BH_FREE(T).  Run as
 * written, T makes a round trip through device memory and the copy reads and writes c again.
 * The rewrite computes T's elements straight into a second buffer of G's size (the twin, at c's
 * places), removes the copy, and lets every later instruction of the flush use the twin for G;
 * after the flush the engine exchanges the two buffers, so G holds the new values and the twin the
 * old ones.  The twin must equal G outside c; when the engine cannot vouch for that (`in_step`
 * false, or G is written earlier in the flush) a full copy G -> twin is inserted first.
 * Conditions: T is written first by a full write of its whole shape and every access of T up to
 * the copy is that full view, T is not synced and is freed in the flush, the copy converts no type,
 * G is not written between T's first write and the copy, no view slides, G is not freed. */
struct TwinUpdate {
    bh_base *target = nullptr;  // G
    bh_base *twin = nullptr;
    bh_base *temp = nullptr;    // T: no longer used by the flush; its BH_FREE is removed, the engine releases it
    bool full_copy = false;           // a copy G -> twin was inserted
    bool in_step_after = false;       // after the exchange, twin == G outside c (nothing else wrote G)
};
std::vector<TwinUpdate> rewrite_in_place_updates(std::vector<bh_instruction *> &list, RewriteStore &store,
                                                 const std::set<bh_base *> &syncs,
                                                 const std::function<bh_base *(bh_base *)> &twin_for,
                                                 const std::function<bool(bh_base *)> &in_step);

} // namespace vehip
} // namespace bohrium
