#include "rewrite.hpp"

#include <map>
#include <set>

#include <bohrium/bh_opcode.h>

namespace bohrium {
namespace vehip {

namespace {

int64_t view_nelem(const bh_view &v) {
    int64_t n = 1;
    for (int64_t d = 0; d < v.ndim; ++d) n *= v.shape[d];
    return n;
}

// A view that covers its whole base in C order.
bool full_contiguous(const bh_view &v) {
    return !v.isConstant() && !v.hasSlide() && v.start == 0 && view_nelem(v) == v.base->nelem() && v.isContiguous();
}

bh_view contiguous_view(bh_base *b, const std::vector<int64_t> &shape) {
    std::vector<int64_t> stride(shape.size());
    int64_t w = 1;
    for (size_t d = shape.size(); d-- > 0;) {
        stride[d] = w;
        w *= shape[d];
    }
    return bh_view(b, 0, static_cast<int64_t>(shape.size()), BhIntVec(shape), BhIntVec(stride));
}

bh_instruction reduce_instr(bh_opcode op, const bh_view &out, const bh_view &in, int64_t axis) {
    bh_instruction r(op, {out, in, bh_view()});
    r.constant.type = bh_type::INT64;
    r.constant.value.int64 = axis;
    return r;
}

} // namespace

int unflatten_full_reductions(std::vector<bh_instruction *> &list, RewriteStore &store) {
    int count = 0;
    for (bool changed = true; changed;) {
        changed = false;
        std::set<const bh_base *> freed;
        for (const bh_instruction *i : list)
            if (i->opcode == BH_FREE) freed.insert(i->operand[0].base);
        // Shape of the latest N-D write that covers each base.
        std::map<const bh_base *, std::vector<int64_t> > nd_shape;
        for (size_t p = 0; p < list.size() && !changed; ++p) {
            bh_instruction *i = list[p];
            if (i->opcode == BH_FREE || i->operand.empty() || i->operand[0].isConstant()) continue;
            if (bh_opcode_is_reduction(i->opcode) && i->operand.size() == 3) {
                const bh_view &in = i->operand[1];
                auto it = nd_shape.find(in.base);
                if (in.ndim == 1 && full_contiguous(in) && it != nd_shape.end() && freed.count(in.base) &&
                    view_nelem(i->operand[0]) == 1) {
                    const std::vector<int64_t> shape = it->second;
                    const std::vector<int64_t> lead(shape.begin(), shape.end() - 1);
                    int64_t nlead = 1;
                    for (int64_t s : lead) nlead *= s;
                    store.bases.emplace_back(new bh_base(nlead, i->operand[0].base->dtype()));
                    bh_base *r = store.bases.back().get();
                    store.instrs.push_back(reduce_instr(i->opcode, contiguous_view(r, lead),
                                                        contiguous_view(in.base, shape),
                                                        static_cast<int64_t>(shape.size()) - 1));
                    bh_instruction *r1 = &store.instrs.back();
                    store.instrs.push_back(reduce_instr(i->opcode, i->operand[0], contiguous_view(r, {nlead}), 0));
                    bh_instruction *r2 = &store.instrs.back();
                    store.instrs.push_back(bh_instruction(BH_FREE, {bh_view(r)}));
                    bh_instruction *fr = &store.instrs.back();
                    list[p] = r1;
                    list.insert(list.begin() + static_cast<std::ptrdiff_t>(p) + 1, {r2, fr});
                    changed = true;
                    ++count;
                    continue;
                }
            }
            const bh_view &out = i->operand[0];
            if (out.ndim >= 2 && full_contiguous(out)) {
                nd_shape[out.base] = std::vector<int64_t>(out.shape.begin(), out.shape.end());
            } else {
                nd_shape.erase(out.base);
            }
        }
    }
    return count;
}

} // namespace vehip
} // namespace bohrium
