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
    // Only a full write of an N-D view can be restated: none, nothing to do (one quick pass).
    bool nd = false;
    for (const bh_instruction *i : list)
        nd |= i->opcode != BH_FREE && !i->operand.empty() && !i->operand[0].isConstant() && i->operand[0].ndim >= 2;
    if (!nd) return 0;
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

namespace bohrium {
namespace vehip {

namespace {

bool touches(const bh_instruction *i, const bh_base *b) {
    for (const bh_view &v : i->operand)
        if (!v.isConstant() && v.base == b) return true;
    return false;
}

// Every computing instruction of the flushes considered writes only its operand 0 (flushes with
// extension methods are left alone).
bool writes(const bh_instruction *i, const bh_base *b) {
    return i->opcode != BH_FREE && !i->operand.empty() && !i->operand[0].isConstant() && i->operand[0].base == b;
}

bool same_shape(const bh_view &a, const bh_view &b) {
    if (a.ndim != b.ndim) return false;
    for (int64_t d = 0; d < a.ndim; ++d)
        if (a.shape[d] != b.shape[d]) return false;
    return true;
}

constexpr size_t kNone = static_cast<size_t>(-1);

} // namespace

std::vector<TwinUpdate> rewrite_in_place_updates(std::vector<bh_instruction *> &list, RewriteStore &store,
                                                 const std::set<bh_base *> &syncs,
                                                 const std::function<bh_base *(bh_base *)> &twin_for,
                                                 const std::function<bool(bh_base *)> &in_step) {
    std::vector<TwinUpdate> out;
    for (const bh_instruction *i : list) {
        if (i->opcode > BH_MAX_OPCODE_ID) return out;
        for (const bh_view &v : i->operand)
            if (!v.isConstant() && v.hasSlide()) return out;
    }
    std::set<const bh_base *> done;
    for (bool changed = true; changed;) {
        changed = false;
        for (size_t q = 0; q < list.size() && !changed; ++q) {
            const bh_instruction *cp = list[q];
            if (cp->opcode != BH_IDENTITY || cp->operand.size() != 2) continue;
            const bh_view &V = cp->operand[0], &in = cp->operand[1];
            if (V.isConstant() || in.isConstant()) continue;
            bh_base *G = V.base, *T = in.base;
            if (G == T || done.count(G) || done.count(T) || G->dtype() != T->dtype() || syncs.count(T)) continue;
            if (!full_contiguous(in) || !same_shape(in, V)) continue;
            size_t free_t = kNone;
            bool free_g = false;
            for (size_t k = 0; k < list.size(); ++k) {
                if (list[k]->opcode != BH_FREE) continue;
                if (list[k]->operand[0].base == T && k > q) free_t = k;
                if (list[k]->operand[0].base == G) free_g = true;
            }
            if (free_t == kNone || free_g) continue;
            size_t p0 = kNone;
            for (size_t k = 0; k < q && p0 == kNone; ++k)
                if (touches(list[k], T)) p0 = k;
            if (p0 == kNone || !writes(list[p0], T)) continue;
            bool ok = true;
            for (size_t o = 1; o < list[p0]->operand.size(); ++o) {
                const bh_view &v = list[p0]->operand[o];
                if (!v.isConstant() && v.base == T) ok = false;  // T's first write does not read T
            }
            for (size_t k = p0; k <= q && ok; ++k)
                for (const bh_view &v : list[k]->operand)
                    if (!v.isConstant() && v.base == T && !(full_contiguous(v) && same_shape(v, V))) ok = false;
            for (size_t k = q + 1; k < list.size() && ok; ++k)
                if (k != free_t && touches(list[k], T)) ok = false;
            for (size_t k = p0; k < q && ok; ++k)
                if (writes(list[k], G)) ok = false;
            if (!ok) continue;
            bh_base *twin = twin_for(G);
            if (twin == nullptr) continue;
            TwinUpdate u;
            u.target = G;
            u.twin = twin;
            u.temp = T;
            bool written_before = false;
            for (size_t k = 0; k < p0; ++k) written_before |= writes(list[k], G);
            u.full_copy = written_before || !in_step(G);
            u.in_step_after = true;
            for (size_t k = q + 1; k < list.size(); ++k) u.in_step_after &= !writes(list[k], G);

            const bh_view at_c(twin, V.start, V.ndim, V.shape, V.stride);
            std::vector<bh_instruction *> nl;
            nl.reserve(list.size() + 1);
            for (size_t k = 0; k < list.size(); ++k) {
                if (k == p0 && u.full_copy) {
                    store.instrs.push_back(bh_instruction(BH_IDENTITY, {bh_view(twin), bh_view(G)}));
                    nl.push_back(&store.instrs.back());
                }
                if (k == q || k == free_t) continue;
                bh_instruction *i = list[k];
                const bool rename = (k >= p0 && k < q && touches(i, T)) || (k > q && touches(i, G));
                if (!rename) {
                    nl.push_back(i);
                    continue;
                }
                store.instrs.push_back(*i);
                bh_instruction *c = &store.instrs.back();
                for (bh_view &v : c->operand) {
                    if (v.isConstant()) continue;
                    if (k < q && v.base == T) v = at_c;
                    else if (k > q && v.base == G) v.base = twin;
                }
                nl.push_back(c);
            }
            list.swap(nl);
            done.insert(G);
            done.insert(twin);
            out.push_back(u);
            changed = true;
        }
    }
    return out;
}

} // namespace vehip
} // namespace bohrium
