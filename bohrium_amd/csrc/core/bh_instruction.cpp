#include <bohrium/bh_instruction.hpp>

#include <numeric>
#include <sstream>
#include <stdexcept>

std::set<bh_base *> bh_instruction::get_bases() const {
    std::set<bh_base *> ret;
    for (const bh_view &v : operand)
        if (!v.isConstant()) ret.insert(v.base);
    return ret;
}

std::vector<const bh_view *> bh_instruction::getViews() const {
    std::vector<const bh_view *> ret;
    for (const bh_view &v : operand)
        if (!v.isConstant()) ret.push_back(&v);
    return ret;
}

std::vector<bh_view *> bh_instruction::getViews() {
    std::vector<bh_view *> ret;
    for (bh_view &v : operand)
        if (!v.isConstant()) ret.push_back(&v);
    return ret;
}

bool bh_instruction::has_constant() const {
    for (const bh_view &v : operand)
        if (v.isConstant()) return true;
    return false;
}

bool bh_instruction::isContiguous() const {
    for (const bh_view *v : getViews())
        if (!v->isContiguous()) return false;
    return true;
}

bool bh_instruction::all_same_shape() const {
    if (operand.empty()) return true;
    const bh_view &first = operand[0];
    for (size_t o = 1; o < operand.size(); ++o)
        if (!operand[o].isConstant() && !bh_view_same_shape(&first, &operand[o])) return false;
    return true;
}

bool bh_instruction::reshapable() const {
    return all_same_shape() && isContiguous() && !bh_opcode_is_sweep(opcode);
}

BhIntVec bh_instruction::shape() const {
    if (bh_opcode_is_sweep(opcode)) return operand[1].shape;
    if (opcode == BH_GATHER || opcode == BH_SCATTER || opcode == BH_COND_SCATTER) return operand[2].shape;
    if (operand.empty()) return BhIntVec();
    return operand[0].shape;
}

int64_t bh_instruction::ndim() const { return static_cast<int64_t>(shape().size()); }

int bh_instruction::sweep_axis() const {
    if (bh_opcode_is_sweep(opcode)) return static_cast<int>(constant.get_int64());
    return BH_MAXDIM;
}

void bh_instruction::reshape(const std::vector<int64_t> &new_shape) {
    if (!reshapable()) throw std::runtime_error("reshape(): instruction not reshapable");
    const int64_t total = std::accumulate(new_shape.begin(), new_shape.end(), int64_t{1}, std::multiplies<int64_t>());
    for (bh_view *v : getViews()) {
        if (v->shape.prod() != total) throw std::runtime_error("reshape(): shape mismatch");
        v->ndim = static_cast<int64_t>(new_shape.size());
        v->shape = BhIntVec(new_shape.begin(), new_shape.end());
        v->stride.assign(new_shape.size(), 0);
        int64_t s = 1;
        for (int64_t i = v->ndim - 1; i >= 0; --i) {
            v->stride[i] = s;
            s *= v->shape[i];
        }
    }
}

// Same axis bookkeeping as the reference (core/bh_instruction.cpp:134-168): inputs drop
// `axis`, gather's data operand and scatter's output keep their own layout, and a
// reduction's output drops the matching axis of the reduced shape.
void bh_instruction::remove_axis(int64_t axis) {
    if (operand.empty()) return;
    for (size_t o = 1; o < operand.size(); ++o) {
        if (!(operand[o].isConstant() || (o == 1 && opcode == BH_GATHER))) operand[o].remove_axis(axis);
    }
    const int sa = sweep_axis();
    if (sa == axis) throw std::runtime_error("remove_axis(): cannot remove a swept axis");
    if (sa > axis && sa < BH_MAXDIM) constant = bh_constant::make(sa - 1, constant.type);
    if (opcode == BH_SCATTER || opcode == BH_COND_SCATTER) return;
    bh_view &out = operand[0];
    if (bh_opcode_is_reduction(opcode)) {
        out.remove_axis(sa < axis ? axis - 1 : axis);
    } else {
        out.remove_axis(axis);
    }
}

// Same axis bookkeeping as the reference (core/bh_instruction.cpp:170-217).
void bh_instruction::transpose(int64_t axis1, int64_t axis2) {
    if (operand.empty()) return;
    for (size_t o = 1; o < operand.size(); ++o) {
        bh_view &v = operand[o];
        if (!v.isConstant() && !(o == 1 && opcode == BH_GATHER)) v.transpose(axis1, axis2);
    }
    bh_view &out = operand[0];
    const int sa = sweep_axis();
    if (sa == axis1) {
        constant = bh_constant::make(axis2, constant.type);
    } else if (sa == axis2) {
        constant = bh_constant::make(axis1, constant.type);
    }
    if (opcode == BH_SCATTER || opcode == BH_COND_SCATTER) return;
    if (bh_opcode_is_reduction(opcode)) {
        if (sa != axis1 && sa != axis2) {
            const int64_t t1 = sa < axis1 ? axis1 - 1 : axis1;
            const int64_t t2 = sa < axis2 ? axis2 - 1 : axis2;
            out.transpose(t1, t2);
        } else {
            if (sa != axis1) std::swap(axis1, axis2);
            out.insert_axis(axis1, 1, 1);
            out.transpose(axis1, axis2);
            out.remove_axis(axis2);
        }
    } else {
        out.transpose(axis1, axis2);
    }
}

bh_type bh_instruction::operand_type(int idx) const {
    const bh_view &v = operand[idx];
    return v.isConstant() ? constant.type : v.base->dtype();
}

std::string bh_instruction::pprint(bool) const {
    std::ostringstream ss;
    ss << (opcode > BH_MAX_OPCODE_ID ? "ExtMethod" : bh_opcode_text(opcode));
    for (const bh_view &v : operand) {
        ss << " ";
        if (v.isConstant()) {
            ss << constant;
        } else {
            ss << v.pprint();
        }
    }
    return ss.str();
}

bool bh_instruction::operator==(const bh_instruction &other) const {
    if (opcode != other.opcode || operand.size() != other.operand.size()) return false;
    for (size_t i = 0; i < operand.size(); ++i) {
        if (operand[i].isConstant() != other.operand[i].isConstant()) return false;
        if (operand[i].isConstant()) {
            if (constant != other.constant) return false;
        } else if (operand[i] != other.operand[i]) {
            return false;
        }
    }
    return true;
}

bool bh_instr_dependency(const bh_instruction *a, const bh_instruction *b) {
    if (a->operand.empty() || b->operand.empty()) return false;
    for (const bh_view &v : a->operand)
        if (!v.isConstant() && !b->operand[0].isConstant() && !bh_view_disjoint(&b->operand[0], &v)) return true;
    for (const bh_view &v : b->operand)
        if (!v.isConstant() && !a->operand[0].isConstant() && !bh_view_disjoint(&a->operand[0], &v)) return true;
    return false;
}
