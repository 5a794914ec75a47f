#include <bohrium/bh_view.hpp>

#include <algorithm>
#include <cassert>
#include <sstream>
#include <stdexcept>

void bh_slide::transpose(int64_t axis1, int64_t axis2) {
    for (bh_slide_dim &d : dims) {
        if (d.rank == axis1) {
            d.rank = axis2;
        } else if (d.rank == axis2) {
            d.rank = axis1;
        }
    }
    auto a1 = resets.find(axis1);
    auto a2 = resets.find(axis2);
    if (a1 != resets.end() && a2 != resets.end()) {
        std::swap(a1->second, a2->second);
    } else if (a1 != resets.end()) {
        resets[axis2] = a1->second;
        resets.erase(axis1);
    } else if (a2 != resets.end()) {
        resets[axis1] = a2->second;
        resets.erase(axis2);
    }
}

bool bh_view::isContiguous() const {
    if (isConstant()) return false;
    int64_t weight = 1;
    for (int64_t d = ndim - 1; d >= 0; --d) {
        if (shape[d] > 1 && stride[d] != weight) return false;
        weight *= shape[d];
    }
    return true;
}

void bh_view::insert_axis(int64_t dim, int64_t size, int64_t st) {
    if (dim < 0 || dim > ndim) throw std::runtime_error("insert_axis(): dim out of range");
    shape.insert(shape.begin() + dim, size);
    stride.insert(stride.begin() + dim, st);
    ++ndim;
}

void bh_view::remove_axis(int64_t dim) {
    if (dim < 0 || dim >= ndim) throw std::runtime_error("remove_axis(): dim out of range");
    shape.erase(shape.begin() + dim);
    stride.erase(stride.begin() + dim);
    --ndim;
}

void bh_view::transpose(int64_t axis1, int64_t axis2) {
    std::swap(shape[axis1], shape[axis2]);
    std::swap(stride[axis1], stride[axis2]);
    slides.transpose(axis1, axis2);
}

std::string bh_view::pprint(bool) const {
    std::ostringstream ss;
    if (isConstant()) return "CONST";
    ss << "a" << reinterpret_cast<uintptr_t>(base) % 100000 << "{start:" << start << ",shape:[";
    for (int64_t i = 0; i < ndim; ++i) ss << (i ? "," : "") << shape[i];
    ss << "],stride:[";
    for (int64_t i = 0; i < ndim; ++i) ss << (i ? "," : "") << stride[i];
    ss << "]}";
    return ss.str();
}

bool bh_view::operator<(const bh_view &o) const {
    if (base != o.base) return base < o.base;
    if (start != o.start) return start < o.start;
    if (ndim != o.ndim) return ndim < o.ndim;
    for (int64_t i = 0; i < ndim; ++i)
        if (shape[i] != o.shape[i]) return shape[i] < o.shape[i];
    for (int64_t i = 0; i < ndim; ++i)
        if (stride[i] != o.stride[i]) return stride[i] < o.stride[i];
    return false;
}

bool bh_view::operator==(const bh_view &o) const {
    if (base == nullptr || o.base == nullptr) return false;
    return base == o.base && ndim == o.ndim && start == o.start && shape == o.shape && stride == o.stride;
}

bool bh_view_same_shape(const bh_view *a, const bh_view *b) {
    if (a->ndim != b->ndim) return false;
    for (int64_t i = 0; i < a->ndim; ++i)
        if (a->shape[i] != b->shape[i]) return false;
    return true;
}

bool bh_view_disjoint(const bh_view *a, const bh_view *b) { return a->base != b->base; }
