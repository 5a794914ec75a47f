/* A strided view into a base array (start, shape and stride in elements).
 *
 * Field meaning, ordering operators and the disjointness rule follow
 * include/bohrium/bh_view.hpp:45-191 and core/bh_view.cpp:188-192 of the reference.
 * Sliding-view metadata (bh_slide) follows include/bohrium/bh_slide.hpp.
 */
#pragma once

#include <cstdint>
#include <map>
#include <string>
#include <tuple>
#include <vector>
#include <bohrium/bh_base.hpp>

constexpr int BH_MAXDIM = 16;

/* Shape/stride vector with the helpers the IR needs. */
class BhIntVec : public std::vector<int64_t> {
public:
    using std::vector<int64_t>::vector;
    BhIntVec() = default;
    BhIntVec(const std::vector<int64_t> &v) : std::vector<int64_t>(v) {}
    int64_t prod() const {
        int64_t p = 1;
        for (int64_t x : *this) p *= x;
        return p;
    }
};

struct bh_slide_dim {
    int64_t rank = 0;
    int64_t offset_change = 0;
    int64_t shape_change = 0;
    int64_t stride = 0;
    int64_t shape = 0;
    int64_t step_delay = 1;
};

struct bh_slide {
    std::vector<bh_slide_dim> dims;
    int64_t iteration_counter = 0;
    std::map<int64_t, std::pair<int64_t, int64_t> > resets;
    void transpose(int64_t axis1, int64_t axis2);
};

struct bh_view {
    bh_base *base = nullptr;
    int64_t start = 0;
    int64_t ndim = 0;
    BhIntVec shape;
    BhIntVec stride;
    bh_slide slides;

    bh_view() = default;
    bh_view(bh_base *base, int64_t start, int64_t ndim, BhIntVec shape, BhIntVec stride)
        : base(base), start(start), ndim(ndim), shape(std::move(shape)), stride(std::move(stride)) {}
    explicit bh_view(bh_base *b) : bh_view(b, 0, 1, BhIntVec{b->nelem()}, BhIntVec{1}) {}

    bool isConstant() const { return base == nullptr; }
    bool is_scalar() const { return shape.prod() == 1; }
    bool isContiguous() const;
    bool hasSlide() const { return !slides.dims.empty(); }

    void insert_axis(int64_t dim, int64_t size, int64_t stride);
    void remove_axis(int64_t dim);
    void transpose(int64_t axis1, int64_t axis2);

    std::string pprint(bool python_notation = true) const;

    bool operator<(const bh_view &other) const;
    bool operator==(const bh_view &other) const;
    bool operator!=(const bh_view &other) const { return !(*this == other); }
};

bool bh_view_same_shape(const bh_view *a, const bh_view *b);
/* Returns true only when the two views provably touch no common element.
 * Views of the same base are always treated as overlapping. */
bool bh_view_disjoint(const bh_view *a, const bh_view *b);
