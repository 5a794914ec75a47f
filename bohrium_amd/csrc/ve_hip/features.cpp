#include "features.hpp"
#include "codegen.hpp"

#include <algorithm>
#include <cmath>
#include <map>
#include <set>
#include <sstream>

#include <bohrium/bh_opcode.h>

namespace bohrium {
namespace vehip {

bool costmodel_predict(const std::vector<float> &x, int64_t n, std::vector<float> &y);  // costmodel.hip
int costmodel_features();
bool launchmodel_predict(const std::vector<float> &x, int64_t n, std::vector<float> &y);
int launchmodel_features();

const char *const kFeatureNames[kNumFeatures] = {
    "log2_elements",   "log2_bytes_in",  "log2_bytes_out",  "log2_bytes_temp", "log2_instrs",   "transcendentals",
    "reductions",      "accumulates",    "gathers_scatters", "max_ndim",       "strided_inputs", "inner_axis_reduce",
    "outer_axis_reduce", "full_reduce",  "input_arrays",    "fp64_fraction"};

namespace {
bool transcendental(int64_t op) {
    switch (op) {
        case BH_EXP: case BH_EXP2: case BH_EXPM1: case BH_LOG: case BH_LOG2: case BH_LOG10: case BH_LOG1P:
        case BH_SQRT: case BH_POWER: case BH_SIN: case BH_COS: case BH_TAN: case BH_SINH: case BH_COSH: case BH_TANH:
        case BH_ARCSIN: case BH_ARCCOS: case BH_ARCTAN: case BH_ARCTAN2: case BH_ARCSINH: case BH_ARCCOSH:
        case BH_ARCTANH: case BH_DIVIDE:
            return true;
        default:
            return false;
    }
}

double view_elems(const bh_view &v) {
    double n = 1;
    for (int64_t d = 0; d < v.ndim; ++d) n *= static_cast<double>(v.shape[d]);
    return n;
}

float lg(double x) { return static_cast<float>(std::log2(1.0 + x)); }
} // namespace

std::vector<float> block_features(const std::vector<jitk::InstrPtr> &instrs) {
    std::set<const bh_base *> freed, constructed;
    for (const jitk::InstrPtr &i : instrs) {
        if (i->opcode == BH_FREE) freed.insert(i->operand[0].base);
        else if (!i->operand.empty() && !i->operand[0].isConstant() && i->constructor) constructed.insert(i->operand[0].base);
    }
    auto is_temp = [&](const bh_base *b) { return freed.count(b) && constructed.count(b); };
    std::map<const bh_base *, double> in_bytes, out_bytes, temp_bytes;
    std::set<const bh_base *> produced;
    std::set<std::pair<const bh_base *, int64_t> > strided;
    double elements = 0, ninstr = 0, trans = 0, red = 0, acc = 0, gs = 0, maxnd = 0, f64 = 0;
    float inner_red = 0, outer_red = 0, full_red = 0;
    for (const jitk::InstrPtr &i : instrs) {
        if (bh_opcode_is_system(i->opcode) || i->operand.empty()) continue;
        ninstr += 1;
        const bh_view &out = i->operand[0];
        elements = std::max(elements, view_elems(bh_opcode_is_sweep(i->opcode) ? i->operand[1] : out));
        if (transcendental(i->opcode)) trans += 1;
        if (bh_opcode_is_reduction(i->opcode)) {
            red += 1;
            const int64_t nd = i->operand[1].ndim;
            if (nd <= 1) full_red = 1;
            else if (i->sweep_axis() == nd - 1) inner_red = 1;
            else outer_red = 1;
        }
        if (bh_opcode_is_accumulate(i->opcode)) acc += 1;
        if (i->opcode == BH_GATHER || i->opcode == BH_SCATTER || i->opcode == BH_COND_SCATTER) gs += 1;
        if (!out.isConstant() && out.base->dtype() == bh_type::FLOAT64) f64 += 1;
        for (size_t o = 0; o < i->operand.size(); ++o) {
            const bh_view &v = i->operand[o];
            if (v.isConstant()) continue;
            maxnd = std::max(maxnd, static_cast<double>(v.ndim));
            const double bytes = view_elems(v) * bh_type_size(v.base->dtype());
            if (o == 0) {
                produced.insert(v.base);
                auto &m = is_temp(v.base) ? temp_bytes : out_bytes;
                m[v.base] = std::max(m[v.base], bytes);
            } else if (!produced.count(v.base) && !is_temp(v.base)) {
                in_bytes[v.base] = std::max(in_bytes[v.base], bytes);
                if (v.ndim > 0 && v.shape[v.ndim - 1] > 1 && v.stride[v.ndim - 1] != 1) strided.insert({v.base, v.start});
            }
        }
    }
    auto total = [](const std::map<const bh_base *, double> &m) {
        double s = 0;
        for (const auto &kv : m) s += kv.second;
        return s;
    };
    std::vector<float> f(kNumFeatures);
    f[0] = lg(elements);
    f[1] = lg(total(in_bytes));
    f[2] = lg(total(out_bytes));
    f[3] = lg(total(temp_bytes));
    f[4] = lg(ninstr);
    f[5] = static_cast<float>(trans);
    f[6] = static_cast<float>(red);
    f[7] = static_cast<float>(acc);
    f[8] = static_cast<float>(gs);
    f[9] = static_cast<float>(maxnd);
    f[10] = static_cast<float>(strided.size());
    f[11] = inner_red;
    f[12] = outer_red;
    f[13] = full_red;
    f[14] = static_cast<float>(in_bytes.size());
    f[15] = ninstr > 0 ? static_cast<float>(f64 / ninstr) : 0.0f;
    return f;
}

std::string features_string(const std::vector<float> &f) {
    std::ostringstream s;
    for (size_t k = 0; k < f.size(); ++k) s << (k ? "," : "") << f[k];
    return s.str();
}

bool predict_block_times(const std::vector<std::vector<jitk::InstrPtr> > &sets, std::vector<double> &us) {
    if (costmodel_features() != kNumFeatures) return false;
    std::vector<float> x;
    x.reserve(sets.size() * kNumFeatures);
    for (const auto &s : sets) {
        const std::vector<float> f = block_features(s);
        x.insert(x.end(), f.begin(), f.end());
    }
    std::vector<float> y;
    if (!costmodel_predict(x, static_cast<int64_t>(sets.size()), y)) return false;
    us.resize(sets.size());
    for (size_t k = 0; k < sets.size(); ++k) us[k] = std::exp(static_cast<double>(y[k]));
    return true;
}

std::vector<float> launch_features(const std::vector<KernelPlan> &plans, int unroll_or_items) {
    const KernelPlan *big = nullptr;
    for (const KernelPlan &p : plans)
        if (big == nullptr || p.work_items > big->work_items) big = &p;
    std::vector<float> f(kNumLaunchFeatures, 0.0f);
    if (big == nullptr) return f;
    const double threads = static_cast<double>(std::max<uint64_t>(1, big->grid)) * static_cast<double>(big->block);
    f[0] = static_cast<float>(std::log2(static_cast<double>(std::max<uint64_t>(1, big->grid))));
    f[1] = static_cast<float>(std::log2(static_cast<double>(std::max(1, unroll_or_items))));
    f[2] = static_cast<float>(std::log2(std::max(1.0, static_cast<double>(big->work_items) / threads)));
    return f;
}

bool predict_launch_times(const std::vector<std::vector<float> > &rows, std::vector<double> &us) {
    const int nf = kNumFeatures + kNumLaunchFeatures;
    if (launchmodel_features() != nf) return false;
    std::vector<float> x;
    x.reserve(rows.size() * static_cast<size_t>(nf));
    for (const auto &r : rows) {
        if (static_cast<int>(r.size()) != nf) return false;
        x.insert(x.end(), r.begin(), r.end());
    }
    std::vector<float> y;
    if (!launchmodel_predict(x, static_cast<int64_t>(rows.size()), y)) return false;
    us.resize(rows.size());
    for (size_t k = 0; k < rows.size(); ++k) us[k] = std::exp(static_cast<double>(y[k]));
    return true;
}

} // namespace vehip
} // namespace bohrium
