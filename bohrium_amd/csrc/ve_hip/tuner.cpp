#include "tuner.hpp"

#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>

namespace bohrium {
namespace vehip {

LaunchKnobs LaunchKnobs::of(const WriterConfig &c) {
    LaunchKnobs k;
    k.max_grid = c.max_grid;
    k.reduce_max_grid = c.reduce_max_grid;
    k.map_unroll = c.map_unroll;
    k.reduce_unroll = c.reduce_unroll;
    k.scan_items = c.scan_items;
    return k;
}

WriterConfig LaunchKnobs::apply(WriterConfig c) const {
    c.max_grid = max_grid;
    c.reduce_max_grid = reduce_max_grid;
    c.map_unroll = map_unroll;
    c.reduce_unroll = reduce_unroll;
    c.scan_items = scan_items;
    return c;
}

bool LaunchKnobs::operator==(const LaunchKnobs &o) const {
    return max_grid == o.max_grid && reduce_max_grid == o.reduce_max_grid && map_unroll == o.map_unroll &&
           reduce_unroll == o.reduce_unroll && scan_items == o.scan_items;
}

std::string LaunchKnobs::str() const {
    std::ostringstream ss;
    ss << "max_grid=" << max_grid << " reduce_max_grid=" << reduce_max_grid << " map_unroll=" << map_unroll
       << " reduce_unroll=" << reduce_unroll << " scan_items=" << scan_items;
    return ss.str();
}

std::string Tuner::signature(const std::vector<KernelPlan> &plans) const {
    const KernelPlan *big = nullptr;
    for (const KernelPlan &p : plans)
        if (big == nullptr || p.work_items > big->work_items) big = &p;
    if (big == nullptr || big->strategy == Strategy::SERIAL || big->work_items < min_items) return "";
    std::ostringstream ss;
    for (const KernelPlan &p : plans) {
        char hex[32];
        std::snprintf(hex, sizeof(hex), "%016llx", static_cast<unsigned long long>(p.source_hash));
        ss << strategy_name(p.strategy) << ":" << hex << ",";
    }
    ss << "log2n=" << static_cast<int>(std::floor(std::log2(static_cast<double>(big->work_items))));
    return ss.str();
}

std::vector<LaunchKnobs> Tuner::candidates(Strategy s, const LaunchKnobs &base) const {
    std::vector<LaunchKnobs> out{base};
    auto add = [&](LaunchKnobs k) {
        for (const LaunchKnobs &o : out)
            if (o == k) return;
        out.push_back(k);
    };
    static const int grids[] = {1024, 2048, 4096, 8192};
    switch (s) {
        case Strategy::MAP:
            for (int u : {1, 2, 4})
                for (int g : grids) {
                    LaunchKnobs k = base;
                    k.max_grid = g;
                    k.map_unroll = u;
                    add(k);
                }
            break;
        case Strategy::FULL_REDUCE:
            for (int u : {1, 2, 4, 8})
                for (int g : grids) {
                    LaunchKnobs k = base;
                    k.reduce_max_grid = g;
                    k.reduce_unroll = u;
                    add(k);
                }
            break;
        case Strategy::ROW_REDUCE:
            for (int g : grids) {
                LaunchKnobs k = base;
                k.max_grid = g;
                add(k);
            }
            break;
        case Strategy::SCAN:
            for (int items : {4, 8, 16}) {
                LaunchKnobs k = base;
                k.scan_items = items;
                add(k);
            }
            break;
        default: break;
    }
    return out;
}

const Tuner::Entry *Tuner::lookup(const std::string &key) const {
    auto it = table.find(key);
    return it == table.end() ? nullptr : &it->second;
}

void Tuner::store(const std::string &key, const Entry &e) {
    table[key] = e;
    if (path.empty()) return;
    FILE *f = std::fopen(path.c_str(), "a");
    if (f == nullptr) return;  // the table still holds it for this process
    std::fprintf(f, "%s %d %d %d %d %d %.6f %.6f\n", key.c_str(), e.knobs.max_grid, e.knobs.reduce_max_grid,
                 e.knobs.map_unroll, e.knobs.reduce_unroll, e.knobs.scan_items, e.best_ms, e.default_ms);
    std::fclose(f);
}

void Tuner::load() {
    if (path.empty()) return;
    std::ifstream in(path);
    std::string line;
    while (std::getline(in, line)) {
        std::istringstream ss(line);
        std::string key;
        Entry e;
        if (!(ss >> key >> e.knobs.max_grid >> e.knobs.reduce_max_grid >> e.knobs.map_unroll >>
              e.knobs.reduce_unroll >> e.knobs.scan_items >> e.best_ms >> e.default_ms))
            continue;
        if (e.knobs.max_grid < 1 || e.knobs.reduce_max_grid < 1 || e.knobs.map_unroll < 1 ||
            e.knobs.reduce_unroll < 1 || e.knobs.scan_items < 1 || e.knobs.scan_items > 64)
            continue;
        table[key] = e;
    }
}

std::string Tuner::dump() const {
    std::ostringstream ss;
    for (const auto &kv : table)
        ss << kv.first << " " << kv.second.knobs.str() << " best_ms=" << kv.second.best_ms
           << " default_ms=" << kv.second.default_ms << "\n";
    return ss.str();
}

bool same_launches(const std::vector<KernelPlan> &a, const std::vector<KernelPlan> &b) {
    if (a.size() != b.size()) return false;
    for (size_t i = 0; i < a.size(); ++i)
        if (a[i].grid != b[i].grid || a[i].block != b[i].block || a[i].source != b[i].source) return false;
    return true;
}

} // namespace vehip
} // namespace bohrium
