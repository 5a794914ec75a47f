/* Batched execution of a flush: launches that share one kernel code run as one launch.
 *
 * A flush with many small uneven blocks of the same structure (e.g. the same
 * map + reduce + scan + gather applied to thousands of arrays) would otherwise
 * cost one launch per block and be bound by launch latency rather than HBM.
 * The plans of the flush are ordered in levels by their base accesses; inside
 * a level, every group of at least batch_min plans with the same kernel source
 * becomes one launch of bh_kernel_batch (codegen.cpp batch_source): workgroup b
 * of the batched grid finds its item by binary search in a table of first
 * workgroups and runs that item's body with the item's own packed arguments.
 * FULL_REDUCE items then get one finalize workgroup each, SCAN items their own
 * zeroed look-back region.  There is no counterpart in the reference engines,
 * which launch every kernel block separately (ve/cuda/engine_cuda.cpp:276-326).
 */
#include "engine.hpp"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <stdexcept>
#include <map>
#include <set>
#include <unordered_map>

#include "../../../include/bhhip.h"

namespace bohrium {
namespace vehip {

namespace {
const char *kPreludeRef() {
    static const char *p =
#include "kernel_prelude.inc"
        "\n#define BH_MATH_FN __device__ __forceinline__\n#define BH_MATH_TABLE static __constant__ const\n"
#include "kernel_math.inc"
        ;
    return p;
}

void check(int rc, const char *what) {
    if (rc != 0) throw std::runtime_error(std::string("hip engine: ") + what + ": " + bhhip_last_error());
}

uint64_t round_up(uint64_t n, uint64_t m) { return (n + m - 1) / m * m; }

/* The workgroup table of a batched launch (batch_source) from the items' workgroup counts; sets the
 * launch's `items` argument and its total workgroups.  Grid-stride and reduction skeletons: the first
 * workgroup of each item and the end.
 *
 * Scans: the item of each workgroup (kNoItem: none, the workgroup returns at once), then from
 * position `items` each item's tile count.  A scan body that gathers reads its source at random;
 * workgroups go to the XCDs round robin (workgroup b to XCD b mod `xcds`), each XCD with its own L2,
 * so all tiles of an item go to one XCD, which then fetches each line of the item's source once
 * rather than once per XCD.  Items are spread over the XCDs largest first, each to the XCD with the
 * fewest tiles so far.  On each XCD the items go in consecutive groups of `group`; inside a group
 * the tiles are interleaved (tile r of every item of the group with more than r tiles, largest
 * items first): a tile waits for its predecessor's prefix, and one item's tiles launched back to back
 * would hold workgroup slots while they wait in line. */
constexpr uint32_t kNoItem = 0xffffffffu;

std::vector<uint32_t> batch_table(Strategy s, const std::vector<uint64_t> &grids, int group, int xcds,
                                  uint32_t *items_arg, uint64_t *blocks) {
    const size_t k = grids.size();
    std::vector<uint32_t> t;
    uint64_t total = 0;
    for (uint64_t g : grids) total += g;
    if (total > 0x7fffffffull) throw std::runtime_error("hip engine: batched grid too large");
    if (s != Strategy::SCAN) {
        *blocks = total;
        t.resize(k + 1);
        uint64_t b = 0;
        for (size_t i = 0; i < k; ++i) {
            t[i] = static_cast<uint32_t>(b);
            b += grids[i];
        }
        t[k] = static_cast<uint32_t>(b);
        *items_arg = static_cast<uint32_t>(k);
        return t;
    }
    const size_t X = static_cast<size_t>(std::max(xcds, 1));
    std::vector<uint32_t> by_size(k);
    for (size_t i = 0; i < k; ++i) by_size[i] = static_cast<uint32_t>(i);
    std::stable_sort(by_size.begin(), by_size.end(), [&](uint32_t a, uint32_t b) { return grids[a] > grids[b]; });
    std::vector<std::vector<uint32_t> > mine(X);  // items of each XCD
    std::vector<uint64_t> load(X, 0);
    for (uint32_t i : by_size) {
        const size_t x = static_cast<size_t>(std::min_element(load.begin(), load.end()) - load.begin());
        mine[x].push_back(i);
        load[x] += grids[i];
    }
    const size_t G = static_cast<size_t>(std::max(group, 1));
    std::vector<std::vector<uint32_t> > seq(X);  // workgroups of each XCD in launch order
    for (size_t x = 0; x < X; ++x) {
        std::vector<uint32_t> &items = mine[x];
        std::sort(items.begin(), items.end());  // the flush's order
        for (size_t g0 = 0; g0 < items.size(); g0 += G) {
            std::vector<uint32_t> order(items.begin() + static_cast<std::ptrdiff_t>(g0),
                                        items.begin() + static_cast<std::ptrdiff_t>(std::min(items.size(), g0 + G)));
            std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return grids[a] > grids[b]; });
            const uint64_t R = order.empty() ? 0 : grids[order[0]];
            size_t c = order.size();
            for (uint64_t r = 0; r < R; ++r) {
                while (c > 0 && grids[order[c - 1]] <= r) --c;
                seq[x].insert(seq[x].end(), order.begin(), order.begin() + static_cast<std::ptrdiff_t>(c));
            }
        }
    }
    size_t L = 0;
    for (const auto &q : seq) L = std::max(L, q.size());
    *blocks = static_cast<uint64_t>(L) * X;
    t.assign(L * X + k, kNoItem);
    for (size_t x = 0; x < X; ++x)
        for (size_t j = 0; j < seq[x].size(); ++j) t[j * X + x] = seq[x][j];
    for (size_t i = 0; i < k; ++i) t[L * X + i] = static_cast<uint32_t>(grids[i]);
    *items_arg = static_cast<uint32_t>(L * X);
    return t;
}
} // namespace

EngineHip::BatchGroup::~BatchGroup() {
    if (dev == nullptr && host == nullptr && scan_scratch == nullptr && uploaded == nullptr) return;
    bhhip_synchronize();
    if (dev != nullptr) bhhip_free(dev);
    if (host != nullptr) bhhip_host_free(host);
    if (scan_scratch != nullptr) bhhip_free(scan_scratch);
    if (uploaded != nullptr) bhhip_event_destroy(uploaded);
}

EngineHip::PersistPlan::~PersistPlan() {
    if (dev == nullptr && host == nullptr && scratch == nullptr && uploaded == nullptr) return;
    bhhip_synchronize();
    if (dev != nullptr) bhhip_free(dev);
    if (host != nullptr) bhhip_host_free(host);
    if (scratch != nullptr) bhhip_free(scratch);
    if (uploaded != nullptr) bhhip_event_destroy(uploaded);
}

std::unique_ptr<EngineHip::PersistPlan> EngineHip::buildPersist(CachedFlush &cf) {
    const Schedule &s = *cf.schedule;
    std::unique_ptr<PersistPlan> pp(new PersistPlan());
    std::unordered_map<uint64_t, uint32_t> code_of;
    std::vector<uint32_t> plan_code;
    // Every plan must fit the persistent kernel: 256-thread workgroups, a skeleton whose body only
    // uses its workgroup number and count (or a ticket), and few distinct codes.
    auto add = [&](const PlanRef &r) -> bool {
        const KernelPlan &p = cf.kernels[r.kernel].plans[r.plan];
        if (p.block != 256 || p.grid == 0 || p.grid > 0x7fffffffull) return false;
        if (p.strategy != Strategy::MAP && p.strategy != Strategy::FULL_REDUCE && p.strategy != Strategy::SCAN &&
            p.strategy != Strategy::ROW_REDUCE)
            return false;
        auto it = code_of.find(p.source_hash);
        if (it == code_of.end()) {
            if (pp->codes.size() >= 32) return false;
            it = code_of.emplace(p.source_hash, static_cast<uint32_t>(pp->codes.size())).first;
            pp->codes.push_back(&p);
        }
        size_t packed = 0;
        arg_offsets(p.args, &packed);
        pp->arg_off.push_back(pp->args_bytes);
        pp->args_bytes += round_up(packed, 16);
        pp->scratch_off.push_back(pp->scratch_bytes);
        pp->scratch_bytes += round_up(static_cast<uint64_t>(std::max<int64_t>(p.scratch_bytes, 0)), 256);
        plan_code.push_back(it->second);
        pp->plans.push_back(r);
        uint64_t nb = p.grid;
        if (p.strategy != Strategy::SCAN && persist_items > 0)
            nb = std::min<uint64_t>(nb, std::max<uint64_t>(1, (p.work_items + persist_items - 1) / persist_items));
        pp->nblk.push_back(static_cast<uint32_t>(nb));
        return true;
    };
    auto task = [&](uint32_t c, uint32_t b, uint32_t nb, uint64_t off) {
        pp->tasks.insert(pp->tasks.end(), {c, b, nb, static_cast<uint32_t>(off / 16)});
    };
    auto close_level = [&](size_t first) {
        const size_t n = pp->tasks.size() / 4;
        if (n == first) return;
        pp->levels.push_back(static_cast<uint32_t>(first));
        pp->max_level_tasks = std::max<uint32_t>(pp->max_level_tasks, static_cast<uint32_t>(n - first));
    };
    for (const Level &lv : s.levels) {
        const size_t first_plan = pp->plans.size();
        std::vector<PlanRef> refs = lv.singles;
        for (size_t g : lv.groups) refs.insert(refs.end(), s.groups[g]->items.begin(), s.groups[g]->items.end());
        for (const PlanRef &r : refs)
            if (!add(r)) return nullptr;
        // Main bodies: the scans first, placed per XCD like a batched scan (their position in the
        // level is their workgroup's position, and the grid is a multiple of 8), then the others.
        size_t first = pp->tasks.size() / 4;
        std::vector<size_t> scans;
        std::vector<uint64_t> grids;
        for (size_t i = first_plan; i < pp->plans.size(); ++i) {
            const KernelPlan &p = cf.kernels[pp->plans[i].kernel].plans[pp->plans[i].plan];
            if (p.strategy == Strategy::SCAN) {
                scans.push_back(i);
                grids.push_back(p.grid);
            }
        }
        if (!scans.empty()) {
            uint32_t items_arg = 0;
            uint64_t blocks = 0;
            const std::vector<uint32_t> table =
                batch_table(Strategy::SCAN, grids, batch_scan_group, batch_xcds, &items_arg, &blocks);
            for (uint64_t b = 0; b < blocks; ++b) {
                if (table[b] == kNoItem) {
                    task(kNoItem, 0, 0, 0);
                    continue;
                }
                const size_t i = scans[table[b]];
                task(2 * plan_code[i], 0, static_cast<uint32_t>(grids[table[b]]), pp->arg_off[i]);
            }
        }
        for (size_t i = first_plan; i < pp->plans.size(); ++i) {
            const KernelPlan &p = cf.kernels[pp->plans[i].kernel].plans[pp->plans[i].plan];
            if (p.strategy == Strategy::SCAN) continue;
            for (uint32_t b = 0; b < pp->nblk[i]; ++b) task(2 * plan_code[i], b, pp->nblk[i], pp->arg_off[i]);
        }
        close_level(first);
        // Finalizes of the level's full reductions, after a grid barrier.
        first = pp->tasks.size() / 4;
        for (size_t i = first_plan; i < pp->plans.size(); ++i) {
            const KernelPlan &p = cf.kernels[pp->plans[i].kernel].plans[pp->plans[i].plan];
            if (p.strategy == Strategy::FULL_REDUCE) task(2 * plan_code[i] + 1, 0, 1, pp->arg_off[i]);
        }
        close_level(first);
    }
    if (pp->plans.empty() || pp->tasks.size() / 4 > 0x7fffffffull) return nullptr;
    pp->levels.push_back(static_cast<uint32_t>(pp->tasks.size() / 4));
    return pp;
}

bool EngineHip::launchPersist(PersistPlan &pp, CachedFlush &cf, const std::vector<bh_base *> &ord2base,
                              const std::vector<bh_instruction *> &instrs) {
    auto t = std::chrono::steady_clock::now();
    if (pp.fn == nullptr) {
        const std::string src = persist_source(pp.codes);
        check(bhhip_compile(kPreludeRef(), src.c_str(), "bh_kernel_persist", cache_dir.c_str(), &pp.fn),
              "kernel compilation");
        // Every workgroup must be resident at once (the grid barriers wait for all of them); a
        // multiple of 8 keeps a task's XCD that of its position in its level.
        int per_cu = 0, cus = 0;
        uint64_t mem = 0;
        check(bhhip_occupancy(pp.fn, 256, &per_cu), "occupancy");
        check(bhhip_device_props(&cus, &mem), "device properties");
        const uint64_t cap = static_cast<uint64_t>(std::max(per_cu, 0)) * static_cast<uint64_t>(std::max(cus, 0));
        uint64_t want = round_up(std::max<uint64_t>(pp.max_level_tasks, 8), 8);
        if (persist_grid > 0) want = std::min<uint64_t>(want, static_cast<uint64_t>(persist_grid));
        pp.grid = static_cast<uint32_t>(std::min(cap, want) / 8 * 8);
        stats.time_compile += std::chrono::duration<double>(std::chrono::steady_clock::now() - t).count();
        t = std::chrono::steady_clock::now();
    }
    if (pp.grid == 0) return false;
    if (pp.scratch == nullptr && pp.scratch_bytes > 0) {
        void *p = nullptr;
        check(bhhip_malloc(&p, pp.scratch_bytes), "persistent scratch");
        check(bhhip_memzero(p, pp.scratch_bytes), "persistent scratch");
        ++scratch_allocs;
        pp.scratch = static_cast<unsigned char *>(p);
    }
    const uint64_t tasks_b = 4 * pp.tasks.size(), levels_b = 4 * pp.levels.size();
    const uint64_t bar_at = round_up(tasks_b + levels_b, 16);
    pp.args_at = bar_at + 16;
    const uint64_t need = pp.args_at + pp.args_bytes;
    if (pp.dev == nullptr) {
        void *d = nullptr;
        check(bhhip_malloc(&d, need), "persistent tables");
        pp.dev = static_cast<unsigned char *>(d);
        pp.dev_bytes = need;
    }
    if (pp.uploaded != nullptr) check(bhhip_event_synchronize(pp.uploaded), "persistent staging");
    if (pp.host == nullptr) {
        void *h = nullptr;
        check(bhhip_host_alloc(&h, need), "persistent staging");
        pp.host = static_cast<uint8_t *>(h);
        pp.host_bytes = need;
        std::memset(pp.host, 0, pp.args_at);
        std::memcpy(pp.host, pp.tasks.data(), tasks_b);
        std::memcpy(pp.host + tasks_b, pp.levels.data(), levels_b);
    }
    uint64_t work = 0;
    for (size_t i = 0; i < pp.plans.size(); ++i) {
        const KernelPlan &p = cf.kernels[pp.plans[i].kernel].plans[pp.plans[i].plan];
        for (int o : p.param_ordinals) ensureResident(ord2base[o]);
        // (a full reduction's finalize folds one partial per task of its main body)
        packArgs(p, ord2base, instrs, pp.scratch == nullptr ? nullptr : pp.scratch + pp.scratch_off[i],
                 pp.host + pp.args_at + pp.arg_off[i], static_cast<size_t>(pp.args_bytes - pp.arg_off[i]),
                 pp.nblk[i]);
        work += p.work_items;
    }
    if (!pp.tables_uploaded) {
        check(bhhip_memcpy_h2d_async(pp.dev, pp.host, need), "persistent tables upload");
        pp.tables_uploaded = true;
    } else {
        check(bhhip_memcpy_h2d_async(pp.dev + pp.args_at, pp.host + pp.args_at, pp.args_bytes), "persistent arguments upload");
    }
    if (pp.uploaded == nullptr) check(bhhip_event_create(&pp.uploaded), "event");
    check(bhhip_event_record(pp.uploaded), "event");

    void *ev0 = nullptr, *ev1 = nullptr;
    if (profile) {
        bhhip_event_create(&ev0);
        bhhip_event_create(&ev1);
        bhhip_event_record(ev0);
    }
    const unsigned char *a_args = pp.dev + pp.args_at;
    const void *a_tasks = pp.dev;
    const void *a_levels = pp.dev + tasks_b;
    unsigned a_nlevels = static_cast<unsigned>(pp.levels.size() - 1);
    unsigned *a_bar = reinterpret_cast<unsigned *>(pp.dev + bar_at);
    unsigned *a_err = counters + 1;
    unsigned a_spin = persist_spin;
    void *params[] = {&a_args, &a_tasks, &a_levels, &a_nlevels, &a_bar, &a_err, &a_spin};
    if (bhhip_launch_cooperative(pp.fn, pp.grid, 256, params) != 0) {
        // (e.g. a grid larger than the device holds at once): the levels run instead, now and later
        if (ev0 != nullptr) bhhip_event_destroy(ev0);
        if (ev1 != nullptr) bhhip_event_destroy(ev1);
        pp.grid = 0;
        ++persist_fallbacks;
        return false;
    }
    // The stop event goes right behind the launch: recorded after the host work below, it would time
    // that work too whenever the device finishes first.
    if (profile) bhhip_event_record(ev1);
    ++stats.launches;
    ++persist_launches;
    ++stats.launches_by_strategy["persistent"];
    errors_unchecked = true;
    for (const PlanRef &r : pp.plans)
        for (int o : cf.kernels[r.kernel].plans[r.plan].written_ordinals) noteWrite(ord2base[o], buffer(ord2base[o]));
    if (profile) {
        pending_events.push_back({ev0, ev1, {"persistent", work, pp.grid, 0.0,
                                             "plans=" + std::to_string(pp.plans.size()) + " tasks=" +
                                                 std::to_string(pp.tasks.size() / 4) + " levels=" +
                                                 std::to_string(pp.levels.size() - 1),
                                             0, ""}});
    }
    stats.time_launch += std::chrono::duration<double>(std::chrono::steady_clock::now() - t).count();
    return true;
}

void EngineHip::buildSchedule(CachedFlush &cf) {
    cf.schedule_built = true;
    if (batch_min <= 0) return;
    std::unordered_map<uint64_t, size_t> count;
    for (const CachedKernel &ck : cf.kernels)
        for (const KernelPlan &p : ck.plans) ++count[p.source_hash];
    bool any = false;
    for (const auto &kv : count) any |= kv.second >= static_cast<size_t>(batch_min);
    if (!any) return;
    auto sched = std::make_shared<Schedule>();

    // Levels from base accesses (ordinals of the flush's bases).
    std::unordered_map<int, int> last_read, last_write;
    auto get = [](const std::unordered_map<int, int> &m, int o) {
        auto it = m.find(o);
        return it == m.end() ? -1 : it->second;
    };
    std::vector<std::vector<int> > plan_level(cf.kernels.size());
    std::vector<int> free_level(cf.kernels.size(), 0);
    int nlevels = 0;
    for (size_t k = 0; k < cf.kernels.size(); ++k) {
        const CachedKernel &ck = cf.kernels[k];
        int kmax = 0;
        for (const KernelPlan &p : ck.plans) {
            int L = 0;
            for (int o : p.param_ordinals) L = std::max(L, get(last_write, o) + 1);
            for (int o : p.written_ordinals) L = std::max(L, std::max(get(last_read, o), get(last_write, o)) + 1);
            for (int o : p.param_ordinals) last_read[o] = std::max(get(last_read, o), L);
            for (int o : p.written_ordinals) last_write[o] = std::max(get(last_write, o), L);
            plan_level[k].push_back(L);
            kmax = std::max(kmax, L);
        }
        // A base is released after the last level that touches it (and after this kernel).
        int F = kmax;
        for (int o : ck.free_ordinals) F = std::max(F, std::max(get(last_read, o), get(last_write, o)));
        for (int o : ck.free_ordinals) last_write[o] = std::max(get(last_write, o), F);
        free_level[k] = F;
        nlevels = std::max(nlevels, F + 1);
    }
    sched->levels.resize(static_cast<size_t>(nlevels));
    std::vector<std::unordered_map<uint64_t, std::vector<PlanRef> > > by_hash(static_cast<size_t>(nlevels));
    std::vector<std::vector<uint64_t> > hash_order(static_cast<size_t>(nlevels));
    for (size_t k = 0; k < cf.kernels.size(); ++k) {
        for (size_t p = 0; p < cf.kernels[k].plans.size(); ++p) {
            const int L = plan_level[k][p];
            const uint64_t h = cf.kernels[k].plans[p].source_hash;
            auto &bucket = by_hash[static_cast<size_t>(L)][h];
            if (bucket.empty()) hash_order[static_cast<size_t>(L)].push_back(h);
            bucket.push_back({static_cast<uint32_t>(k), static_cast<uint32_t>(p)});
        }
        sched->levels[static_cast<size_t>(free_level[k])].free_kernels.push_back(static_cast<uint32_t>(k));
    }
    for (size_t L = 0; L < sched->levels.size(); ++L) {
        Level &lv = sched->levels[L];
        for (uint64_t h : hash_order[L]) {
            std::vector<PlanRef> &items = by_hash[L][h];
            if (items.size() >= static_cast<size_t>(batch_min)) {
                sched->groups.emplace_back(new BatchGroup());
                sched->groups.back()->items = std::move(items);
                lv.groups.push_back(sched->groups.size() - 1);
            } else {
                lv.singles.insert(lv.singles.end(), items.begin(), items.end());
            }
        }
        // Singles keep their original order.
        std::sort(lv.singles.begin(), lv.singles.end(), [](const PlanRef &a, const PlanRef &b) {
            return a.kernel != b.kernel ? a.kernel < b.kernel : a.plan < b.plan;
        });
    }
    cf.schedule = sched;
}

void EngineHip::runSchedule(CachedFlush &cf, const std::vector<bh_base *> &ord2base,
                            const std::vector<bh_instruction *> &instrs) {
    stats.kernels += cf.kernels.size();
    Schedule &s = *cf.schedule;
    // (a recorded iteration keeps the level launches: graph replay makes ordinary launches)
    if (persistent && recording == nullptr) {
        if (!s.persist_checked) {
            s.persist = buildPersist(cf);
            s.persist_checked = true;
        }
        if (s.persist && launchPersist(*s.persist, cf, ord2base, instrs)) {
            for (Level &lv : s.levels)
                for (uint32_t k : lv.free_kernels)
                    for (int o : cf.kernels[k].free_ordinals) release(ord2base[o]);
            return;
        }
    }
    for (Level &lv : s.levels) {
        for (const PlanRef &r : lv.singles) {
            kernel_seq = r.kernel + 1000000000ull;
            launchPlan(cf.kernels[r.kernel].plans[r.plan], ord2base, instrs);
        }
        for (size_t g : lv.groups) launchBatch(*s.groups[g], cf, ord2base, instrs);
        for (uint32_t k : lv.free_kernels)
            for (int o : cf.kernels[k].free_ordinals) release(ord2base[o]);
    }
}

// Compile-only mode with BH_HIP_DUMP_DIR: one file per batched launch with the batch source, the
// first-workgroup table and every item's packed arguments (bases as ordinal + 1 shifted by 32 bits),
// so the batched launch can be replayed on the host with checked arrays.
void EngineHip::dumpSchedule(CachedFlush &cf, const std::vector<bh_base *> &ord2base,
                             const std::vector<bh_instruction *> &instrs, const char *dir) {
    static uint64_t seq = 0;
    if (persistent) {
        // the persistent plan: its source compiled (compile-only), its levels and tasks
        std::unique_ptr<PersistPlan> pp = buildPersist(cf);
        const std::string path = std::string(dir) + "/persist_" + std::to_string(seq++) + ".txt";
        FILE *f = std::fopen(path.c_str(), "w");
        if (f == nullptr) throw std::runtime_error("hip engine: cannot write " + path);
        if (!pp) {
            std::fprintf(f, "persistent none\n");
        } else {
            const std::string src = persist_source(pp->codes);
            check(bhhip_compile(kPreludeRef(), src.c_str(), "bh_kernel_persist", cache_dir.c_str(), nullptr),
                  "kernel compilation");
            std::fprintf(f, "plans %zu\ncodes %zu\nlevels", pp->plans.size(), pp->codes.size());
            for (uint32_t v : pp->levels) std::fprintf(f, " %u", v);
            std::fprintf(f, "\ncode_strategies");
            for (const KernelPlan *c : pp->codes) std::fprintf(f, " %s", strategy_name(c->strategy));
            std::fprintf(f, "\ngrids");
            for (const PlanRef &r : pp->plans) std::fprintf(f, " %llu", static_cast<unsigned long long>(cf.kernels[r.kernel].plans[r.plan].grid));
            std::fprintf(f, "\nnblk");
            for (uint32_t v : pp->nblk) std::fprintf(f, " %u", v);
            std::fprintf(f, "\narg_off");
            for (uint64_t v : pp->arg_off) std::fprintf(f, " %llu", static_cast<unsigned long long>(v));
            std::fprintf(f, "\ntasks");
            for (uint32_t v : pp->tasks) std::fprintf(f, " %u", v);
            std::fprintf(f, "\ncode\n%s", src.c_str());
        }
        std::fclose(f);
    }
    for (const auto &gp : cf.schedule->groups) {
        const BatchGroup &g = *gp;
        const KernelPlan &p0 = cf.kernels[g.items[0].kernel].plans[g.items[0].plan];
        size_t packed = 0;
        arg_offsets(p0.args, &packed);
        const uint64_t stride = round_up(packed, 16);
        const std::string path = std::string(dir) + "/batch_" + std::to_string(seq++) + ".txt";
        FILE *f = std::fopen(path.c_str(), "w");
        if (f == nullptr) throw std::runtime_error("hip engine: cannot write " + path);
        std::vector<uint64_t> grids;
        for (const PlanRef &r : g.items) grids.push_back(cf.kernels[r.kernel].plans[r.plan].grid);
        uint32_t items_arg = 0;
        uint64_t blocks = 0;
        const std::vector<uint32_t> table = batch_table(p0.strategy, grids, batch_scan_group, batch_xcds, &items_arg, &blocks);
        std::fprintf(f, "strategy %s\nitems %zu\nitems_arg %u\nblocks %llu\nblock %u\nstride %llu\ntable",
                     strategy_name(p0.strategy), g.items.size(), items_arg, static_cast<unsigned long long>(blocks),
                     p0.block, static_cast<unsigned long long>(stride));
        for (uint32_t v : table) std::fprintf(f, " %u", v);
        std::fprintf(f, "\n");
        std::vector<uint8_t> buf(stride);
        std::set<int> ords;
        for (const PlanRef &r : g.items) {
            const KernelPlan &p = cf.kernels[r.kernel].plans[r.plan];
            std::fill(buf.begin(), buf.end(), 0);
            packArgs(p, ord2base, instrs, nullptr, buf.data(), stride);
            std::fprintf(f, "item");
            for (uint8_t b : buf) std::fprintf(f, " %u", b);
            std::fprintf(f, "\n");
            ords.insert(p.param_ordinals.begin(), p.param_ordinals.end());
        }
        for (int o : ords)
            std::fprintf(f, "base %d %lld %d\n", o + 1, static_cast<long long>(ord2base[o]->nelem()),
                         bh_type_size(ord2base[o]->dtype()));
        const std::string src = batch_source(p0);
        check(bhhip_compile(kPreludeRef(), src.c_str(), "bh_kernel_batch", cache_dir.c_str(), nullptr),
              "kernel compilation");
        std::fprintf(f, "code\n%s", src.c_str());
        std::fclose(f);
    }
}

void EngineHip::launchBatch(BatchGroup &g, CachedFlush &cf, const std::vector<bh_base *> &ord2base,
                            const std::vector<bh_instruction *> &instrs) {
    auto t = std::chrono::steady_clock::now();
    KernelPlan &p0 = cf.kernels[g.items[0].kernel].plans[g.items[0].plan];
    if (g.fn == nullptr) {
        const std::string src = batch_source(p0);
        check(bhhip_compile(kPreludeRef(), src.c_str(), "bh_kernel_batch", cache_dir.c_str(), &g.fn), "kernel compilation");
        if (!p0.fin_name.empty())
            check(bhhip_compile(kPreludeRef(), src.c_str(), "bh_kernel_fin_batch", cache_dir.c_str(), &g.fin_fn),
                  "kernel compilation");
        stats.time_compile += std::chrono::duration<double>(std::chrono::steady_clock::now() - t).count();
        t = std::chrono::steady_clock::now();
    }
    const size_t k = g.items.size();
    size_t packed = 0;
    arg_offsets(p0.args, &packed);
    const uint64_t stride = round_up(packed, 16);
    const uint64_t table_off = stride * k;
    std::vector<uint64_t> grids(k);
    for (size_t i = 0; i < k; ++i) grids[i] = cf.kernels[g.items[i].kernel].plans[g.items[i].plan].grid;
    uint32_t items_arg = 0;
    uint64_t blocks = 0;
    const std::vector<uint32_t> table = batch_table(p0.strategy, grids, batch_scan_group, batch_xcds, &items_arg, &blocks);
    const uint64_t need = table_off + 4 * table.size();

    // Scratch regions, one per item.
    std::vector<uint64_t> scratch_off(k, 0);
    uint64_t scratch_total = 0;
    for (size_t i = 0; i < k; ++i) {
        const KernelPlan &p = cf.kernels[g.items[i].kernel].plans[g.items[i].plan];
        scratch_off[i] = scratch_total;
        scratch_total += round_up(static_cast<uint64_t>(std::max<int64_t>(p.scratch_bytes, 0)), 256);
    }
    unsigned char *scratch_base = nullptr;
    if (p0.strategy == Strategy::FULL_REDUCE) {
        ensureScratch(reduce_scratch, reduce_scratch_bytes, scratch_total, false);
        scratch_base = reduce_scratch;
    } else if (p0.strategy == Strategy::SCAN) {
        if (g.scan_bytes < scratch_total) {
            if (g.scan_scratch != nullptr) {
                check(bhhip_synchronize(), "synchronize");
                bhhip_free(g.scan_scratch);
                g.scan_scratch = nullptr;
            }
            void *p = nullptr;
            check(bhhip_malloc(&p, scratch_total), "batch scan state");
            check(bhhip_memzero(p, scratch_total), "batch scan state");
            ++scratch_allocs;
            g.scan_scratch = static_cast<unsigned char *>(p);
            g.scan_bytes = scratch_total;
        }
        scratch_base = g.scan_scratch;
    }

    // Staging: wait until the previous upload from it has completed, grow when needed.
    if (g.uploaded != nullptr) check(bhhip_event_synchronize(g.uploaded), "batch staging");
    if (g.host_bytes < need) {
        if (g.host != nullptr) bhhip_host_free(g.host);
        void *h = nullptr;
        check(bhhip_host_alloc(&h, need), "batch staging");
        g.host = static_cast<uint8_t *>(h);
        g.host_bytes = need;
    }
    if (g.dev_bytes < need) {
        if (g.dev != nullptr) {
            check(bhhip_synchronize(), "synchronize");
            bhhip_free(g.dev);
        }
        void *d = nullptr;
        check(bhhip_malloc(&d, need), "batch arguments");
        g.dev = static_cast<unsigned char *>(d);
        g.dev_bytes = need;
    }
    std::memcpy(g.host + table_off, table.data(), 4 * table.size());
    uint64_t work = 0;
    for (size_t i = 0; i < k; ++i) {
        KernelPlan &p = cf.kernels[g.items[i].kernel].plans[g.items[i].plan];
        for (int o : p.param_ordinals) ensureResident(ord2base[o]);
        packArgs(p, ord2base, instrs, scratch_base == nullptr ? nullptr : scratch_base + scratch_off[i],
                 g.host + stride * i, stride);
        work += p.work_items;
    }
    check(bhhip_memcpy_h2d_async(g.dev, g.host, need), "batch arguments upload");
    if (g.uploaded == nullptr) check(bhhip_event_create(&g.uploaded), "event");
    check(bhhip_event_record(g.uploaded), "event");

    void *ev0 = nullptr, *ev1 = nullptr;
    if (profile) {
        bhhip_event_create(&ev0);
        bhhip_event_create(&ev1);
        bhhip_event_record(ev0);
    }
    struct {
        const unsigned char *args;
        const uint32_t *first;
        uint32_t items, stride;
    } a = {g.dev, reinterpret_cast<const uint32_t *>(g.dev + table_off), items_arg, static_cast<uint32_t>(stride)};
    check(bhhip_launch(g.fn, blocks, p0.block, &a, sizeof(a)), "batched kernel launch");
    errors_unchecked |= p0.uses_error_flag;
    if (recording != nullptr)
        recording->push_back({g.fn, blocks, p0.block,
                              {reinterpret_cast<uint8_t *>(&a), reinterpret_cast<uint8_t *>(&a) + sizeof(a)}});
    ++stats.launches;
    if (g.fin_fn != nullptr) {
        struct {
            const unsigned char *args;
            uint32_t stride, pad;
        } f = {g.dev, static_cast<uint32_t>(stride), 0u};
        check(bhhip_launch(g.fin_fn, k, p0.block, &f, sizeof(f)), "batched kernel launch");
        if (recording != nullptr)
            recording->push_back({g.fin_fn, k, p0.block,
                                  {reinterpret_cast<uint8_t *>(&f), reinterpret_cast<uint8_t *>(&f) + sizeof(f)}});
        ++stats.launches;
    }
    if (profile) bhhip_event_record(ev1);  // right behind the launches (see launchPersist)
    ++batch_launches;
    batched_items += k;
    stats.launches_by_strategy[std::string("batch-") + strategy_name(p0.strategy)] += 1;
    for (const PlanRef &r : g.items) {
        for (int o : cf.kernels[r.kernel].plans[r.plan].written_ordinals) noteWrite(ord2base[o], buffer(ord2base[o]));
    }
    if (profile) {
        pending_events.push_back({ev0, ev1, {std::string("batch-") + strategy_name(p0.strategy), work, blocks, 0.0,
                                             "items=" + std::to_string(k), 0, ""}});
    }
    stats.time_launch += std::chrono::duration<double>(std::chrono::steady_clock::now() - t).count();
}

} // namespace vehip
} // namespace bohrium
