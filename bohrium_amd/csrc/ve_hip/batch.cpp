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
#include <set>
#include <unordered_map>

#include "../../../include/bhhip.h"

namespace bohrium {
namespace vehip {

namespace {
const char *kPreludeRef() {
    static const char *p =
#include "kernel_prelude.inc"
        "\n#define BH_MATH_FN __device__ __forceinline__\n"
#include "kernel_math.inc"
        ;
    return p;
}

void check(int rc, const char *what) {
    if (rc != 0) throw std::runtime_error(std::string("hip engine: ") + what + ": " + bhhip_last_error());
}

uint64_t round_up(uint64_t n, uint64_t m) { return (n + m - 1) / m * m; }
} // namespace

EngineHip::BatchGroup::~BatchGroup() {
    if (dev == nullptr && host == nullptr && scan_scratch == nullptr && uploaded == nullptr) return;
    bhhip_synchronize();
    if (dev != nullptr) bhhip_free(dev);
    if (host != nullptr) bhhip_host_free(host);
    if (scan_scratch != nullptr) bhhip_free(scan_scratch);
    if (uploaded != nullptr) bhhip_event_destroy(uploaded);
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
    auto sched = std::make_shared<Schedule>();
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
    for (const auto &gp : cf.schedule->groups) {
        const BatchGroup &g = *gp;
        const KernelPlan &p0 = cf.kernels[g.items[0].kernel].plans[g.items[0].plan];
        size_t packed = 0;
        arg_offsets(p0.args, &packed);
        const uint64_t stride = round_up(packed, 16);
        const std::string path = std::string(dir) + "/batch_" + std::to_string(seq++) + ".txt";
        FILE *f = std::fopen(path.c_str(), "w");
        if (f == nullptr) throw std::runtime_error("hip engine: cannot write " + path);
        std::fprintf(f, "strategy %s\nitems %zu\nblock %u\nstride %llu\nfirst", strategy_name(p0.strategy),
                     g.items.size(), p0.block, static_cast<unsigned long long>(stride));
        uint64_t blocks = 0;
        for (const PlanRef &r : g.items) {
            std::fprintf(f, " %llu", static_cast<unsigned long long>(blocks));
            blocks += cf.kernels[r.kernel].plans[r.plan].grid;
        }
        std::fprintf(f, " %llu\n", static_cast<unsigned long long>(blocks));
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
        std::fprintf(f, "code\n%s", batch_source(p0).c_str());
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
    const uint64_t need = table_off + 4 * (k + 1);

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
    uint32_t *first = reinterpret_cast<uint32_t *>(g.host + table_off);
    uint64_t blocks = 0;
    uint64_t work = 0;
    for (size_t i = 0; i < k; ++i) {
        KernelPlan &p = cf.kernels[g.items[i].kernel].plans[g.items[i].plan];
        for (int o : p.param_ordinals) ensureResident(ord2base[o]);
        packArgs(p, ord2base, instrs, scratch_base == nullptr ? nullptr : scratch_base + scratch_off[i],
                 g.host + stride * i, stride);
        first[i] = static_cast<uint32_t>(blocks);
        blocks += p.grid;
        work += p.work_items;
        if (blocks > 0x7fffffffull) throw std::runtime_error("hip engine: batched grid too large");
    }
    first[k] = static_cast<uint32_t>(blocks);
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
    } a = {g.dev, reinterpret_cast<const uint32_t *>(g.dev + table_off), static_cast<uint32_t>(k),
           static_cast<uint32_t>(stride)};
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
        ++stats.launches;
    }
    ++batch_launches;
    batched_items += k;
    stats.launches_by_strategy[std::string("batch-") + strategy_name(p0.strategy)] += 1;
    for (const PlanRef &r : g.items) {
        for (int o : cf.kernels[r.kernel].plans[r.plan].written_ordinals) noteWrite(ord2base[o], buffer(ord2base[o]));
    }
    if (profile) {
        bhhip_event_record(ev1);
        pending_events.push_back({ev0, ev1, {std::string("batch-") + strategy_name(p0.strategy), work, blocks, 0.0,
                                             "items=" + std::to_string(k), 0, ""}});
    }
    stats.time_launch += std::chrono::duration<double>(std::chrono::steady_clock::now() - t).count();
}

} // namespace vehip
} // namespace bohrium
