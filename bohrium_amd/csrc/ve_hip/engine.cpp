#include "engine.hpp"

#include <bohrium/flat_ptr_map.hpp>

#include <sys/mman.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <deque>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <iomanip>
#include <sstream>
#include <stdexcept>

#include <bohrium/bh_main_memory.hpp>
#include "../../../include/bhhip.h"
#include "../../../include/bh_ve_hip.h"
#include "../../../include/bh_mem_signal.h"

namespace bohrium {
namespace vehip {

namespace {
const char *kPrelude =
#include "kernel_prelude.inc"
    "\n#define BH_MATH_FN __device__ __forceinline__\n#define BH_MATH_TABLE static __constant__ const\n"
#include "kernel_math.inc"
    ;

void check(int rc, const char *what) {
    if (rc != 0) throw std::runtime_error(std::string("hip engine: ") + what + ": " + bhhip_last_error());
}

double seconds_since(std::chrono::steady_clock::time_point t) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t).count();
}

uint64_t fnv1a(const std::string &s) {
    uint64_t h = 1469598103934665603ull;
    for (unsigned char c : s) {
        h ^= c;
        h *= 1099511628211ull;
    }
    return h;
}

// Host writes into guarded pages seen by the fault handler (all engines of the process).
std::atomic<uint64_t> g_host_write_faults{0};

// Fault handler of a guarded host copy (signal context, the faulting thread): make the pages
// writable again and, when the host copy was current, make it the newest (the device copy is
// uploaded again before a kernel reads it).  A write into a stale host copy leaves the device
// copy the newest: such writes are lost, as they are when the reference engines hold the array.
int on_host_write(void *, void *idx) {
    DeviceBuffer *d = static_cast<DeviceBuffer *>(idx);
    void *addr = d->guard_addr;
    if (addr == nullptr || mprotect(addr, d->guard_bytes, PROT_READ | PROT_WRITE) != 0) return 0;
    d->guard_addr = nullptr;
    d->guard_bytes = 0;
    bh_mem_signal_detach(addr);
    if (d->host_valid) d->dev_valid = false;
    g_host_write_faults.fetch_add(1);
    return 1;
}

std::string default_cache_dir() {
    const char *xdg = std::getenv("XDG_CACHE_HOME");
    const char *home = std::getenv("HOME");
    if (xdg != nullptr && *xdg) return std::string(xdg) + "/bohrium_amd";
    if (home != nullptr && *home) return std::string(home) + "/.cache/bohrium_amd";
    return "/tmp/bohrium_amd_cache";
}
} // namespace

// ------------------------------------------------------------------ pool

void *DevicePool::alloc(uint64_t bytes) {
    bytes = std::max<uint64_t>(256, (bytes + 255) / 256 * 256);
    // Last released first: a loop that frees and re-creates a temporary gets the same buffer
    // every iteration, so its launches (and a graph of them) stay identical.
    auto it = _free.find(bytes);
    if (it != _free.end() && !it->second.empty()) {
        void *p = it->second.back();
        it->second.pop_back();
        _cached -= bytes;
        ++hits;
        return p;
    }
    ++misses;
    void *p = nullptr;
    if (bhhip_malloc(&p, bytes) != 0) {
        trim(0);
        check(bhhip_malloc(&p, bytes), "device allocation");
    }
    allocated += bytes;
    peak = std::max(peak, allocated);
    return p;
}

void DevicePool::release(void *ptr, uint64_t bytes) {
    if (ptr == nullptr) return;
    bytes = std::max<uint64_t>(256, (bytes + 255) / 256 * 256);
    _free[bytes].push_back(ptr);
    _cached += bytes;
    if (_cached > limit) trim(limit / 2);
}

void DevicePool::trim(uint64_t keep) {
    bool synced = false;
    while (_cached > keep) {
        // the largest size first, the last released buffer of it first
        auto largest = _free.end();
        for (auto it = _free.begin(); it != _free.end(); ++it)
            if (!it->second.empty() && (largest == _free.end() || it->first > largest->first)) largest = it;
        if (largest == _free.end()) break;
        if (!synced) {
            bhhip_synchronize();
            synced = true;
        }
        bhhip_free(largest->second.back());
        largest->second.pop_back();
        ++frees;
        _cached -= largest->first;
        allocated -= largest->first;
        if (largest->second.empty()) _free.erase(largest);
    }
}

DevicePool::~DevicePool() {
    for (auto &e : _free)
        for (void *p : e.second) bhhip_free(p);
}

// ------------------------------------------------------------------ engine

EngineHip::EngineHip(component::ComponentVE &comp) : comp(comp) {
    const ConfigParser &cfg = comp.config;
    const char *dry_env = std::getenv("BH_HIP_DRY_RUN");
    dry_run = cfg.defaultGet<bool>("dry_run", false) || (dry_env != nullptr && std::string(dry_env) == "1");
    if (dry_run) {
        check(bhhip_set_arch(cfg.defaultGet<std::string>("dry_run_arch", "gfx950").c_str()), "set arch");
    } else {
        check(bhhip_init(cfg.defaultGet<int>("device_number", 0)), "init");
        check(bhhip_device_props(&num_cus, nullptr), "device properties");
    }
    fusion_config.pre_fuser = cfg.defaultGet<std::string>("pre_fuser", "lossy");
    fusion_config.fuser_list =
        cfg.defaultGetList("fuser_list", {"greedy", "push_reductions_inwards", "collapse_redundant_axes"});
    fusion_config.greedy_threshold = cfg.defaultGet<uint64_t>("greedy_threshold", 10000);
    fusion_config.avoid_rank0_sweep = false;
    fusion_config.merge_filter = [](const jitk::LoopB &a, const jitk::LoopB &b) { return root_merge_allowed(a, b); };
    fusion_config.accumulate_alone = true;
    fusion_config.one_sweep_axis = true;
    fusion_config.learned_cost = [](const std::vector<std::vector<jitk::InstrPtr> > &sets, std::vector<double> &us) {
        return predict_block_times(sets, us);
    };
    wcfg.max_grid = cfg.defaultGet<int>("max_grid", 2048);
    wcfg.reduce_max_grid = cfg.defaultGet<int>("reduce_max_grid", 4096);
    wcfg.resident_grid = cfg.defaultGet<bool>("resident_grid", false);
    wcfg.reduce_min_items = std::max(1, cfg.defaultGet<int>("reduce_min_items", 4));
    wcfg.scan_items = cfg.defaultGet<int>("scan_items", 8);
    wcfg.map_unroll = std::max(1, cfg.defaultGet<int>("map_unroll", 1));
    wcfg.reduce_unroll = std::max(1, cfg.defaultGet<int>("reduce_unroll", 1));
    wcfg.row_reduce = cfg.defaultGet<bool>("row_reduce", true);
    wcfg.waves_per_eu = std::max(0, cfg.defaultGet<int>("waves_per_eu", 0));
    wcfg.xcd_remap = cfg.defaultGet<bool>("xcd_remap", false);
    wcfg.xcd_parts = cfg.defaultGet<bool>("xcd_parts", true);
    unflatten = cfg.defaultGet<bool>("unflatten_reductions", true);
    resident = cfg.defaultGet<bool>("resident", true);
    track_host_writes = cfg.defaultGet<bool>("host_write_tracking", true);
    graph_repeats = cfg.defaultGet<bool>("graph_repeats", true);
    batch_min = cfg.defaultGet<int>("batch_min", 8);
    batch_scan_group = std::max(1, cfg.defaultGet<int>("batch_scan_group", 8));
    batch_xcds = std::max(1, cfg.defaultGet<int>("batch_xcds", 8));
    persistent = cfg.defaultGet<bool>("batch_persistent", false);
    persist_items = std::max(0, cfg.defaultGet<int>("persist_items", 16384));
    persist_grid = std::max(0, cfg.defaultGet<int>("persist_grid", 0));
    profile = cfg.defaultGet<bool>("prof", false);
    cache_dir = cfg.defaultGet<std::string>("cache_dir", "");
    if (cache_dir.empty() || cache_dir == "NONE") cache_dir = default_cache_dir();
    if (cache_dir == "off") cache_dir.clear();
    pool.limit = cfg.defaultGet<uint64_t>("malloc_cache_limit_bytes", 64ull << 30);
    twin_updates = cfg.defaultGet<bool>("twin_updates", true);
    twin_min_bytes = cfg.defaultGet<uint64_t>("twin_min_bytes", 1u << 20);
    tuner.enabled = cfg.defaultGet<bool>("autotune", true) && !dry_run;
    tuner.samples = std::max(1, cfg.defaultGet<int>("autotune_samples", 3));
    tuner.min_items = cfg.defaultGet<uint64_t>("autotune_min_items", 1u << 20);
    tuner.min_gain = cfg.defaultGet<double>("autotune_min_gain", 0.02);
    {
        const std::string m = cfg.defaultGet<std::string>("autotune_mode", "measure");
        tuner.mode = m == "model" ? Tuner::Mode::MODEL : m == "prune" ? Tuner::Mode::PRUNE : Tuner::Mode::MEASURE;
        tuner.prune_keep = std::max(1, cfg.defaultGet<int>("autotune_prune_keep", 2));
    }
    if (!cache_dir.empty() && cfg.defaultGet<bool>("autotune_persist", true)) {
        tuner.path = cache_dir + "/autotune_" + cfg.defaultGet<std::string>("dry_run_arch", "gfx950") + ".txt";
        tuner.load();
    }
    if (dry_run) return;
    void *c = nullptr;
    check(bhhip_malloc(&c, 256), "allocation of counters");
    check(bhhip_memzero(c, 256), "zeroing counters");
    counters = static_cast<unsigned int *>(c);
}

EngineHip::~EngineHip() {
    for (auto &kv : bufs) unguardHost(kv.second);
    if (dry_run) return;
    bhhip_synchronize();
    for (auto &kv : graph_cache) bhhip_graph_destroy(kv.second);
    graph_cache.clear();
    for (auto &kv : bufs)
        if (!kv.second.external) pool.release(kv.second.ptr, kv.second.bytes);
    bufs.clear();
    if (reduce_scratch) bhhip_free(reduce_scratch);
    if (scan_scratch) bhhip_free(scan_scratch);
    if (counters) bhhip_free(counters);
}

DeviceBuffer &EngineHip::buffer(bh_base *base) {
    auto it = bufs.find(base);
    if (it != bufs.end()) {
        // Buffers are filed by base address; one of another size means a freed base was never
        // released and a new base took its address.  Launching with it would write out of bounds.
        const DeviceBuffer &d = it->second;
        if (d.ptr != nullptr && !d.external && d.bytes != static_cast<uint64_t>(base->nbytes()))
            throw std::logic_error("hip engine: a device buffer of " + std::to_string(d.bytes) +
                                   " bytes is filed under a base of " + std::to_string(base->nbytes()) + " bytes");
        return it->second;
    }
    DeviceBuffer &b = bufs[base];
    b.host_valid = base->getDataPtr() != nullptr;
    return b;
}

void EngineHip::ensureResident(bh_base *base) {
    DeviceBuffer &b = buffer(base);
    if (dry_run) return;
    if (b.ptr == nullptr) {
        b.bytes = static_cast<uint64_t>(base->nbytes());
        b.ptr = pool.alloc(b.bytes);
    }
    if (!b.dev_valid && b.host_valid && base->getDataPtr() != nullptr) {
        const auto t = std::chrono::steady_clock::now();
        check(bhhip_memcpy_h2d(b.ptr, base->getDataPtr(), b.bytes), "copy to device");
        stats.h2d_bytes += b.bytes;
        stats.time_copy += seconds_since(t);
        b.dev_valid = true;
        ++b.writes;
    }
}

void EngineHip::copyToHost(bh_base *base) {
    auto it = bufs.find(base);
    if (it == bufs.end()) return;
    DeviceBuffer &b = it->second;
    if (dry_run) {
        bh_data_malloc(base);
        return;
    }
    if (!b.dev_valid || b.host_valid || b.ptr == nullptr) return;
    const auto t = std::chrono::steady_clock::now();
    unguardHost(b);
    bh_data_malloc(base);
    check(bhhip_memcpy_d2h(base->getDataPtr(), b.ptr, b.bytes), "copy to host");
    stats.d2h_bytes += b.bytes;
    stats.time_copy += seconds_since(t);
    b.host_valid = true;
    // Kernels that raise errors without a check at the end of their flush (scans) are checked at
    // the next copy to the host, which has waited for them anyway.
    if (errors_unchecked) checkErrors();
}

void EngineHip::copyAllToHost() {
    // The host copies become the newest: another component works on them from here on.
    for (auto &kv : bufs) {
        copyToHost(const_cast<bh_base *>(kv.first));
        unguardHost(kv.second);
        if (kv.second.host_valid) kv.second.dev_valid = false;
    }
}

void EngineHip::guardHost(bh_base *base, DeviceBuffer &d) {
    if (d.guard_addr != nullptr) return;
    void *p = base->getDataPtr();
    const uint64_t n = static_cast<uint64_t>(base->nbytes());
    const long page = sysconf(_SC_PAGESIZE);
    // Only pages that hold nothing but this base's data: a live allocation of the engine's allocator
    // (a mapping of its own).  Anything else is handed out as the newest copy instead.
    if (p == nullptr || n == 0 || page <= 0 || !bh_memory_owns(p, n)) {
        if (d.host_valid) d.dev_valid = false;
        return;
    }
    const uint64_t len = (n + page - 1) / page * page;
    if (mprotect(p, len, PROT_READ) != 0) {
        if (d.host_valid) d.dev_valid = false;
        return;
    }
    d.guard_addr = p;
    d.guard_bytes = len;
    if (bh_mem_signal_attach(&d, p, len, on_host_write) != 0) {
        mprotect(p, len, PROT_READ | PROT_WRITE);
        d.guard_addr = nullptr;
        d.guard_bytes = 0;
        if (d.host_valid) d.dev_valid = false;
        return;
    }
    ++host_guards;
}

void EngineHip::unguardHost(DeviceBuffer &d) {
    void *p = d.guard_addr;
    if (p == nullptr) return;
    bh_mem_signal_detach(p);
    if (mprotect(p, d.guard_bytes, PROT_READ | PROT_WRITE) != 0)
        throw std::runtime_error("hip engine: cannot make a guarded host copy writable again");
    d.guard_addr = nullptr;
    d.guard_bytes = 0;
}

void EngineHip::eraseBuffer(std::unordered_map<const bh_base *, DeviceBuffer>::iterator it) {
    unguardHost(it->second);
    if (!it->second.external) pool.release(it->second.ptr, it->second.bytes);
    bufs.erase(it);
}

void EngineHip::handOutHost(bh_base *base) {
    auto it = bufs.find(base);
    if (it == bufs.end() || dry_run) return;
    DeviceBuffer &d = it->second;
    if (!d.host_valid || base->getDataPtr() == nullptr) return;
    if (!d.dev_valid || d.ptr == nullptr) return;  // the host copy is the only current one anyway
    if (track_host_writes) {
        guardHost(base, d);
    } else {
        d.dev_valid = false;
    }
}

void EngineHip::release(bh_base *base) {
    released_now.insert(base);
    ++releases;
    auto it = bufs.find(base);
    if (it != bufs.end()) eraseBuffer(it);
    bh_data_free(base);
    auto tw = twins.find(base);
    if (tw != twins.end()) {
        auto tb = bufs.find(tw->second.base.get());
        if (tb != bufs.end()) eraseBuffer(tb);
        twins.erase(tw);
    }
}

void EngineHip::exchangeTwin(const TwinUpdate &u) {
    DeviceBuffer &g = buffer(u.target);
    DeviceBuffer &t = buffer(u.twin);
    unguardHost(g);
    std::swap(g.ptr, t.ptr);
    noteWrite(u.target, g);
    noteWrite(u.twin, t);
    twins[u.target].in_step_at = u.in_step_after ? g.writes : ~0ull;
    flush_exchanges.push_back(u);
    ++twin_exchanges;
}

void EngineHip::applyEffects(const FlushEffects &f) {
    for (bh_base *b : f.written) ++buffer(b).writes;
    for (const TwinUpdate &u : f.exchanges) exchangeTwin(u);
    flush_written.clear();
    flush_exchanges.clear();
}

void EngineHip::ensureScratch(unsigned char *&ptr, uint64_t &have, uint64_t need, bool zero) {
    if (need <= have) return;
    const uint64_t bytes = std::max<uint64_t>(need, 1 << 20);
    if (ptr) {
        check(bhhip_synchronize(), "synchronize");
        bhhip_free(ptr);
    }
    void *p = nullptr;
    check(bhhip_malloc(&p, bytes), "scratch allocation");
    ++scratch_allocs;
    if (zero) check(bhhip_memzero(p, bytes), "scratch zeroing");
    ptr = static_cast<unsigned char *>(p);
    have = bytes;
}

void EngineHip::launchPlan(KernelPlan &plan, const std::vector<bh_base *> &ord2base,
                           const std::vector<bh_instruction *> &instrs) {
    auto t = std::chrono::steady_clock::now();
    if (dry_run) {
        check(bhhip_compile(kPrelude, plan.source.c_str(), plan.name.c_str(), cache_dir.c_str(), nullptr),
              "kernel compilation");
        stats.time_compile += seconds_since(t);
        if (!plan.fin_name.empty()) ++stats.launches;
        char hex[32];
        std::snprintf(hex, sizeof(hex), "%016llx", static_cast<unsigned long long>(plan.source_hash));
        dry_run_log.push_back(std::string(strategy_name(plan.strategy)) + " " + std::to_string(plan.grid) + " " +
                              std::to_string(plan.block) + " " + hex);
        ++stats.launches;
        ++stats.launches_by_strategy[strategy_name(plan.strategy)];
        for (int o : plan.written_ordinals) noteWrite(ord2base[o], buffer(ord2base[o]));
        static const char *dump = std::getenv("BH_HIP_DUMP_DIR");
        if (dump != nullptr && *dump) dumpLaunch(dump, plan, ord2base, instrs);
        return;
    }
    if (plan.function == nullptr) {
        compilePlan(plan);
        stats.time_compile += seconds_since(t);
        t = std::chrono::steady_clock::now();
    }
    for (int o : plan.param_ordinals) ensureResident(ord2base[o]);
    if (plan.strategy == Strategy::FULL_REDUCE) ensureScratch(reduce_scratch, reduce_scratch_bytes, plan.scratch_bytes, false);
    if (plan.strategy == Strategy::SCAN) ensureScratch(scan_scratch, scan_scratch_bytes, plan.scratch_bytes, true);
    uint8_t buf[4096];
    const size_t n = packArgs(plan, ord2base, instrs, plan.strategy == Strategy::SCAN ? scan_scratch : reduce_scratch,
                              buf, sizeof(buf));

    void *ev0 = nullptr, *ev1 = nullptr;
    if (profile) {
        bhhip_event_create(&ev0);
        bhhip_event_create(&ev1);
        bhhip_event_record(ev0);
    }
    check(bhhip_launch(plan.function, plan.grid, plan.block, buf, n), "kernel launch");
    ++stats.launches;
    errors_unchecked |= plan.uses_error_flag;
    if (recording != nullptr) recording->push_back({plan.function, plan.grid, plan.block, {buf, buf + n}});
    if (plan.fin_function != nullptr) {
        check(bhhip_launch(plan.fin_function, 1, plan.block, buf, n), "kernel launch");
        ++stats.launches;
        if (recording != nullptr) recording->push_back({plan.fin_function, 1, plan.block, {buf, buf + n}});
    }
    if (profile) bhhip_event_record(ev1);  // right behind the launches, before the host work below
    ++stats.launches_by_strategy[strategy_name(plan.strategy)];
    for (int o : plan.written_ordinals) noteWrite(ord2base[o], buffer(ord2base[o]));
    if (profile) {
        char hex[32];
        std::snprintf(hex, sizeof(hex), "%016llx", static_cast<unsigned long long>(plan.source_hash));
        pending_events.push_back({ev0, ev1, {strategy_name(plan.strategy), plan.work_items, plan.grid, 0.0, hex,
                                             kernel_seq, features_string(plan.features)}});
    }
    stats.time_launch += seconds_since(t);
}

// Compile-only mode with BH_HIP_DUMP_DIR: one file per planned launch with its device code, shape and
// the resolved value of every argument (bases as element count and element size), so the launch can
// be replayed on the host by a CPU build of the same code.
void EngineHip::dumpLaunch(const char *dir, const KernelPlan &plan, const std::vector<bh_base *> &ord2base,
                           const std::vector<bh_instruction *> &instrs) {
    static uint64_t seq = 0;
    const std::string path = std::string(dir) + "/launch_" + std::to_string(seq++) + ".txt";
    FILE *f = std::fopen(path.c_str(), "w");
    if (f == nullptr) throw std::runtime_error("hip engine: cannot write " + path);
    std::fprintf(f, "strategy %s\ngrid %llu\nblock %u\nscratch %lld\n", strategy_name(plan.strategy),
                 static_cast<unsigned long long>(plan.grid), plan.block, static_cast<long long>(plan.scratch_bytes));
    for (size_t i = 0; i < plan.args.size(); ++i) {
        const ArgSpec &a = plan.args[i];
        std::fprintf(f, "param %s %s ", plan.param_types[i].c_str(), plan.param_names[i].c_str());
        if (a.kind == ArgSpec::BASE_PTR) {
            const bh_base *b = ord2base[a.base_ordinal];
            std::fprintf(f, "base %lld %d %d\n", static_cast<long long>(b->nelem()), bh_type_size(b->dtype()),
                         a.base_ordinal);
        } else if (a.kind == ArgSpec::GRID_U32) {
            const uint32_t g = static_cast<uint32_t>(plan.grid);
            std::fprintf(f, "bytes %u %u %u %u\n", g & 255u, (g >> 8) & 255u, (g >> 16) & 255u, g >> 24);
        } else if (a.kind != ArgSpec::LITERAL) {
            std::fprintf(f, "engine\n");
        } else {
            uint8_t buf[64];
            const uint8_t *src = a.bytes.data();
            int64_t start = 0;
            if (a.view_origin >= 0 && a.view_origin < static_cast<int64_t>(instrs.size())) {
                start = instrs[a.view_origin]->operand[a.view_operand].start;
                src = reinterpret_cast<const uint8_t *>(&start);
            } else if (a.const_origin >= 0 && a.const_origin < static_cast<int64_t>(instrs.size())) {
                const bh_constant &c = instrs[a.const_origin]->constant;
                src = a.const_part == 0 ? reinterpret_cast<const uint8_t *>(&c.value.r123.start)
                    : a.const_part == 1 ? reinterpret_cast<const uint8_t *>(&c.value.r123.key)
                                        : reinterpret_cast<const uint8_t *>(&c.value);
            }
            std::memcpy(buf, src, static_cast<size_t>(a.size));
            std::fprintf(f, "bytes");
            for (int k = 0; k < a.size; ++k) std::fprintf(f, " %u", buf[k]);
            std::fprintf(f, "\n");
        }
    }
    std::fprintf(f, "code\n%s", plan.device_code.c_str());
    std::fclose(f);
}

void EngineHip::compilePlan(KernelPlan &plan) {
    check(bhhip_compile(kPrelude, plan.source.c_str(), plan.name.c_str(), cache_dir.c_str(), &plan.function),
          "kernel compilation");
    if (!plan.fin_name.empty())
        check(bhhip_compile(kPrelude, plan.source.c_str(), plan.fin_name.c_str(), cache_dir.c_str(), &plan.fin_function),
              "kernel compilation");
    // A grid-stride launch with more workgroups than fit on the device at once runs a second, partial
    // round on a fraction of the CUs; one resident round of workgroups does the same work without it.
    const bool grid_stride = plan.strategy == Strategy::MAP || plan.strategy == Strategy::FULL_REDUCE ||
                             plan.strategy == Strategy::ROW_REDUCE;
    if (wcfg.resident_grid && grid_stride && num_cus > 0) {
        int per_cu = 0;
        check(bhhip_occupancy(plan.function, plan.block, &per_cu), "occupancy query");
        const uint64_t cap = static_cast<uint64_t>(num_cus) * static_cast<uint64_t>(std::max(1, per_cu));
        if (plan.grid > cap) plan.grid = cap;
    }
}

size_t EngineHip::packArgs(const KernelPlan &plan, const std::vector<bh_base *> &ord2base,
                           const std::vector<bh_instruction *> &instrs, void *scratch, uint8_t *out, size_t cap,
                           uint32_t grid_arg) {
    size_t n = 0;
    auto put = [&](const void *p, int size, int align) {
        while (n % static_cast<size_t>(align)) {
            if (n >= cap) throw std::runtime_error("hip engine: kernel arguments exceed their buffer");
            out[n++] = 0;
        }
        if (n + static_cast<size_t>(size) > cap) throw std::runtime_error("hip engine: kernel arguments exceed their buffer");
        std::memcpy(out + n, p, static_cast<size_t>(size));
        n += static_cast<size_t>(size);
    };
    for (const ArgSpec &a : plan.args) {
        void *ptr = nullptr;
        switch (a.kind) {
            case ArgSpec::LITERAL:
                if (a.view_origin >= 0 && a.view_origin < static_cast<int64_t>(instrs.size())) {
                    put(&instrs[a.view_origin]->operand[a.view_operand].start, 8, 8);
                } else if (a.const_origin >= 0 && a.const_origin < static_cast<int64_t>(instrs.size())) {
                    const bh_constant &c = instrs[a.const_origin]->constant;
                    const void *src = a.const_part == 0 ? static_cast<const void *>(&c.value.r123.start)
                                    : a.const_part == 1 ? static_cast<const void *>(&c.value.r123.key)
                                                        : static_cast<const void *>(&c.value);
                    put(src, a.size, a.align);
                } else {
                    put(a.bytes.data(), a.size, a.align);
                }
                continue;
            case ArgSpec::BASE_PTR:
                // (compile-only dumps stand a base's ordinal + 1, shifted, in for its address)
                ptr = dry_run ? reinterpret_cast<void *>(static_cast<uintptr_t>(a.base_ordinal + 1) << 32)
                              : bufs.at(ord2base[a.base_ordinal]).ptr;
                // A base without memory (no elements) stands on zeroed device memory: a gather reads
                // element 0 of its source whatever the index, and raises the index error instead.
                if (ptr == nullptr && !dry_run) ptr = counters + 16;
                break;
            case ArgSpec::GRID_U32: {
                const uint32_t g = grid_arg != 0 ? grid_arg : static_cast<uint32_t>(plan.grid);
                put(&g, 4, 4);
                continue;
            }
            case ArgSpec::SCRATCH_PTR: ptr = scratch; break;
            case ArgSpec::COUNTER_PTR: ptr = counters; break;
            case ArgSpec::ERRFLAG_PTR: ptr = counters + 1; break;
        }
        put(&ptr, 8, 8);
    }
    while (n % 8) {
        if (n >= cap) throw std::runtime_error("hip engine: kernel arguments exceed their buffer");
        out[n++] = 0;
    }
    return n;
}

std::string EngineHip::flushKey(const std::vector<bh_instruction *> &instrs, std::vector<bh_base *> &ord2base) const {
    std::vector<int64_t> words;  // the key as 64-bit words (one string made at the end)
    words.reserve(instrs.size() * 24);
    words.push_back(0);  // the hash of the words after it (FlushKeyHash reads it instead of the whole key)
    FlatPtrMap base2ord(instrs.size());
    auto put = [&](int64_t v) { words.push_back(v); };
    // Equal constants share a kernel parameter: the key holds, for each constant, the position of the
    // first constant of equal type and value in the flush.
    struct ConstKey {
        uint64_t lo, hi;
        int type;
        bool operator==(const ConstKey &o) const { return lo == o.lo && hi == o.hi && type == o.type; }
    };
    struct ConstHash {
        size_t operator()(const ConstKey &k) const {
            return std::hash<uint64_t>()(k.lo * 0x9e3779b97f4a7c15ull ^ k.hi ^ (static_cast<uint64_t>(k.type) << 56));
        }
    };
    std::unordered_map<ConstKey, int64_t, ConstHash> first_equal;
    int64_t nconst = 0;
    for (const bh_instruction *i : instrs) {
        put(i->opcode);
        put(i->constructor ? 1 : 0);
        put(static_cast<int64_t>(i->operand.size()));
        for (const bh_view &v : i->operand) {
            if (v.isConstant()) {
                put(-1);
                put(static_cast<int64_t>(i->constant.type));
                // A sweep's constant is its axis, which shapes the loop nest.
                if (bh_opcode_is_sweep(i->opcode)) put(i->constant.get_int64());
                if (i->constant.type != bh_type::R123) {
                    uint64_t w[2] = {0, 0};
                    std::memcpy(w, &i->constant.value,
                                std::min<size_t>(sizeof(w), static_cast<size_t>(bh_type_size(i->constant.type))));
                    put(first_equal.emplace(ConstKey{w[0], w[1], static_cast<int>(i->constant.type)}, nconst).first->second);
                }
                ++nconst;
                continue;
            }
            const std::pair<int, bool> r = base2ord.insert(v.base, static_cast<int>(ord2base.size()));
            const int o = r.first;
            if (r.second) {
                ord2base.push_back(v.base);
                put(-2);
                put(static_cast<int64_t>(v.base->dtype()));
                put(v.base->nelem());
            }
            put(o);
            // A sliding view's start changes every repeat; it is a launch argument, not structure.
            put(v.hasSlide() ? -3 : v.start);
            put(v.ndim);
            for (int64_t d = 0; d < v.ndim; ++d) {
                put(v.shape[d]);
                put(v.stride[d]);
            }
        }
    }
    uint64_t h = 0xcbf29ce484222325ull;
    for (size_t i = 1; i < words.size(); ++i) h = (h ^ static_cast<uint64_t>(words[i])) * 0x100000001b3ull;
    words[0] = static_cast<int64_t>(h ^ (h >> 29));
    return std::string(reinterpret_cast<const char *>(words.data()), words.size() * sizeof(int64_t));
}

void EngineHip::resolveProfile() {
    for (PendingEvent &p : pending_events) {
        float ms = 0;
        bhhip_event_elapsed_ms(p.start, p.stop, &ms);
        bhhip_event_destroy(p.start);
        bhhip_event_destroy(p.stop);
        p.rec.device_ms = ms;
        stats.device_ms_by_strategy[p.rec.strategy] += ms;
        last_launches.push_back(p.rec);
    }
    pending_events.clear();
}

std::vector<uint64_t> EngineHip::quietToken() const {
    return {pool.misses, pool.frees, stats.h2d_bytes, stats.d2h_bytes, scratch_allocs, last_index_checks ? 1u : 0u,
            flush_cache_misses, batch_launches};
}

void *EngineHip::graphFor(const std::vector<IssuedLaunch> &launches) {
    std::string key;
    for (const IssuedLaunch &l : launches) {
        key.append(reinterpret_cast<const char *>(&l.function), sizeof(l.function));
        key.append(reinterpret_cast<const char *>(&l.grid), sizeof(l.grid));
        key.append(reinterpret_cast<const char *>(&l.block), sizeof(l.block));
        const uint64_t n = l.args.size();
        key.append(reinterpret_cast<const char *>(&n), sizeof(n));
        key.append(reinterpret_cast<const char *>(l.args.data()), l.args.size());
    }
    auto it = graph_cache.find(key);
    if (it != graph_cache.end()) return it->second;
    if (graph_cache.size() >= 64) {
        check(bhhip_synchronize(), "synchronize");
        for (auto &kv : graph_cache) bhhip_graph_destroy(kv.second);
        graph_cache.clear();
    }
    check(bhhip_capture_begin(), "graph capture");
    for (const IssuedLaunch &l : launches) {
        if (bhhip_launch(l.function, l.grid, l.block, l.args.data(), l.args.size()) != 0) {
            void *discard = nullptr;
            bhhip_capture_end(&discard);
            throw std::runtime_error(std::string("hip engine: launch during graph capture failed: ") + bhhip_last_error());
        }
    }
    void *exec = nullptr;
    check(bhhip_capture_end(&exec), "graph instantiation");
    graph_cache[key] = exec;
    return exec;
}

void EngineHip::graphLaunch(void *graph, size_t nlaunches, uint64_t iterations) {
    check(bhhip_graph_launch(graph), "graph launch");
    graph_replays += iterations;
    stats.launches += nlaunches;
    stats.flushes += iterations;
}

void EngineHip::checkErrors() {
    if (dry_run) return;
    errors_unchecked = false;
    unsigned int flag = 0;
    check(bhhip_memcpy_d2h(&flag, counters + 1, sizeof(flag)), "reading the error flag");
    if (flag != 0) {
        const unsigned int zero = 0;
        check(bhhip_memcpy_h2d(counters + 1, &zero, sizeof(zero)), "clearing the error flag");
        if (flag & 1u) throw std::out_of_range("hip engine: gather/scatter index out of bounds");
        if (flag & 4u) {
            // the barrier counters of that launch may be left mid-count: no more persistent launches
            persistent = false;
            throw std::runtime_error("hip engine: a workgroup of a persistent launch gave up waiting at a grid "
                                     "barrier (its results are wrong; persistent launches are now off)");
        }
        throw std::runtime_error("hip engine: a scan tile gave up waiting for the prefix of an earlier tile "
                                 "(its accumulate results are wrong)");
    }
}

void EngineHip::handleExecution(BhIR *bhir) {
    const auto t_total = std::chrono::steady_clock::now();
    ++stats.flushes;
    released_now.clear();
    flush_written.clear();
    flush_exchanges.clear();
    std::set<bh_base *> frees;
    std::vector<bh_instruction *> instr_list = jitk::remove_non_computed_system_instr(bhir->instr_list, frees);
    for (bh_base *b : frees) release(b);
    rewrite_store.clear();
    if (unflatten) unflatten_full_reductions(instr_list, rewrite_store);
    std::vector<TwinUpdate> exchanges;
    if (twin_updates) {
        std::set<bh_base *> syncs;
        for (bh_base *b : bhir->getSyncs()) syncs.insert(b);
        auto twin_for = [&](bh_base *g) -> bh_base * {
            if (static_cast<uint64_t>(g->nbytes()) < twin_min_bytes) return nullptr;
            auto b = bufs.find(g);
            if (b != bufs.end() && b->second.external) return nullptr;
            Twin &tw = twins[g];
            if (!tw.base) tw.base.reset(new bh_base(g->nelem(), g->dtype()));
            return tw.base.get();
        };
        auto in_step = [&](bh_base *g) {
            auto tw = twins.find(g);
            auto b = bufs.find(g);
            return tw != twins.end() && b != bufs.end() && tw->second.in_step_at == b->second.writes;
        };
        exchanges = rewrite_in_place_updates(instr_list, rewrite_store, syncs, twin_for, in_step);
        for (const TwinUpdate &u : exchanges) twin_full_copies += u.full_copy ? 1 : 0;
    }
    jitk::set_constructor_flag(instr_list);
    stats.time_prepare += seconds_since(t_total);

    auto t = std::chrono::steady_clock::now();
    std::vector<bh_base *> ord2base;
    const std::string key = flushKey(instr_list, ord2base);
    stats.time_key += seconds_since(t);
    t = std::chrono::steady_clock::now();  // time_fuse: the lookup, and on a miss the fusion
    auto hit = flush_cache.find(key);
    static const bool debug = std::getenv("BH_HIP_DEBUG") != nullptr;
    if (hit == flush_cache.end()) {
        ++flush_cache_misses;
        std::unordered_map<const bh_base *, int> base2ord;  // the flush's base ordinals (inverse of ord2base)
        for (size_t o = 0; o < ord2base.size(); ++o) base2ord.emplace(ord2base[o], static_cast<int>(o));
        std::vector<jitk::LoopB> kernels = fold_serial_tails(jitk::get_kernel_list(instr_list, fusion_config));
        // A kernel releases the bases freed in it when it ends: no later kernel may use them.
        {
            std::set<const bh_base *> released;
            for (const jitk::LoopB &k : kernels) {
                for (const bh_base *b : k.getAllBases())
                    if (released.count(b))
                        throw std::logic_error("hip engine: a kernel uses a base that an earlier kernel frees:\n" +
                                               k.pprint());
                for (bh_base *b : k.getAllFrees()) released.insert(b);
            }
        }
        stats.time_fuse += seconds_since(t);
        t = std::chrono::steady_clock::now();
        CachedFlush cf;
        for (const jitk::LoopB &kernel : kernels) {
            if (debug) std::fprintf(stderr, "[hip] kernel:\n%s", kernel.pprint().c_str());
            CachedKernel ck;
            if (!kernel.isSystemOnly()) {
                const std::string bad = verify_kernel(kernel);
                if (!bad.empty())
                    throw std::logic_error("hip engine: kernel loops do not match its views: " + bad + "\n" +
                                           kernel.pprint());
                const std::vector<float> feats = block_features(jitk::allInstr(kernel));
                // Ordinals, features and hashes of a kernel's plans (under any knob set).
                auto finish = [&](std::vector<KernelPlan> &plans) {
                    for (KernelPlan &p : plans) {
                        p.features = feats;
                        for (const ArgSpec &a : p.args)
                            if (a.view_origin >= static_cast<int64_t>(instr_list.size())) p.reusable = false;
                        cf.reusable &= p.reusable;
                        for (bh_base *b : p.params) p.param_ordinals.push_back(base2ord.at(b));
                        for (bh_base *b : p.written) p.written_ordinals.push_back(base2ord.at(b));
                        for (ArgSpec &a : p.args)
                            if (a.kind == ArgSpec::BASE_PTR) a.base_ordinal = base2ord.at(a.base);
                        p.source_hash = fnv1a(p.source);
                        cf.index_checks |= p.uses_index_check;
                        if (debug)
                            std::fprintf(stderr, "[hip] plan %s grid %llu\n%s\n", strategy_name(p.strategy),
                                         static_cast<unsigned long long>(p.grid),
                                         std::getenv("BH_HIP_DEBUG")[0] == '2' ? p.source.c_str() : "");
                    }
                };
                ck.plans = plan_kernel(kernel, wcfg);
                finish(ck.plans);
                if (tuner.enabled) tuneKernel(ck, kernel, finish);
            }
            for (bh_base *b : kernel.getAllFrees()) ck.free_ordinals.push_back(base2ord.at(b));
            cf.kernels.push_back(std::move(ck));
        }
        stats.time_codegen += seconds_since(t);
        if (flush_cache.size() > 4096) flush_cache.clear();
        if (!cf.reusable) flush_cache.erase(key);
        hit = flush_cache.emplace(cf.reusable ? key : key + "#once", std::move(cf)).first;
    } else {
        ++flush_cache_hits;
        stats.time_fuse += seconds_since(t);
    }
    CachedFlush &cf = hit->second;
    if (!cf.schedule_built) buildSchedule(cf);
    static const char *dump_dir = std::getenv("BH_HIP_DUMP_DIR");
    if (cf.schedule && dry_run && dump_dir != nullptr && *dump_dir) dumpSchedule(cf, ord2base, instr_list, dump_dir);
    if (cf.schedule && !dry_run) {
        runSchedule(cf, ord2base, instr_list);
    } else {
        for (CachedKernel &ck : cf.kernels) {
            ++stats.kernels;
            ++kernel_seq;
            if (ck.trial && !settleTrial(ck)) {
                runTrial(ck, ord2base, instr_list);
            } else {
                for (KernelPlan &p : ck.plans) launchPlan(p, ord2base, instr_list);
            }
            for (int o : ck.free_ordinals) release(ord2base[o]);
        }
    }
    const auto t_finish = std::chrono::steady_clock::now();
    // The updated arrays take the twins' buffers (their new values); the twins keep the old ones.
    for (const TwinUpdate &u : exchanges) {
        exchangeTwin(u);
        release(u.temp);
    }
    // The bridge deletes a freed base after this flush and may reuse its address for a new base;
    // a buffer still filed under that address would then be handed to the new base.
    for (const bh_instruction &i : bhir->instr_list) {
        if (i.opcode != BH_FREE || released_now.count(i.operand[0].base)) continue;
        ++missed_frees;
        if (debug) std::fprintf(stderr, "[hip] BH_FREE of a%p was not released by its kernel\n",
                                static_cast<void *>(i.operand[0].base));
        release(i.operand[0].base);
    }
    const bool index_checks = cf.index_checks;
    last_index_checks = index_checks;
    if (!cf.reusable) flush_cache.erase(hit);
    if (index_checks) checkErrors();
    for (bh_base *b : bhir->getSyncs()) {
        copyToHost(b);
        if (!resident) {
            auto it = bufs.find(b);
            if (it != bufs.end()) eraseBuffer(it);
        } else {
            handOutHost(b);
        }
    }
    stats.time_finish += seconds_since(t_finish);
    stats.time_total += seconds_since(t_total);
}

// ------------------------------------------------------------------ autotuner

EngineHip::Trial::~Trial() {
    for (Pending &p : pending) {
        bhhip_event_destroy(p.ev0);
        bhhip_event_destroy(p.ev1);
    }
}

void EngineHip::tuneKernel(CachedKernel &ck, const jitk::LoopB &kernel,
                           const std::function<void(std::vector<KernelPlan> &)> &finish) {
    const std::string key = tuner.signature(ck.plans);
    if (key.empty()) return;
    const LaunchKnobs configured = LaunchKnobs::of(wcfg);
    if (const Tuner::Entry *e = tuner.lookup(key)) {
        ++tuner.table_hits;
        if (!(e->knobs == configured)) {
            std::vector<KernelPlan> plans = plan_kernel(kernel, e->knobs.apply(wcfg));
            finish(plans);
            ck.plans = std::move(plans);
        }
        return;
    }
    const KernelPlan *big = &ck.plans.front();
    for (const KernelPlan &p : ck.plans)
        if (p.work_items > big->work_items) big = &p;
    auto t = std::make_shared<Trial>();
    t->key = key;
    t->knobs.push_back(configured);
    t->variants.push_back(ck.plans);
    for (const LaunchKnobs &k : tuner.candidates(big->strategy, configured)) {
        if (k == configured) continue;
        std::vector<KernelPlan> plans = plan_kernel(kernel, k.apply(wcfg));
        finish(plans);
        bool dup = false;
        for (const std::vector<KernelPlan> &v : t->variants) dup |= same_launches(v, plans);
        if (dup) continue;
        t->knobs.push_back(k);
        t->variants.push_back(std::move(plans));
    }
    if (t->variants.size() < 2) return;
    // Block + launch features of each candidate, and the launch-shape model's predictions.
    const std::vector<float> block = ck.plans.front().features;
    for (size_t c = 0; c < t->variants.size(); ++c) {
        const LaunchKnobs &k = t->knobs[c];
        const int per_thread = big->strategy == Strategy::MAP           ? k.map_unroll
                               : big->strategy == Strategy::FULL_REDUCE ? k.reduce_unroll
                               : big->strategy == Strategy::SCAN        ? k.scan_items
                                                                        : 1;
        std::vector<float> row = block;
        const std::vector<float> lf = launch_features(t->variants[c], per_thread);
        row.insert(row.end(), lf.begin(), lf.end());
        t->features.push_back(std::move(row));
    }
    if (!predict_launch_times(t->features, t->predicted_ms)) t->predicted_ms.clear();
    if (tuner.mode == Tuner::Mode::MODEL && !t->predicted_ms.empty()) {
        size_t pick = 0;
        for (size_t c = 1; c < t->predicted_ms.size(); ++c)
            if (t->predicted_ms[c] < t->predicted_ms[pick]) pick = c;
        ck.plans = t->variants[pick];
        ++tuner.model_picks;
        return;
    }
    if (tuner.mode == Tuner::Mode::PRUNE && !t->predicted_ms.empty() &&
        t->variants.size() > static_cast<size_t>(tuner.prune_keep) + 1) {
        // The configured knobs and the model's best-predicted candidates go on trial; the others not.
        std::vector<size_t> order;
        for (size_t c = 1; c < t->variants.size(); ++c) order.push_back(c);
        std::stable_sort(order.begin(), order.end(),
                         [&](size_t a, size_t b) { return t->predicted_ms[a] < t->predicted_ms[b]; });
        std::vector<size_t> keep = {0};
        keep.insert(keep.end(), order.begin(), order.begin() + tuner.prune_keep);
        std::sort(keep.begin(), keep.end());
        auto pick = [&](auto &v) {
            std::remove_reference_t<decltype(v)> kept;
            for (size_t c : keep) kept.push_back(std::move(v[c]));
            v.swap(kept);
        };
        tuner.pruned += t->variants.size() - keep.size();
        pick(t->knobs);
        pick(t->variants);
        pick(t->features);
        pick(t->predicted_ms);
    }
    t->ms.resize(t->variants.size());
    t->inflight.assign(t->variants.size(), 0);
    ck.trial = std::move(t);
}

void EngineHip::pollTrial(Trial &t, bool wait) {
    size_t keep = 0;
    for (size_t i = 0; i < t.pending.size(); ++i) {
        Trial::Pending &p = t.pending[i];
        int done = 1;
        if (!wait) check(bhhip_event_done(p.ev1, &done), "event query");
        if (!done) {
            t.pending[keep++] = p;
            continue;
        }
        float ms = 0;
        check(bhhip_event_elapsed_ms(p.ev0, p.ev1, &ms), "event timing");
        t.ms[p.cand].push_back(ms);
        if (tuner.samples_log.size() < tuner.log_limit && p.cand < t.features.size())
            tuner.samples_log.push_back({t.features[p.cand], static_cast<double>(ms)});
        --t.inflight[p.cand];
        bhhip_event_destroy(p.ev0);
        bhhip_event_destroy(p.ev1);
    }
    t.pending.resize(keep);
}

// Adopt the table's choice when another flush settled the signature meanwhile, else the fastest
// candidate once every one has its samples.  True when the kernel's plans are final.
bool EngineHip::settleTrial(CachedKernel &ck) {
    Trial &t = *ck.trial;
    pollTrial(t, false);
    if (const Tuner::Entry *e = tuner.lookup(t.key)) {
        size_t pick = 0;
        for (size_t c = 0; c < t.knobs.size(); ++c)
            if (t.knobs[c] == e->knobs) pick = c;
        ck.plans = t.variants[pick];
        ck.trial.reset();
        return true;
    }
    for (size_t c = 0; c < t.ms.size(); ++c)
        if (static_cast<int>(t.ms[c].size()) < tuner.samples) return false;
    std::vector<double> best(t.ms.size());
    for (size_t c = 0; c < t.ms.size(); ++c) best[c] = *std::min_element(t.ms[c].begin(), t.ms[c].end());
    size_t pick = 0;
    for (size_t c = 1; c < best.size(); ++c)
        if (best[c] < best[pick]) pick = c;
    if (pick != 0 && best[pick] > best[0] * (1.0 - tuner.min_gain)) pick = 0;
    if (pick == 0) ++tuner.kept_default;
    if (tuner.outcomes_log.size() < tuner.log_limit) tuner.outcomes_log.push_back({t.key, best, t.predicted_ms, pick});
    ++tuner.decisions;
    tuner.saved_ms += best[0] - best[pick];
    Tuner::Entry e;
    e.knobs = t.knobs[pick];
    e.best_ms = best[pick];
    e.default_ms = best[0];
    tuner.store(t.key, e);
    ck.plans = t.variants[pick];
    ck.trial.reset();
    return true;
}

void EngineHip::runTrial(CachedKernel &ck, const std::vector<bh_base *> &ord2base,
                         const std::vector<bh_instruction *> &instrs) {
    Trial &t = *ck.trial;
    size_t c = 0;
    for (size_t i = 1; i < t.variants.size(); ++i)
        if (t.ms[i].size() + t.inflight[i] < t.ms[c].size() + t.inflight[c]) c = i;
    // Compile and make the operands resident before the first event: the events time the device's
    // timeline, so host work between them (a hiprtc compile, an upload) would count as kernel time.
    for (KernelPlan &plan : t.variants[c]) {
        if (plan.function == nullptr) compilePlan(plan);
        for (int o : plan.param_ordinals) ensureResident(ord2base[o]);
        if (plan.strategy == Strategy::FULL_REDUCE)
            ensureScratch(reduce_scratch, reduce_scratch_bytes, plan.scratch_bytes, false);
        if (plan.strategy == Strategy::SCAN) ensureScratch(scan_scratch, scan_scratch_bytes, plan.scratch_bytes, true);
    }
    Trial::Pending p{c, nullptr, nullptr};
    check(bhhip_event_create(&p.ev0), "event creation");
    check(bhhip_event_create(&p.ev1), "event creation");
    check(bhhip_event_record(p.ev0), "event record");
    for (KernelPlan &plan : t.variants[c]) launchPlan(plan, ord2base, instrs);
    check(bhhip_event_record(p.ev1), "event record");
    t.pending.push_back(p);
    ++t.inflight[c];
    ++tuner.trials;
}

void EngineHip::resolveTuning() {
    if (dry_run) return;
    for (auto &kv : flush_cache)
        for (CachedKernel &ck : kv.second.kernels)
            if (ck.trial) {
                pollTrial(*ck.trial, true);
                settleTrial(ck);
            }
}

void EngineHip::handleExtmethod(BhIR *bhir) {
    // Extension methods have the opcodes after the built-in ones: none, nothing to split.
    bool any = false;
    for (const bh_instruction &instr : bhir->instr_list) any |= instr.opcode > BH_MAX_OPCODE_ID;
    if (!any) return;
    std::vector<bh_instruction> instr_list;
    for (bh_instruction &instr : bhir->instr_list) {
        auto ext = comp.extmethods.find(instr.opcode);
        if (ext != comp.extmethods.end()) {
            BhIR b(std::move(instr_list), bhir->getSyncs());
            handleExecution(&b);
            instr_list.clear();
            for (const bh_view &v : instr.operand)
                if (!v.isConstant()) ensureResident(v.base);
            bh_hip_ext_context ctx;
            ctx.version = 1;
            ctx.dry_run = dry_run ? 1 : 0;
            ctx.engine = this;
            ctx.device_pointer = [](void *engine, const void *base) -> void * {
                EngineHip *e = static_cast<EngineHip *>(engine);
                bh_base *b = const_cast<bh_base *>(static_cast<const bh_base *>(base));
                return e->getMemoryPointer(*b, false, false, false);
            };
            ctx.stream = dry_run ? nullptr : bhhip_stream();
            ext->second.execute(&instr, &ctx);
            // An extension method may write any of its operands (blas_trmm/blas_trsm write B, not C):
            // the device copies of all of them are now the current ones.
            for (const bh_view &v : instr.operand) {
                if (v.isConstant()) continue;
                noteWrite(v.base, buffer(v.base));
            }
        } else if (comp.child_extmethods.count(instr.opcode)) {
            throw std::runtime_error("hip engine: extension method is only known to a child component");
        } else {
            instr_list.push_back(std::move(instr));
        }
    }
    bhir->instr_list = std::move(instr_list);
}

void *EngineHip::getMemoryPointer(bh_base &base, bool copy2host, bool force_alloc, bool nullify) {
    bh_base *b = &base;
    if (!copy2host) {
        ensureResident(b);
        if (dry_run) return nullptr;
        return bufs.at(b).ptr;
    }
    copyToHost(b);
    if (force_alloc && b->getDataPtr() == nullptr) {
        bh_data_malloc(b);
        buffer(b);
    }
    void *ret = b->getDataPtr();
    auto it = bufs.find(b);
    if (!resident && it != bufs.end()) {
        eraseBuffer(it);
        it = bufs.end();
    }
    if (nullify) {
        // The caller takes the host buffer over: no guard may stay on its pages.
        if (it != bufs.end()) {
            unguardHost(it->second);
            it->second.host_valid = false;
        }
        b->resetDataPtr(nullptr);
    } else if (it != bufs.end()) {
        handOutHost(b);
    }
    return ret;
}

namespace {
// "option: 1, 2 3" inside a "; "-separated parameter string -> {1, 2, 3}; {} when absent or malformed.
std::vector<uint64_t> param_integer_list(const std::string &option, const std::string &param) {
    size_t at = 0;
    while ((at = param.find(option, at)) != std::string::npos) {
        const bool starts = at == 0 || param[at - 1] == ' ' || param[at - 1] == ';' || param[at - 1] == '\t';
        size_t p = at + option.size();
        while (p < param.size() && (param[p] == ' ' || param[p] == '\t')) ++p;
        if (starts && p < param.size() && param[p] == ':') {
            std::vector<uint64_t> ret;
            ++p;
            while (p < param.size() && param[p] != ';') {
                const char c = param[p];
                if (c == ' ' || c == ',' || c == '\t') {
                    ++p;
                } else if (c >= '0' && c <= '9') {
                    uint64_t v = 0;
                    while (p < param.size() && param[p] >= '0' && param[p] <= '9') v = v * 10 + (param[p++] - '0');
                    ret.push_back(v);
                } else {
                    return {};
                }
            }
            return ret;
        }
        at += option.size();
    }
    return {};
}
} // namespace

std::string EngineHip::userKernel(const std::string &kernel, std::vector<bh_view> &operands, const std::string &param) {
    for (const bh_view &v : operands)
        if (v.isConstant()) return "[HIP] userKernel: operands cannot be constants";
    const std::vector<uint64_t> gws = param_integer_list("global_work_size", param);
    const std::vector<uint64_t> lws = param_integer_list("local_work_size", param);
    if (gws.size() != lws.size())
        return "[HIP] userKernel-param dimension of global_work_size and local_work_size must be the same";
    if (gws.empty() || gws.size() > 3)
        return "[HIP] userKernel-param needs global_work_size and local_work_size of one to three dimensions";
    uint32_t grid[3] = {1, 1, 1}, block[3] = {1, 1, 1};
    uint64_t threads = 1;
    for (size_t d = 0; d < gws.size(); ++d) {
        if (lws[d] == 0 || lws[d] > 1024) return "[HIP] userKernel-param local_work_size must be in 1..1024";
        const uint64_t g = (gws[d] + lws[d] - 1) / lws[d];
        if (g > 0x7fffffffull) return "[HIP] userKernel-param global_work_size too large";
        grid[d] = static_cast<uint32_t>(g);
        block[d] = static_cast<uint32_t>(lws[d]);
        threads *= lws[d];
    }
    if (threads > 1024) return "[HIP] userKernel-param local_work_size has more than 1024 threads";
    const auto t = std::chrono::steady_clock::now();
    if (dry_run) {
        if (bhhip_compile("", kernel.c_str(), "execute", cache_dir.c_str(), nullptr) != 0)
            return std::string("[HIP] ") + bhhip_last_error();
        stats.time_compile += seconds_since(t);
        return "";
    }
    void *fn = nullptr;
    if (bhhip_compile("", kernel.c_str(), "execute", cache_dir.c_str(), &fn) != 0)
        return std::string("[HIP] ") + bhhip_last_error();
    stats.time_compile += seconds_since(t);
    std::vector<void *> ptrs;
    for (const bh_view &v : operands) {
        ensureResident(v.base);
        ptrs.push_back(bufs.at(v.base).ptr);
    }
    void *none = nullptr;
    if (bhhip_launch_3d(fn, grid, block, ptrs.empty() ? static_cast<const void *>(&none) : ptrs.data(),
                        ptrs.size() * sizeof(void *)) != 0)
        return std::string("[HIP] ") + bhhip_last_error();
    ++stats.launches;
    ++stats.launches_by_strategy["user"];
    // The kernel may write any operand.
    for (const bh_view &v : operands) markDeviceWritten(v.base);
    return "";
}

void EngineHip::markHostWritten(bh_base *base) {
    auto it = bufs.find(base);
    if (it == bufs.end() || base->getDataPtr() == nullptr) return;
    unguardHost(it->second);
    it->second.host_valid = true;
    it->second.dev_valid = false;
}

void EngineHip::markDeviceWritten(bh_base *base) {
    DeviceBuffer &d = buffer(base);
    d.dev_valid = true;
    d.host_valid = false;
    ++d.writes;
}

void EngineHip::setMemoryPointer(bh_base *base, bool host_ptr, void *mem) {
    DeviceBuffer &d = buffer(base);
    if (host_ptr) {
        unguardHost(d);
        if (base->getDataPtr() != nullptr && base->getDataPtr() != mem) bh_data_free(base);
        base->resetDataPtr(mem);
        d.host_valid = mem != nullptr;
        d.dev_valid = false;
    } else {
        if (d.ptr && !d.external) pool.release(d.ptr, d.bytes);
        d.ptr = mem;
        d.bytes = static_cast<uint64_t>(base->nbytes());
        d.external = true;
        d.dev_valid = true;
        d.host_valid = false;
        ++d.writes;
    }
}

std::string EngineHip::info() const {
    std::ostringstream ss;
    ss << "----\nHIP vector engine:\n  Device: " << bhhip_device_info() << "\n  Kernel cache: "
       << (cache_dir.empty() ? "memory only" : cache_dir) << "\n  Fuser list:";
    for (const std::string &f : fusion_config.fuser_list) ss << " " << f;
    ss << "\n  Resident arrays: " << (resident ? "yes" : "no") << "\n";
    return ss.str();
}

std::string EngineHip::statistics() const {
    uint64_t compiles = 0, mhits = 0, dhits = 0;
    bhhip_compile_stats(&compiles, &mhits, &dhits);
    std::ostringstream ss;
    ss << std::fixed << std::setprecision(6);
    ss << "flushes: " << stats.flushes << "\nkernels: " << stats.kernels << "\nlaunches: " << stats.launches << "\n";
    for (const auto &kv : stats.launches_by_strategy) ss << "launches_" << kv.first << ": " << kv.second << "\n";
    ss << "compiles: " << compiles << "\ncompile_memory_hits: " << mhits << "\ncompile_disk_hits: " << dhits << "\n";
    ss << "h2d_bytes: " << stats.h2d_bytes << "\nd2h_bytes: " << stats.d2h_bytes << "\n";
    ss << "time_fuse: " << stats.time_fuse << "\ntime_codegen: " << stats.time_codegen
       << "\ntime_compile: " << stats.time_compile << "\ntime_launch: " << stats.time_launch
       << "\ntime_copy: " << stats.time_copy << "\ntime_total: " << stats.time_total
       << "\ntime_prepare: " << stats.time_prepare << "\ntime_key: " << stats.time_key
       << "\ntime_finish: " << stats.time_finish << "\n";
    for (const auto &kv : stats.device_ms_by_strategy) ss << "device_ms_" << kv.first << ": " << kv.second << "\n";
    ss << "flush_cache_hits: " << flush_cache_hits << "\nflush_cache_misses: " << flush_cache_misses << "\n";
    ss << "missed_frees: " << missed_frees << "\n";
    ss << "buffers: " << bufs.size() << "\nreleases: " << releases << "\n";
    ss << "host_guards: " << host_guards << "\nhost_write_faults: " << g_host_write_faults.load() << "\n";
    ss << "graph_replays: " << graph_replays << "\ngraphs_cached: " << graph_cache.size() << "\n";
    ss << "graph_declined_changed: " << graph_declined_changed << "\ngraph_declined_busy: " << graph_declined_busy << "\n";
    ss << "batch_launches: " << batch_launches << "\nbatched_items: " << batched_items << "\n";
    ss << "persist_launches: " << persist_launches << "\npersist_fallbacks: " << persist_fallbacks << "\n";
    ss << "twin_exchanges: " << twin_exchanges << "\ntwin_full_copies: " << twin_full_copies << "\n";
    ss << "autotune_model_picks: " << tuner.model_picks << "\nautotune_pruned: " << tuner.pruned << "\n";
    ss << "autotune_trials: " << tuner.trials << "\nautotune_decisions: " << tuner.decisions
       << "\nautotune_kept_configured: " << tuner.kept_default << "\nautotune_table_hits: " << tuner.table_hits
       << "\nautotune_saved_ms: " << tuner.saved_ms << "\n";
    ss << "pool_allocated: " << pool.allocated << "\npool_peak: " << pool.peak << "\npool_hits: " << pool.hits
       << "\npool_misses: " << pool.misses << "\n";
    return ss.str();
}

// ------------------------------------------------------------------ component

namespace {

// Advance views that slide between repeats (reference include/bohrium/jitk/engines/dyn_view.hpp:21-67).
void slide_views(BhIR *bhir) {
    for (bh_instruction &instr : bhir->instr_list) {
        for (bh_view &view : instr.operand) {
            if (!view.hasSlide()) continue;
            for (const bh_slide_dim &dim : view.slides.dims) {
                if (dim.step_delay == 1 || (view.slides.iteration_counter % dim.step_delay == dim.step_delay - 1)) {
                    if (dim.stride) {
                        int64_t change = dim.offset_change * dim.stride;
                        const int64_t max_rel = dim.stride * dim.shape;
                        int64_t rel = view.start % (dim.stride * dim.shape);
                        rel += change;
                        if (rel < 0) {
                            change += max_rel;
                        } else if (rel >= max_rel) {
                            change -= max_rel;
                        }
                        view.start += change;
                        auto it = view.slides.resets.find(dim.rank);
                        if (it != view.slides.resets.end()) {
                            const int64_t reset_at = it->second.first;
                            int64_t &since = it->second.second;
                            since += change;
                            if (view.slides.iteration_counter > 0 &&
                                (view.slides.iteration_counter / dim.step_delay) % reset_at == reset_at - 1) {
                                view.start -= since;
                                since = 0;
                                view.shape[dim.rank] -= reset_at * dim.shape_change;
                            }
                        }
                    }
                    view.shape[dim.rank] += dim.shape_change;
                    if (view.shape[dim.rank] < 0) view.shape[dim.rank] = 0;
                }
            }
            view.slides.iteration_counter += 1;
        }
    }
}

// Periods of a replayed repeat body captured in one graph.
constexpr uint64_t kGraphRounds = 8;

class Impl : public component::ComponentVE {
public:
    EngineHip engine;

    explicit Impl(int stack_level) : ComponentVE(stack_level), engine(*this) {}
    ~Impl() override = default;

    static bool has_slides(const BhIR &bhir) {
        for (const bh_instruction &i : bhir.instr_list)
            for (const bh_view &v : i.operand)
                if (!v.isConstant() && v.hasSlide()) return true;
        return false;
    }

    bool has_extmethods(const BhIR &bhir) const {
        for (const bh_instruction &i : bhir.instr_list)
            if (extmethods.count(i.opcode) || child_extmethods.count(i.opcode)) return true;
        return false;
    }

    void execute(BhIR *bhir) override {
        if (disabled) {
            child.execute(bhir);
            return;
        }
        bh_base *cond = bhir->getRepeatCondition();
        const uint64_t nrepeats = bhir->getNRepeats();
        // A repeat body that only launches kernels is captured as a graph and replayed for the
        // remaining iterations: once two successive iterations launched the same sequence (period
        // 1), or once an iteration repeats the one before the previous (period 2: an array updated
        // in place through its twin alternates between two buffers).  A replayed iteration also
        // repeats what its flush did to engine state besides launching (write counts, twin
        // exchanges).
        const bool with_ext = has_extmethods(*bhir);
        const bool graphable = engine.graph_repeats && !engine.profile && !engine.dry_run && cond == nullptr &&
                               nrepeats >= 4 && !has_slides(*bhir) && !with_ext;
        struct Iteration {
            std::vector<EngineHip::IssuedLaunch> launches;
            bool quiet;
            EngineHip::FlushEffects effects;
        };
        std::deque<Iteration> hist;
        std::vector<EngineHip::IssuedLaunch> phase_launches[2];
        std::vector<EngineHip::IssuedLaunch> cur;
        void *graphs[2] = {nullptr, nullptr};
        size_t graph_launches[2] = {0, 0};
        EngineHip::FlushEffects graph_effects[2];
        int period = 0, phase = 0;
        // Several periods in one graph: fewer graph launches (each costs a submission and a gap on
        // the device); `rounds` periods per launch of `multi`.
        void *multi = nullptr;
        uint64_t rounds = 0;
        size_t multi_launches = 0;
        for (uint64_t i = 0; i < nrepeats; ++i) {
            if (period > 0) {
                if (multi == nullptr && phase == 0 && nrepeats - i >= 2 * static_cast<uint64_t>(period)) {
                    rounds = std::min<uint64_t>(kGraphRounds, (nrepeats - i) / static_cast<uint64_t>(period));
                    std::vector<EngineHip::IssuedLaunch> seq;
                    for (uint64_t r = 0; r < rounds; ++r)
                        for (int ph = 0; ph < period; ++ph) {
                            const std::vector<EngineHip::IssuedLaunch> &l = phase_launches[ph];
                            seq.insert(seq.end(), l.begin(), l.end());
                        }
                    multi = engine.graphFor(seq);
                    multi_launches = seq.size();
                }
                if (multi != nullptr && phase == 0 && nrepeats - i >= rounds * static_cast<uint64_t>(period)) {
                    engine.graphLaunch(multi, multi_launches, rounds * static_cast<uint64_t>(period));
                    for (uint64_t r = 0; r < rounds; ++r)
                        for (int ph = 0; ph < period; ++ph) engine.applyEffects(graph_effects[ph]);
                    i += rounds * static_cast<uint64_t>(period) - 1;
                    continue;
                }
                engine.graphLaunch(graphs[phase], graph_launches[phase]);
                engine.applyEffects(graph_effects[phase]);
                phase = (phase + 1) % period;
                continue;
            }
            const bool rec = graphable && i >= 1;
            const std::vector<uint64_t> q0 = engine.quietToken();
            cur.clear();
            if (rec) engine.recording = &cur;
            // Running extension methods takes them out of the list: a later iteration runs them again.
            std::vector<bh_instruction> body;
            if (with_ext && i + 1 < nrepeats) body = bhir->instr_list;
            try {
                engine.handleExtmethod(bhir);
                engine.handleExecution(bhir);
            } catch (...) {
                engine.recording = nullptr;
                throw;
            }
            engine.recording = nullptr;
            if (with_ext && i + 1 < nrepeats) bhir->instr_list = std::move(body);
            EngineHip::FlushEffects fx = engine.takeEffects();
            if (rec) {
                hist.push_back({cur, engine.quietToken() == q0, std::move(fx)});
                if (hist.size() > 3) hist.pop_front();
                const size_t n = hist.size();
                const Iteration &last = hist[n - 1];
                if (n >= 2 && !last.launches.empty()) {
                    if (last.launches == hist[n - 2].launches) {
                        if (!last.quiet) {
                            ++engine.graph_declined_busy;
                        } else {
                            graphs[0] = engine.graphFor(last.launches);
                            graph_launches[0] = last.launches.size();
                            graph_effects[0] = last.effects;
                            phase_launches[0] = last.launches;
                            period = 1;
                        }
                    } else if (n >= 3 && last.launches == hist[n - 3].launches) {
                        if (!last.quiet || !hist[n - 2].quiet) {
                            ++engine.graph_declined_busy;
                        } else {
                            // the next iteration repeats hist[n - 2], the one after it `last`
                            graphs[0] = engine.graphFor(hist[n - 2].launches);
                            graph_launches[0] = hist[n - 2].launches.size();
                            graph_effects[0] = hist[n - 2].effects;
                            graphs[1] = engine.graphFor(last.launches);
                            graph_launches[1] = last.launches.size();
                            graph_effects[1] = last.effects;
                            phase_launches[0] = hist[n - 2].launches;
                            phase_launches[1] = last.launches;
                            period = 2;
                        }
                    } else {
                        ++engine.graph_declined_changed;
                    }
                }
            }
            if (cond != nullptr) {
                engine.copyToHost(cond);
                if (cond->getDataPtr() != nullptr && !static_cast<bool *>(cond->getDataPtr())[0]) break;
            }
            slide_views(bhir);
        }
    }

    void extmethod(const std::string &name, bh_opcode opcode) override {
        try {
            // Library paths with a directory part are relative to the config file, like `impl`.
            std::vector<std::string> libs = config.defaultGetList("libs", {});
            for (std::string &l : libs)
                if (!l.empty() && l[0] != '/' && l.find('/') != std::string::npos) l = config.file_dir + "/" + l;
            extmethods.insert(std::make_pair(opcode, extmethod::ExtmethodFace(libs, name)));
        } catch (const std::out_of_range &) {
            if (!child.initiated()) throw;
            child.extmethod(name, opcode);
            child_extmethods.insert(opcode);
        }
    }

    std::string message(const std::string &msg) override {
        std::ostringstream ss;
        if (msg == "statistic_enable_and_reset") {
            engine.resetStatistics();
            engine.profile = true;
        } else if (msg == "statistic") {
            engine.resolveProfile();
            ss << engine.statistics();
        } else if (msg == "info") {
            ss << engine.info();
        } else if (msg == "GPU: disable") {
            if (!child.initiated()) throw std::runtime_error("hip engine: cannot disable the GPU without a child engine");
            engine.copyAllToHost();
            disabled = true;
        } else if (msg == "GPU: enable") {
            disabled = false;
        } else if (msg == "profile: on") {
            engine.profile = true;
        } else if (msg == "profile: off") {
            engine.profile = false;
        } else if (msg.rfind("codegen:", 0) == 0) {
            // "codegen: max_grid=G map_unroll=U reduce_unroll=V" (any subset); drops cached plans.
            std::istringstream in(msg.substr(8));
            std::string kv;
            while (in >> kv) {
                const size_t eq = kv.find('=');
                if (eq == std::string::npos) throw std::invalid_argument("hip engine: bad codegen setting: " + kv);
                const std::string k = kv.substr(0, eq);
                const int v = std::stoi(kv.substr(eq + 1));
                if (v < 0 || (v < 1 && k != "row_reduce" && k != "waves_per_eu" && k != "resident_grid" &&
                              k != "math64" && k != "xcd_remap" && k != "xcd_parts"))
                    throw std::invalid_argument("hip engine: codegen setting must be >= 1: " + kv);
                if (k == "max_grid") {
                    engine.codegenConfig().max_grid = v;
                } else if (k == "map_unroll") {
                    engine.codegenConfig().map_unroll = v;
                } else if (k == "reduce_unroll") {
                    engine.codegenConfig().reduce_unroll = v;
                } else if (k == "resident_grid") {
                    engine.codegenConfig().resident_grid = v != 0;
                } else if (k == "math64") {
                    engine.codegenConfig().math64 = v != 0;
                } else if (k == "xcd_remap") {
                    engine.codegenConfig().xcd_remap = v != 0;
                } else if (k == "xcd_parts") {
                    engine.codegenConfig().xcd_parts = v != 0;
                } else if (k == "reduce_max_grid") {
                    engine.codegenConfig().reduce_max_grid = v;
                } else if (k == "reduce_min_items") {
                    engine.codegenConfig().reduce_min_items = v;
                } else if (k == "scan_items") {
                    engine.codegenConfig().scan_items = v;
                } else if (k == "row_reduce") {
                    engine.codegenConfig().row_reduce = v != 0;
                } else if (k == "waves_per_eu") {
                    engine.codegenConfig().waves_per_eu = v;
                } else {
                    throw std::invalid_argument("hip engine: unknown codegen setting: " + k);
                }
            }
            const WriterConfig &w = engine.codegenConfig();
            ss << "max_grid=" << w.max_grid << " map_unroll=" << w.map_unroll << " reduce_unroll=" << w.reduce_unroll
               << " reduce_max_grid=" << w.reduce_max_grid << " reduce_min_items=" << w.reduce_min_items
               << " scan_items=" << w.scan_items << " row_reduce=" << (w.row_reduce ? 1 : 0)
               << " waves_per_eu=" << w.waves_per_eu << " resident_grid=" << (w.resident_grid ? 1 : 0)
               << " math64=" << (w.math64 ? 1 : 0) << " xcd_remap=" << (w.xcd_remap ? 1 : 0)
               << " xcd_parts=" << (w.xcd_parts ? 1 : 0);
        } else if (msg.rfind("batch_min:", 0) == 0) {
            // Smallest group of same-code launches run as one batched launch (0: never); drops cached plans.
            engine.setBatchMin(std::stoi(msg.substr(10)));
        } else if (msg.rfind("batch_scan_group:", 0) == 0) {
            // Batched scans: items whose tiles are interleaved (1: each item's tiles in a row).
            engine.batch_scan_group = std::max(1, std::stoi(msg.substr(17)));
            engine.dropPlans();
        } else if (msg == "batch_persistent: on" || msg == "batch_persistent: off") {
            // Batched schedules as one persistent launch, or one launch per level and group.
            engine.persistent = msg == "batch_persistent: on";
        } else if (msg.rfind("persist_spin:", 0) == 0) {
            // Polls of a persistent launch's grid-barrier wait before it raises the error flag.
            engine.persist_spin = static_cast<unsigned>(std::stoul(msg.substr(13)));
        } else if (msg == "batch_persistent") {
            ss << (engine.persistent ? "on" : "off");
        } else if (msg.rfind("persist_items:", 0) == 0 || msg.rfind("persist_grid:", 0) == 0) {
            // Elements per task of grid-stride plans, or the cap on the persistent grid (0: none);
            // drops cached plans (and with them their persistent plans).
            const bool items = msg.rfind("persist_items:", 0) == 0;
            const int v = std::max(0, std::stoi(msg.substr(items ? 14 : 13)));
            (items ? engine.persist_items : engine.persist_grid) = v;
            engine.dropPlans();
        } else if (msg.rfind("batch_xcds:", 0) == 0) {
            // Batched scans: XCDs the workgroups are dealt to round robin (1: no XCD placement).
            engine.batch_xcds = std::max(1, std::stoi(msg.substr(11)));
            engine.dropPlans();
        } else if (msg == "plans: drop") {
            engine.dropPlans();
        } else if (msg == "autotune: on" || msg == "autotune: off") {
            // Tuning applies to kernels planned from here on; cached plans are dropped.
            engine.tuner.enabled = msg == "autotune: on" && !engine.dry_run;
            engine.dropPlans();
        } else if (msg == "autotune: forget") {
            // Forget the in-memory table (its file is kept) and every cached plan.
            engine.tuner.clear();
            engine.dropPlans();
        } else if (msg.rfind("autotune_samples:", 0) == 0) {
            engine.tuner.samples = std::max(1, std::stoi(msg.substr(17)));
        } else if (msg.rfind("autotune_min_items:", 0) == 0) {
            engine.tuner.min_items = std::stoull(msg.substr(19));
            engine.dropPlans();
        } else if (msg == "autotune_mode: measure" || msg == "autotune_mode: model" || msg == "autotune_mode: prune") {
            engine.tuner.mode = msg == "autotune_mode: model"   ? Tuner::Mode::MODEL
                                : msg == "autotune_mode: prune" ? Tuner::Mode::PRUNE
                                                                : Tuner::Mode::MEASURE;
            engine.dropPlans();
        } else if (msg == "autotune samples") {
            // Timed trial launches since the last read: "f=<block and launch features> ms=<device ms>".
            for (const Tuner::Sample &x : engine.tuner.samples_log) {
                ss << "f=";
                for (size_t k = 0; k < x.features.size(); ++k) ss << (k ? "," : "") << x.features[k];
                ss << " ms=" << x.ms << "\n";
            }
            engine.tuner.samples_log.clear();
        } else if (msg == "autotune outcomes") {
            // Settled trials since the last read: every candidate's best time, the model's predictions.
            for (const Tuner::Outcome &o : engine.tuner.outcomes_log) {
                ss << "key=" << o.key << " chosen=" << o.chosen << " ms=";
                for (size_t k = 0; k < o.best_ms.size(); ++k) ss << (k ? "," : "") << o.best_ms[k];
                ss << " predicted_ms=";
                if (o.predicted_ms.empty()) ss << "none";
                for (size_t k = 0; k < o.predicted_ms.size(); ++k) ss << (k ? "," : "") << o.predicted_ms[k] * 1e-3;
                ss << "\n";
            }
            engine.tuner.outcomes_log.clear();
        } else if (msg == "autotune") {
            // Settle what can be settled (synchronizes), then list the table.
            engine.resolveTuning();
            ss << engine.tuner.dump();
        } else if (msg == "host_write_tracking: on") {
            engine.track_host_writes = true;
        } else if (msg == "host_write_tracking: off") {
            engine.track_host_writes = false;
        } else if (msg == "graph: on") {
            engine.graph_repeats = true;
        } else if (msg == "graph: off") {
            engine.graph_repeats = false;
        } else if (msg == "graph replays") {
            ss << engine.graph_replays;
        } else if (msg == "plan log") {
            for (const std::string &l : engine.dry_run_log) ss << l << "\n";
            engine.dry_run_log.clear();
        } else if (msg == "launches") {
            engine.resolveProfile();
            for (const auto &r : engine.last_launches)
                ss << r.strategy << " " << r.work_items << " " << r.grid << " " << r.device_ms << " " << r.kernel_hash
                   << " k=" << r.kernel_id << " f=" << r.features << "\n";
            engine.last_launches.clear();
        } else if (msg.rfind("fuser_list:", 0) == 0) {
            std::vector<std::string> list;
            std::istringstream in(msg.substr(11));
            std::string item;
            while (std::getline(in, item, ',')) {
                item.erase(0, item.find_first_not_of(' '));
                item.erase(item.find_last_not_of(' ') + 1);
                if (!item.empty()) list.push_back(item);
            }
            engine.fusion_config.fuser_list = list;
            engine.dropPlans();
        }
        if (child.initiated()) ss << child.message(msg);
        return ss.str();
    }

    void *getMemoryPointer(bh_base &base, bool copy2host, bool force_alloc, bool nullify) override {
        return engine.getMemoryPointer(base, copy2host, force_alloc, nullify);
    }

    void setMemoryPointer(bh_base *base, bool host_ptr, void *mem) override {
        engine.setMemoryPointer(base, host_ptr, mem);
    }

    void memCopy(bh_view &src, bh_view &dst, const std::string &) override {
        // A contiguous whole-base copy between two bases.
        if (src.base->nbytes() != dst.base->nbytes()) throw std::runtime_error("memCopy: size mismatch");
        void *s = engine.getMemoryPointer(*src.base, false, false, false);
        void *d = engine.getMemoryPointer(*dst.base, false, false, false);
        check(bhhip_memcpy_d2d(d, s, static_cast<size_t>(src.base->nbytes())), "memCopy");
        engine.markDeviceWritten(dst.base);
    }

    void *getDeviceContext() override { return nullptr; }
    void setDeviceContext(void *) override {}

    std::string userKernel(const std::string &kernel, std::vector<bh_view> &operand_list, const std::string &compile_cmd,
                           const std::string &tag, const std::string &param) override {
        if (tag != "hip") {
            if (!child.initiated()) throw std::runtime_error("No backend with tag \"" + tag + "\" found");
            for (const bh_view &v : operand_list) engine.copyToHost(v.base);
            const std::string ret = child.userKernel(kernel, operand_list, compile_cmd, tag, param);
            // The child may have written any operand's host data.
            for (const bh_view &v : operand_list) engine.markHostWritten(v.base);
            return ret;
        }
        return engine.userKernel(kernel, operand_list, param);
    }
};

} // namespace

} // namespace vehip
} // namespace bohrium

extern "C" bohrium::component::ComponentImpl *create(int stack_level) {
    return new bohrium::vehip::Impl(stack_level);
}

extern "C" void destroy(bohrium::component::ComponentImpl *self) { delete self; }
