/* The hip vector engine: BhIR -> fused kernels -> hiprtc -> launches on one
 * HIP stream, with arrays kept resident in device memory between flushes.
 *
 * Per-flush flow mirrors the reference GPU engine's handleExecution
 * (include/bohrium/jitk/engines/engine_gpu.hpp:81-160): drop non-computing
 * system instructions and free what they release, set constructor flags, fuse
 * into kernels, run each kernel, release arrays freed inside it, copy sync'ed
 * arrays to the host.  Differences by design: every kernel runs on the device
 * (no CPU offload), launches are stream-ordered without a per-launch
 * synchronize, and a sync leaves the device copy valid.
 */
#pragma once

#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <set>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>
#include <bohrium/bh_component.hpp>
#include <bohrium/jitk/fuser.hpp>
#include "codegen.hpp"
#include "rewrite.hpp"
#include "features.hpp"
#include "tuner.hpp"

namespace bohrium {
namespace vehip {

struct DeviceBuffer {
    void *ptr = nullptr;
    uint64_t bytes = 0;
    bool dev_valid = false;   // device holds the newest data
    bool host_valid = false;  // host buffer holds the newest data
    bool external = false;    // device memory supplied by the caller (not pooled)
    /* Host pages made read-only after the host copy was handed out while the device copy stayed
     * current; the first host write faults, and the fault handler makes the host copy the newest. */
    void *guard_addr = nullptr;
    uint64_t guard_bytes = 0;
    /* Counts every change of the device copy's contents (launch, upload, exchange with a twin). */
    uint64_t writes = 0;
};

/* Size-keyed reuse of device allocations (hipMalloc is far slower than a lookup). */
class DevicePool {
    /* Released buffers by exact size, the last released at the back (constant time per release and
     * reuse: a flush of thousands of temporaries releases and takes thousands of buffers). */
    std::unordered_map<uint64_t, std::vector<void *> > _free;
    uint64_t _cached = 0;

public:
    uint64_t limit = 64ull << 30;
    uint64_t allocated = 0, peak = 0, hits = 0, misses = 0, frees = 0;
    void *alloc(uint64_t bytes);
    void release(void *ptr, uint64_t bytes);
    void trim(uint64_t keep);
    ~DevicePool();
};

struct EngineStats {
    uint64_t flushes = 0, kernels = 0, launches = 0, h2d_bytes = 0, d2h_bytes = 0;
    std::map<std::string, uint64_t> launches_by_strategy;
    double time_fuse = 0, time_codegen = 0, time_compile = 0, time_launch = 0, time_copy = 0, time_total = 0;
    /* Host phases of a flush: system instructions, frees and rewrites before the plan lookup; the
     * lookup key; after the launches (releases, twin exchanges, error check, syncs). */
    double time_prepare = 0, time_key = 0, time_finish = 0;
    /* Per-launch device time (ms) keyed by strategy, recorded when profiling is enabled. */
    std::map<std::string, double> device_ms_by_strategy;
};

class EngineHip {
public:
    explicit EngineHip(component::ComponentVE &comp);
    ~EngineHip();

    void handleExecution(BhIR *bhir);
    void handleExtmethod(BhIR *bhir);
    void *getMemoryPointer(bh_base &base, bool copy2host, bool force_alloc, bool nullify);
    void setMemoryPointer(bh_base *base, bool host_ptr, void *mem);
    /* The device buffer of `base` now holds its newest data (written outside a kernel launch). */
    void markDeviceWritten(bh_base *base);
    /* The host data of `base` now holds its newest data (written by a child component). */
    void markHostWritten(bh_base *base);
    void copyAllToHost();
    /* Compile the user's `extern "C" __global__ void execute(T0 *a0, ...)` with hiprtc and launch it on the
     * operands' device buffers with the shape in `param` ("global_work_size: X [Y [Z]]; local_work_size: ...").
     * Returns "" or an error message, like the reference's OpenCL engine
     * (ve/opencl/engine_opencl.cpp:611-669). */
    std::string userKernel(const std::string &kernel, std::vector<bh_view> &operands, const std::string &param);
    void copyToHost(bh_base *base);
    /* The caller receives the host copy of `base` (a sync, or getMemoryPointer with copy2host) and may
     * write it.  With host_write_tracking the device copy stays current and the host pages become
     * read-only until the first write; without it (or for host memory the engine cannot protect) the
     * host copy becomes the newest at once, as in the reference GPU engines (ve/cuda/main.cpp:91-108). */
    void handOutHost(bh_base *base);
    bool track_host_writes = true;
    std::string info() const;
    std::string statistics() const;
    void resetStatistics() { stats = EngineStats(); }

    /* Last flush's launches, for cost-model training and tests: (strategy, work items, device ms). */
    struct LaunchRecord {
        std::string strategy;
        uint64_t work_items;
        uint64_t grid;
        double device_ms;
        std::string kernel_hash;
        uint64_t kernel_id = 0;   // executed jitk kernel this launch belongs to
        std::string features;     // its cost-model features (comma separated)
    };
    std::vector<LaunchRecord> last_launches;
    /* Resolve the events of profiled launches into last_launches (synchronizes). */
    void resolveProfile();

    /* One launch as issued: kernel, shape and the packed argument bytes. */
    struct IssuedLaunch {
        void *function = nullptr;
        uint64_t grid = 0;
        uint32_t block = 0;
        std::vector<uint8_t> args;
        bool operator==(const IssuedLaunch &o) const {
            return function == o.function && grid == o.grid && block == o.block && args == o.args;
        }
    };
    /* When set, every launch is also appended here (used to find repeat bodies that can be graph-replayed). */
    std::vector<IssuedLaunch> *recording = nullptr;
    /* Changes whenever the engine allocates device memory, copies between host and device or checks
     * index errors: a repeat iteration whose token is unchanged did nothing but launch kernels. */
    std::vector<uint64_t> quietToken() const;
    /* A graph holding `launches` in order (captured once per distinct sequence, then cached). */
    void *graphFor(const std::vector<IssuedLaunch> &launches);
    /* Launch a captured graph that stands for `iterations` repeat iterations (for the statistics). */
    void graphLaunch(void *graph, size_t nlaunches, uint64_t iterations = 1);
    bool graph_repeats = true;
    uint64_t graph_replays = 0;
    uint64_t kernel_seq = 0;  // id of the jitk kernel whose plans are being launched
    /* Batched launches: groups of at least `batch_min` launches of one kernel code (0 disables). */
    int batch_min = 8;
    /* Batched scans: items whose tiles are interleaved with each other (1: each item's tiles in a row). */
    int batch_scan_group = 8;
    /* Batched scans: XCDs the workgroups go to round robin (all tiles of an item on one XCD, whose L2
     * then serves the item's gathers: a quarter of the HBM reads on the batched-arrays workload, at
     * the same time; 1: no placement). */
    int batch_xcds = 8;
    uint64_t batch_launches = 0, batched_items = 0;
    void setBatchMin(int n) {
        batch_min = n;
        flush_cache.clear();
    }
    /* Forget every cached flush plan (after a change of fusion settings or of the cost model). */
    void dropPlans() { flush_cache.clear(); }
    /* Repeat iterations not captured: launches differed from the previous iteration / the engine also
     * allocated, copied or checked errors. */
    uint64_t graph_declined_changed = 0, graph_declined_busy = 0;
    /* Kernel-writer settings; changing them through this reference drops every cached plan. */
    WriterConfig &codegenConfig() {
        flush_cache.clear();
        return wcfg;
    }

    /* What a flush did to engine state besides its launches; a replayed graph of that flush
     * repeats it (bases written: their write counts; twin exchanges). */
    struct FlushEffects {
        std::vector<bh_base *> written;
        std::vector<TwinUpdate> exchanges;
    };
    FlushEffects takeEffects() {
        FlushEffects f;
        f.written.assign(flush_written.begin(), flush_written.end());
        f.exchanges.swap(flush_exchanges);
        flush_written.clear();
        return f;
    }
    void applyEffects(const FlushEffects &f);
    /* In-place updates through a temporary run as buffer exchanges with a twin (rewrite.hpp). */
    bool twin_updates = true;
    uint64_t twin_min_bytes = 1u << 20;
    uint64_t twin_exchanges = 0, twin_full_copies = 0;

    /* Launch shapes per kernel signature (see tuner.hpp). */
    Tuner tuner;
    /* Read every outstanding tuning measurement (synchronizes) and settle finished trials. */
    void resolveTuning();

    jitk::FusionConfig fusion_config;
    bool profile = false;
    /* Compile-only mode: plan and hiprtc-compile every kernel, touch no device. */
    bool dry_run = false;
    std::vector<std::string> dry_run_log;  // "strategy grid block hash" per planned launch

private:
    component::ComponentVE &comp;
    WriterConfig wcfg;
    bool unflatten = true;       // restate full reductions of N-D temporaries on their shape
    RewriteStore rewrite_store;  // instructions and temporaries added to the current flush
    std::string cache_dir;
    bool resident = true;
    int num_cus = 0;  // compute units of the device (0 in compile-only mode)
    std::unordered_map<const bh_base *, DeviceBuffer> bufs;
    /* The second buffer of an array updated in place (rewrite_in_place_updates), and the write
     * count of the array at which the twin last equalled it outside the updated view. */
    struct Twin {
        std::unique_ptr<bh_base> base;
        uint64_t in_step_at = ~0ull;
    };
    std::unordered_map<const bh_base *, Twin> twins;
    std::set<bh_base *> flush_written;
    std::vector<TwinUpdate> flush_exchanges;
    void exchangeTwin(const TwinUpdate &u);
    void noteWrite(bh_base *base, DeviceBuffer &d) {
        d.dev_valid = true;
        d.host_valid = false;
        ++d.writes;
        if (recording != nullptr) flush_written.insert(base);  // only a recorded flush's effects are kept
    }
    /* Bases released during the current flush; a BH_FREE that no kernel released is caught with it. */
    std::unordered_set<const bh_base *> released_now;
    DevicePool pool;
    unsigned char *reduce_scratch = nullptr;
    uint64_t reduce_scratch_bytes = 0;
    unsigned char *scan_scratch = nullptr;
    uint64_t scan_scratch_bytes = 0;
    unsigned int *counters = nullptr;  // [0]: reduce ticket, [1]: error flag, [16..63]: zeros (stand-in base)
    EngineStats stats;
    /* Autotuning of one kernel in progress: its plans under each distinct candidate knob set,
     * device times measured so far, and launches whose events are not read yet. */
    struct Trial {
        std::string key;
        std::vector<LaunchKnobs> knobs;  // [0]: the configured knobs
        std::vector<std::vector<KernelPlan> > variants;
        std::vector<std::vector<float> > features;  // block + launch features per candidate
        std::vector<double> predicted_ms;           // the launch-shape model's, when installed
        std::vector<std::vector<double> > ms;
        std::vector<int> inflight;
        struct Pending {
            size_t cand;
            void *ev0, *ev1;
        };
        std::vector<Pending> pending;
        ~Trial();
    };
    /* Plans of an already seen flush, keyed by its canonical encoding. */
    struct CachedKernel {
        std::vector<KernelPlan> plans;
        std::vector<int> free_ordinals;
        std::shared_ptr<Trial> trial;  // set while the kernel's launch shape is being tuned
    };
    /* Many launches of one kernel code in a flush run as one batched launch (bh_kernel_batch):
     * each item keeps its own packed arguments and workgroup count.  Launches are ordered in
     * levels; a launch's level is one more than that of every earlier launch it conflicts with
     * (reads what that one writes, or writes what it reads or writes). */
    struct PlanRef {
        uint32_t kernel, plan;
    };
    struct BatchGroup {
        std::vector<PlanRef> items;
        void *fn = nullptr, *fin_fn = nullptr;
        unsigned char *dev = nullptr;  // packed arguments of every item, then first-workgroup table
        uint64_t dev_bytes = 0;
        uint8_t *host = nullptr;  // page-locked staging of the same
        uint64_t host_bytes = 0;
        void *uploaded = nullptr;  // event recorded after the last upload from `host`
        unsigned char *scan_scratch = nullptr;  // zeroed look-back state, one region per item (SCAN)
        uint64_t scan_bytes = 0;
        ~BatchGroup();
    };
    struct Level {
        std::vector<PlanRef> singles;
        std::vector<size_t> groups;
        std::vector<uint32_t> free_kernels;  // kernels whose frees are released after this level
    };
    /* The whole schedule as one cooperative launch of bh_kernel_persist (persist_source): the
     * levels' launches become tasks of a table, run level by level with a grid barrier between
     * levels (and between a level's full-reduction main bodies and their finalizes). */
    struct PersistPlan {
        std::vector<PlanRef> plans;         // every plan of the schedule (argument blocks in this order)
        std::vector<const KernelPlan *> codes;  // one plan per distinct kernel code
        std::vector<uint64_t> arg_off;      // per plan: byte offset of its packed arguments (16-aligned)
        uint64_t args_bytes = 0;
        std::vector<uint64_t> scratch_off;  // per plan: byte offset of its scratch region
        std::vector<uint32_t> nblk;         // per plan: its workgroups (tasks) in the persistent launch
        uint64_t scratch_bytes = 0;
        std::vector<uint32_t> tasks;        // 4 words per task: case, workgroup, workgroups, args / 16
        std::vector<uint32_t> levels;       // first task of each level, then the end
        uint32_t max_level_tasks = 0;
        void *fn = nullptr;
        uint32_t grid = 0;
        unsigned char *dev = nullptr;  // tasks | levels | barrier counters (16 bytes) | arguments
        uint64_t dev_bytes = 0, args_at = 0;
        bool tables_uploaded = false;
        uint8_t *host = nullptr;  // page-locked staging
        uint64_t host_bytes = 0;
        void *uploaded = nullptr;
        unsigned char *scratch = nullptr;  // zeroed once (scan look-back state persists across launches)
        ~PersistPlan();
    };
    struct Schedule {
        std::vector<Level> levels;
        std::vector<std::unique_ptr<BatchGroup> > groups;
        std::unique_ptr<PersistPlan> persist;
        bool persist_checked = false;  // persist built, or found not to apply
    };
    struct CachedFlush {
        std::vector<CachedKernel> kernels;
        bool index_checks = false;
        bool reusable = true;
        bool schedule_built = false;
        std::shared_ptr<Schedule> schedule;  // null: launch plan by plan
    };
    void buildSchedule(CachedFlush &cf);
    void runSchedule(CachedFlush &cf, const std::vector<bh_base *> &ord2base, const std::vector<bh_instruction *> &instrs);
    void dumpSchedule(CachedFlush &cf, const std::vector<bh_base *> &ord2base, const std::vector<bh_instruction *> &instrs,
                      const char *dir);
    void launchBatch(BatchGroup &g, CachedFlush &cf, const std::vector<bh_base *> &ord2base,
                     const std::vector<bh_instruction *> &instrs);
    /* Build the schedule's persistent plan (null when a plan does not fit the persistent kernel). */
    std::unique_ptr<PersistPlan> buildPersist(CachedFlush &cf);
    /* False when the cooperative launch could not be made (the caller runs the levels). */
    bool launchPersist(PersistPlan &pp, CachedFlush &cf, const std::vector<bh_base *> &ord2base,
                       const std::vector<bh_instruction *> &instrs);
public:
    /* Batched schedules run as one persistent launch (false, the default: one launch per level and
     * group, which measures faster on MI355X: the grid barriers' cache write-back and invalidation and
     * the task loop cost more than the launches they replace). */
    bool persistent = false;
    /* Grid-stride plans (map, full and row reductions) take one task per `persist_items` elements (at
     * most their own workgroup count): fewer, longer tasks than their standalone grid.  0: the plan's
     * own grid. */
    int persist_items = 16384;
    int persist_grid = 0;  // >0: at most this many workgroups in the persistent launch
    /* Polls of a grid barrier's wait before it gives up and raises the error flag (default 2^22,
     * seconds of waiting; a tiny value forces the error path from the host). */
    unsigned persist_spin = 1u << 22;
    uint64_t persist_launches = 0, persist_fallbacks = 0;
private:
    /* `grid_arg`: the value packed for a GRID_U32 argument (0: the plan's grid). */
    size_t packArgs(const KernelPlan &plan, const std::vector<bh_base *> &ord2base,
                    const std::vector<bh_instruction *> &instrs, void *scratch, uint8_t *out, size_t cap,
                    uint32_t grid_arg = 0);
    void compilePlan(KernelPlan &plan);
    void dumpLaunch(const char *dir, const KernelPlan &plan, const std::vector<bh_base *> &ord2base,
                    const std::vector<bh_instruction *> &instrs);
    /* Flush keys start with a 64-bit hash of the rest (flushKey): hashing reads those 8 bytes, not the
     * whole key (megabytes for flushes of tens of thousands of instructions); equality still compares
     * the whole key. */
    struct FlushKeyHash {
        size_t operator()(const std::string &k) const noexcept {
            uint64_t h = 0;
            if (k.size() >= sizeof h) std::memcpy(&h, k.data(), sizeof h);
            return static_cast<size_t>(h);
        }
    };
    std::unordered_map<std::string, CachedFlush, FlushKeyHash> flush_cache;
    uint64_t flush_cache_hits = 0, flush_cache_misses = 0;
    uint64_t missed_frees = 0;
    uint64_t releases = 0;  // bases released (BH_FREE handled) since the engine started
    uint64_t host_guards = 0;  // host copies handed out read-only
    bool last_index_checks = false;
    bool errors_unchecked = false;  // a launch since the last error-flag check can raise an error
    uint64_t scratch_allocs = 0;
    std::map<std::string, void *> graph_cache;
    std::string flushKey(const std::vector<bh_instruction *> &instrs, std::vector<bh_base *> &ord2base) const;
    void launchPlan(KernelPlan &plan, const std::vector<bh_base *> &ord2base,
                    const std::vector<bh_instruction *> &instrs);
    /* Autotuner: set up a trial for a kernel just planned (or take the table's choice), run one
     * launch of the next candidate, read finished measurements, and settle on a candidate. */
    void tuneKernel(CachedKernel &ck, const jitk::LoopB &kernel,
                    const std::function<void(std::vector<KernelPlan> &)> &finish);
    void runTrial(CachedKernel &ck, const std::vector<bh_base *> &ord2base,
                  const std::vector<bh_instruction *> &instrs);
    void pollTrial(Trial &t, bool wait);
    bool settleTrial(CachedKernel &ck);
    struct PendingEvent {
        void *start, *stop;
        LaunchRecord rec;
    };
    std::vector<PendingEvent> pending_events;

    DeviceBuffer &buffer(bh_base *base);
    void guardHost(bh_base *base, DeviceBuffer &d);
    void unguardHost(DeviceBuffer &d);
    void eraseBuffer(std::unordered_map<const bh_base *, DeviceBuffer>::iterator it);
    void ensureResident(bh_base *base);
    void release(bh_base *base);
    void ensureScratch(unsigned char *&ptr, uint64_t &have, uint64_t need, bool zero);

    void checkErrors();
};

} // namespace vehip
} // namespace bohrium
