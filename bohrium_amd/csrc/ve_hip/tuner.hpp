/* Autotuner of the hip vector engine: launch shapes per block signature, found by measurement.
 *
 * The reference engines launch every block with fixed work-group shapes from the config
 * (config.ini.in:204-209, ve/cuda/engine_cuda.cpp:40-45).  Here the shape of a fused kernel is a
 * set of skeleton knobs (workgroups of the grid-stride skeletons, elements per thread issued
 * together, elements per thread of a scan tile) and the best set differs per kernel: a stencil
 * streams best from fewer, longer-running workgroups, a transcendental-heavy map from more.
 *
 * A kernel's signature is the skeleton and source hash of each of its launches plus log2 of its
 * iteration space (sources do not depend on sizes).  The first time a signature with at least
 * `min_items` elements is executed, the engine plans the kernel under every candidate knob set,
 * drops candidates whose launches come out identical, and runs the distinct ones in turn on the
 * live flushes (round robin, `samples` launches each, device time from HIP events read without
 * waiting).  The fastest candidate is kept when it beats the configured knobs by `min_gain`;
 * the choice goes into a table keyed by signature, kept in memory and appended to a file in the
 * kernel cache directory, so later flushes and later processes launch the chosen shape at once.
 */
#pragma once

#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "codegen.hpp"

namespace bohrium {
namespace vehip {

/* The skeleton knobs the tuner varies (a subset of WriterConfig). */
struct LaunchKnobs {
    int max_grid = 2048;
    int reduce_max_grid = 4096;
    int map_unroll = 1;
    int reduce_unroll = 1;
    int scan_items = 8;

    static LaunchKnobs of(const WriterConfig &c);
    WriterConfig apply(WriterConfig c) const;
    bool operator==(const LaunchKnobs &o) const;
    std::string str() const;
};

class Tuner {
public:
    bool enabled = true;
    /* measure: trials decide (the default); model: the launch-shape model (bhcm_install_launch)
     * picks the candidate with the lowest predicted time and no trial runs; prune: trials of the
     * configured knobs and the model's `prune_keep` best-predicted candidates only (model and prune
     * measure every candidate when no model is installed). */
    enum class Mode { MEASURE, MODEL, PRUNE } mode = Mode::MEASURE;
    int prune_keep = 2;
    int samples = 3;                  // timed launches per candidate
    uint64_t min_items = 1u << 20;    // smaller kernels keep the configured knobs
    double min_gain = 0.02;           // a candidate must beat the configured knobs by this fraction
    std::string path;                 // table file ("" = memory only)

    struct Entry {
        LaunchKnobs knobs;
        double best_ms = 0, default_ms = 0;
    };

    /* Signature of a kernel from its (default-knob) launch plans; "" when it is not tuned. */
    std::string signature(const std::vector<KernelPlan> &plans) const;
    /* Candidate knob sets for a kernel whose largest launch uses skeleton `s` (the configured
     * set first). */
    std::vector<LaunchKnobs> candidates(Strategy s, const LaunchKnobs &base) const;
    const Entry *lookup(const std::string &key) const;
    void store(const std::string &key, const Entry &e);
    /* Read `path` (later lines of one key win); missing file: empty table. */
    void load();
    void clear() { table.clear(); }
    std::string dump() const;

    uint64_t trials = 0, decisions = 0, table_hits = 0, kept_default = 0, model_picks = 0, pruned = 0;
    double saved_ms = 0;  // sum over decisions of (configured - chosen) per-launch device time

    /* Training data of the launch-shape model: block + launch features of each timed trial launch
     * and its device time; and per settled trial, every candidate's best time and (when a launch
     * model is installed) its prediction.  Kept up to `log_limit` entries until read. */
    struct Sample {
        std::vector<float> features;
        double ms;
    };
    struct Outcome {
        std::string key;
        std::vector<double> best_ms, predicted_ms;  // per candidate ([0]: configured knobs)
        size_t chosen = 0;
    };
    std::vector<Sample> samples_log;
    std::vector<Outcome> outcomes_log;
    size_t log_limit = 200000;

private:
    std::map<std::string, Entry> table;
};

/* The launch plans of one kernel are the same shape under two knob sets (same sources, grids and
 * workgroup sizes): trying both would measure the same launches twice. */
bool same_launches(const std::vector<KernelPlan> &a, const std::vector<KernelPlan> &b);

} // namespace vehip
} // namespace bohrium
