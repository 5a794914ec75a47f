/* HIP kernel writer for the hip vector engine.
 *
 * Each top-level block of a jitk kernel becomes one launch of one of these
 * skeletons, into which the block's scalar body is spliced:
 *   MAP          grid-stride loop over the flattened leading parallel ranks;
 *                inner ranks (and inner sweeps) run as sequential loops.
 *   FULL_REDUCE  rank-0 reductions: per-thread accumulators, wave64 butterfly,
 *                LDS across waves, per-workgroup partials; a one-workgroup
 *                finalize kernel folds them in workgroup order (deterministic)
 *                and runs the small scalar blocks that follow in thread 0.
 *   SCAN         rank-0 add/multiply accumulate: wave64 shuffle scan, LDS
 *                across waves and a decoupled look-back across workgroups.
 *   ROW_REDUCE   a map whose innermost parallel loop only seeds and reduces an inner
 *                loop (row sums of an expression): one wave per row, lanes stride the
 *                row (coalesced), wave64 butterfly, lane 0 stores.
 *   SERIAL       anything else, executed by one thread (always correct).
 * Per-opcode scalar semantics restate the reference's write_operation
 * (core/jitk/instruction.cpp:58-544) and engine.cpp:251-336 special forms.
 */
#pragma once

#include <cstdint>
#include <set>
#include <string>
#include <vector>
#include <bohrium/jitk/block.hpp>

namespace bohrium {
namespace vehip {

enum class Strategy { MAP, FULL_REDUCE, SCAN, SERIAL, ROW_REDUCE };
const char *strategy_name(Strategy s);

/* One kernel argument: either literal bytes known at plan time, or a pointer
 * the engine resolves at launch time. */
struct ArgSpec {
    /* GRID_U32: the launch's workgroup count as uint32 (the finalize kernel's partial count). */
    enum Kind { LITERAL, BASE_PTR, SCRATCH_PTR, COUNTER_PTR, ERRFLAG_PTR, GRID_U32 } kind = LITERAL;
    const bh_base *base = nullptr;
    int64_t scratch_offset = 0;  // byte offset into the engine's scratch area
    std::vector<uint8_t> bytes;  // LITERAL value
    int size = 8;
    int align = 8;
    /* A LITERAL that is an instruction's constant: the instruction's position in the
     * flushed list (origin id) and which part of an R123 constant (0 start, 1 key, -1 whole).
     * A cached plan re-reads these from the current flush. */
    int64_t const_origin = -1;
    int const_part = -1;
    /* BASE_PTR: the base's ordinal in the flush (first-appearance order), set when cached. */
    int base_ordinal = -1;
    /* A LITERAL that is the start of a sliding view: read from operand `view_operand` of the
     * instruction at position `view_origin` in the current flush. */
    int64_t view_origin = -1;
    int view_operand = -1;
};

struct KernelPlan {
    Strategy strategy = Strategy::SERIAL;
    std::string source;  // kernel source without the prelude
    std::string name;
    /* FULL_REDUCE: a second kernel of the same module, launched with one workgroup right
     * after `name` with the same arguments (folds the per-workgroup partials). */
    std::string fin_name;
    std::vector<ArgSpec> args;
    /* The body as device functions (bh_main_, and bh_fin_ for FULL_REDUCE) and their parameters,
     * one per ArgSpec: declaration, value type and name.  `source` adds the kernels calling them. */
    std::string device_code;
    std::string attributes;  // extra attributes of the main kernel (occupancy hint)
    std::vector<std::string> param_decls, param_types, param_names;
    uint64_t grid = 1;
    uint32_t block = 256;
    /* Bases that must be resident on the device for this launch, and those it writes. */
    std::vector<bh_base *> params;
    std::set<bh_base *> written;
    /* Scratch bytes (partials / look-back state) this launch needs. */
    int64_t scratch_bytes = 0;
    bool uses_index_check = false;  // gather/scatter: the flush checks the error flag when it ends
    bool uses_error_flag = false;   // takes bh_err (index checks, bounded look-back waits)
    bool uses_gather_flag = false;  // a gather marks bad indices in bh_bad_ (raised when the body ends)
    uint64_t work_items = 0;  // elements of the iteration space (for statistics / cost model features)
    /* Filled when the plan is cached: ordinals of params/written bases and the compiled function. */
    std::vector<int> param_ordinals, written_ordinals;
    void *function = nullptr;
    void *fin_function = nullptr;
    /* Cost-model features of the jitk kernel this launch belongs to (recorded with profiled launches). */
    std::vector<float> features;
    uint64_t source_hash = 0;
    /* False when a sliding view's start cannot be re-read from the flush (plan not reusable). */
    bool reusable = true;
};

/* Hooks the engine supplies to the writer. */
struct WriterConfig {
    int max_grid = 2048;     // workgroups for the MAP skeleton
    int reduce_max_grid = 4096;  // workgroups for FULL_REDUCE
    int reduce_min_items = 4;    // FULL_REDUCE: elements per thread before another workgroup is added
    int scan_items = 8;      // elements per thread per scan tile
    int map_unroll = 1;      // grid-stride elements per thread issued together (MAP)
    bool row_reduce = true;  // run row reductions with one wave per row (ROW_REDUCE)
    int waves_per_eu = 0;    // >0: amdgpu_waves_per_eu occupancy hint on the main kernels
    int reduce_unroll = 1;   // the same for FULL_REDUCE
    /* Grid-stride skeletons (MAP, FULL_REDUCE, ROW_REDUCE) launch at most the workgroups that are
     * resident on the device at once (CUs x occupancy of the compiled kernel). */
    bool resident_grid = false;
    /* float64 exp and log through the prelude's bh_exp64 / bh_log64 (false: the HIP math library's). */
    bool math64 = true;
    /* Grid-stride skeletons renumber their workgroups so that each XCD runs a contiguous range. */
    bool xcd_remap = false;
    /* Grid-stride skeletons of stencils (a base read through several views) walk the iteration space
     * in 8 contiguous parts, one per XCD (workgroup b runs on XCD b mod 8): a row's neighbours are
     * read under the same L2. */
    bool xcd_parts = true;
    uint32_t epoch = 1;      // look-back flag generation (changes every launch)
};

/* Append each kernel made only of small scalar blocks to the kernel before it when
 * that one ends in a FULL_REDUCE loop: its finalize kernel then runs them. */
std::vector<jitk::LoopB> fold_serial_tails(std::vector<jitk::LoopB> kernels);

/* Check that every operand dimension is iterated within its extent by the loop of the matching
 * rank (what the generated indexing assumes); returns "" or a description of each violation. */
std::string verify_kernel(const jitk::LoopB &kernel);

/* Split a jitk kernel (rank -1 LoopB) into launch plans. */
std::vector<KernelPlan> plan_kernel(const jitk::LoopB &kernel, const WriterConfig &cfg);

/* Which skeleton runs a rank-0 block (used by the fuser to keep blocks on parallel skeletons). */
Strategy classify_root_block(const jitk::LoopB &l0);

/* True when merging two root blocks keeps the result on a parallel skeleton. */
bool root_merge_allowed(const jitk::LoopB &l1, const jitk::LoopB &l2);

/* Byte offset of each argument in the packed kernel-argument buffer (natural alignment),
 * and the buffer's total size (a multiple of 8). */
std::vector<size_t> arg_offsets(const std::vector<ArgSpec> &args, size_t *total);

/* Source of the batched form of `plan`: bh_kernel_batch runs many items that share the plan's
 * code, each with its own packed arguments and workgroup count (and bh_kernel_fin_batch, one
 * workgroup per item, for FULL_REDUCE). */
std::string batch_source(const KernelPlan &plan);

/* Source of a persistent launch running many plans of a flush (codes: one plan per distinct kernel
 * code): bh_kernel_persist, launched cooperatively with 256-thread workgroups, runs a table of tasks
 * level by level with a grid barrier between levels (kernel_prelude bh_grid_sync). */
std::string persist_source(const std::vector<const KernelPlan *> &codes);

/* HIP spelling of an element type inside generated kernels. */
const char *hip_type(bh_type t);

} // namespace vehip
} // namespace bohrium
