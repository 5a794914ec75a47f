/* Instruction list -> kernel list.
 *
 * The pipeline and the names of its stages follow the reference's jitk
 * (core/jitk/apply_fusion.cpp:48-259, fuser.cpp:35-392, graph.cpp:52-365,
 * transformer.cpp:145-246): simplify, pre-fuse ("lossy" or "singleton"), apply
 * the configured transformer list ("greedy", "serial", "push_reductions_inwards",
 * "collapse_redundant_axes", plus this engine's "learned"), then insert the sweep
 * identities and split into kernels.
 */
#pragma once

#include <functional>
#include <string>
#include <vector>
#include <bohrium/jitk/block.hpp>

namespace bohrium {
namespace jitk {

/* Benefit of merging block `a` into `b` (larger is merged first).  The
 * analytic default is the reference's `weight()`: bytes of arrays that become
 * temporaries (graph.cpp:136-150). */
typedef std::function<double(const Block &a, const Block &b)> MergeWeight;

struct FusionConfig {
    std::string pre_fuser = "lossy";
    std::vector<std::string> fuser_list = {"greedy", "collapse_redundant_axes"};
    uint64_t greedy_threshold = 10000;
    bool avoid_rank0_sweep = false;
    /* Used by the "learned" transformer; when empty it falls back to the analytic weight. */
    MergeWeight learned_weight;
    /* Used by the "learned" fuser when set: the predicted device time of a block made of each
     * given instruction set (false: no model, fall back to the analytic greedy fuser). */
    std::function<bool(const std::vector<std::vector<InstrPtr> > &, std::vector<double> &)> learned_cost;
    /* Engine-specific veto on merging two blocks (e.g. keep root blocks on a parallel skeleton). */
    std::function<bool(const LoopB &, const LoopB &)> merge_filter;
    /* Keep an accumulate out of the pre-fuser's groups that hold any other sweep (a GPU scan
     * skeleton cannot also reduce); the greedy fuser applies merge_filter instead. */
    bool accumulate_alone = false;
    /* Keep sweeps along different axes of one shape in different pre-fuser groups (a loop nest can
     * make only one of them its inner reduction). */
    bool one_sweep_axis = false;
    bool can_merge(const Block &a, const Block &b) const {
        if (!mergeable(a, b, avoid_rank0_sweep)) return false;
        return !merge_filter || a.isInstr() || b.isInstr() || merge_filter(a.getLoop(), b.getLoop());
    }
};

std::vector<InstrPtr> simplify_instr_list(const std::vector<bh_instruction *> &instr_list);
void simplify_instr(bh_instruction &instr);
std::vector<Block> pre_fuser_lossy(const std::vector<bh_instruction *> &instr_list, bool accumulate_alone = false,
                                   bool one_sweep_axis = false);
std::vector<Block> fuser_singleton(const std::vector<bh_instruction *> &instr_list);
void fuser_serial(std::vector<Block> &block_list, const FusionConfig &config);
void fuser_greedy(const FusionConfig &config, std::vector<Block> &block_list, const MergeWeight &weight_fn);
/* Greedy fusion scored by a cost model: repeatedly merge the pair whose merge is predicted to
 * save the most time, cost(a) + cost(b) - cost(a u b), while that saving is positive.  Returns
 * false (block_list untouched) when config.learned_cost has no model. */
bool fuser_greedy_learned(const FusionConfig &config, std::vector<Block> &block_list);
void push_reductions_inwards(std::vector<Block> &block_list);
void collapse_redundant_axes(std::vector<Block> &block_list);

/* The analytic costs of the reference fuser. */
uint64_t weight(const Block &b1, const Block &b2);
uint64_t block_cost(const Block &block);

std::vector<Block> get_block_list(const std::vector<bh_instruction *> &instr_list, const FusionConfig &config);
std::vector<LoopB> get_kernel_list(const std::vector<bh_instruction *> &instr_list, const FusionConfig &config);
/* Turn an already fused block list into kernels (identity instructions inserted before every sweep). */
std::vector<LoopB> kernels_from_blocks(std::vector<Block> block_list, int64_t &origin_count);

std::vector<bh_instruction *> remove_non_computed_system_instr(std::vector<bh_instruction> &instr_list,
                                                              std::set<bh_base *> &frees);
void set_constructor_flag(std::vector<bh_instruction *> &instr_list);

} // namespace jitk
} // namespace bohrium
