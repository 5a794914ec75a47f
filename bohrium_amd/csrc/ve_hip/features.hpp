/* Features of a fused block (an instruction set) for the learned cost model.
 *
 * The model predicts the log of the block's device time from these values; the
 * engine records them with every profiled launch (the training data) and the
 * learned fuser computes them for each candidate merge.  All are functions of the
 * instructions alone, so a block and the union of two blocks are described the
 * same way.  A temporary is a base written first (constructor) and freed inside
 * the set; it costs no HBM traffic once the block is one kernel.
 */
#pragma once

#include <string>
#include <vector>

#include <bohrium/jitk/block.hpp>

namespace bohrium {
namespace vehip {

constexpr int kNumFeatures = 16;

/* Names in order (kept in sync with the Python side, bohrium_amd/costmodel.py). */
extern const char *const kFeatureNames[kNumFeatures];

std::vector<float> block_features(const std::vector<jitk::InstrPtr> &instrs);

/* Comma-separated rendering used in the engine's launch records. */
std::string features_string(const std::vector<float> &f);

/* Predicted device time (microseconds) of each instruction set with the installed cost model;
 * false when no model is installed. */
bool predict_block_times(const std::vector<std::vector<jitk::InstrPtr> > &sets, std::vector<double> &us);

/* Launch features of a kernel's plans under one knob set: log2 of the largest launch's workgroups,
 * log2 of the elements a thread of it issues together (unroll, or scan items), log2 of its
 * elements per thread (work items / (workgroups x workgroup size)). */
constexpr int kNumLaunchFeatures = 3;
std::vector<float> launch_features(const std::vector<struct KernelPlan> &plans, int unroll_or_items);

/* Predicted device time (microseconds) of each row of block features + launch features with the
 * installed launch-shape model; false when none is installed. */
bool predict_launch_times(const std::vector<std::vector<float> > &rows, std::vector<double> &us);

} // namespace vehip
} // namespace bohrium
