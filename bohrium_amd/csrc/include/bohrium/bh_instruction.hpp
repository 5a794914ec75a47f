/* One bytecode instruction: opcode, operand views, an optional constant.
 *
 * Members and the shape/axis helpers mirror include/bohrium/bh_instruction.hpp:36-166
 * and core/bh_instruction.cpp:71-237 of the reference (principal shape, sweep axis,
 * reshape, remove_axis, transpose) because the fuser and the code generators
 * rely on exactly those meanings.
 */
#pragma once

#include <set>
#include <string>
#include <vector>
#include <bohrium/bh_opcode.h>
#include <bohrium/bh_view.hpp>
#include <bohrium/bh_constant.hpp>

struct bh_instruction {
    bh_opcode opcode = -1;
    std::vector<bh_view> operand;
    bh_constant constant;
    /* Set by the engine: this instruction is the first writer of its output
     * base in the current flush (array contraction may keep it in a register). */
    bool constructor = false;
    /* Position in the flushed instruction list; survives reshapes/transposes. */
    int64_t origin_id = -1;

    bh_instruction() = default;
    bh_instruction(bh_opcode opcode, std::vector<bh_view> operands) : opcode(opcode), operand(std::move(operands)) {}

    std::set<bh_base *> get_bases() const;
    std::vector<const bh_view *> getViews() const;
    std::vector<bh_view *> getViews();
    bool has_constant() const;
    bool isContiguous() const;
    bool all_same_shape() const;
    bool reshapable() const;
    /* Shape of the iteration space of this instruction (the input's shape for sweeps). */
    BhIntVec shape() const;
    int64_t ndim() const;
    /* Axis swept by a reduction/accumulate, BH_MAXDIM otherwise. */
    int sweep_axis() const;
    void reshape(const std::vector<int64_t> &shape);
    void remove_axis(int64_t axis);
    void transpose(int64_t axis1, int64_t axis2);
    bh_type operand_type(int operand_index) const;
    std::string pprint(bool python_notation = true) const;

    bool operator==(const bh_instruction &other) const;
    bool operator!=(const bh_instruction &other) const { return !(*this == other); }
};

/* True when `a` and `b` must keep their relative order (one writes what the other touches). */
bool bh_instr_dependency(const bh_instruction *a, const bh_instruction *b);
