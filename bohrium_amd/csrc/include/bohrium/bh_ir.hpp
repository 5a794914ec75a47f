/* A flushed batch of instructions: the unit the bridge hands down the stack.
 * Same members as include/bohrium/bh_ir.hpp:30-107 of the reference.
 */
#pragma once

#include <set>
#include <vector>
#include <bohrium/bh_instruction.hpp>

class BhIR {
public:
    std::vector<bh_instruction> instr_list;
    std::set<bh_base *> _syncs;
    uint64_t _nrepeats;
    /* Repeat while this single-element BH_BOOL base holds true (nullptr: always). */
    bh_base *_repeat_condition;

    BhIR(std::vector<bh_instruction> instr_list, std::set<bh_base *> syncs, uint64_t nrepeats = 1,
         bh_base *repeat_condition = nullptr)
        : instr_list(std::move(instr_list)), _syncs(std::move(syncs)), _nrepeats(nrepeats),
          _repeat_condition(repeat_condition) {}

    const std::set<bh_base *> getSyncs() const { return _syncs; }
    uint64_t getNRepeats() const { return _nrepeats; }
    bh_base *getRepeatCondition() { return _repeat_condition; }
};
