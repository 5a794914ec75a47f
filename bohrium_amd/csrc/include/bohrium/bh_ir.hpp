/* A flushed batch of instructions: the unit the bridge hands down the stack.
 * Same members as include/bohrium/bh_ir.hpp:30-107 of the reference.
 */
#pragma once

#include <map>
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

    /* De-serialize an archive written by writeSerializedArchive (reference core/bh_ir.cpp:37-117).
     * Base arrays are named by the sender's base addresses ("remote" bases): `remote2local` maps
     * them to local bases and gains an entry for each base the archive introduces; those that
     * arrive with data are appended to `data_recv` (their bytes follow the archive, in that order);
     * the remote bases freed by the batch's BH_FREEs are added to `frees`.  Throws
     * std::runtime_error on a malformed archive. */
    BhIR(const std::vector<char> &serialized_archive, std::map<const bh_base *, bh_base> &remote2local,
         std::vector<bh_base *> &data_recv, std::set<bh_base *> &frees);
    /* Serialize this batch (reference core/bh_ir.cpp:119-154): repeats, repeat condition,
     * instructions, syncs, and every base not yet in `known_base_arrays` (added to it), in the
     * order of first appearance.  The bases among them that hold host data are appended to
     * `new_data`; the sender sends their bytes after the archive. */
    std::vector<char> writeSerializedArchive(std::set<bh_base *> &known_base_arrays, std::vector<bh_base *> &new_data);
    const std::set<bh_base *> getSyncs() const { return _syncs; }
    uint64_t getNRepeats() const { return _nrepeats; }
    bh_base *getRepeatCondition() { return _repeat_condition; }
};
