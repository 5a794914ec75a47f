/* Loop-nest Block IR consumed by the engine's code generators.
 *
 * A kernel is a LoopB of rank -1; a LoopB of rank r >= 0 is a loop over axis r
 * of its instructions' principal shape; an InstrB is one instruction placed at
 * rank ndim.  The members (rank, size, _block_list, _sweeps, _news, _frees,
 * _reshapable) and their meaning follow include/bohrium/jitk/block.hpp:54-277
 * of the reference, so fusion partitions and array contraction are defined the
 * same way.
 */
#pragma once

#include <memory>
#include <set>
#include <string>
#include <variant>
#include <vector>
#include <bohrium/bh_instruction.hpp>

namespace bohrium {
namespace jitk {

typedef std::shared_ptr<const bh_instruction> InstrPtr;

class Block;

class LoopB {
public:
    int rank = -42;
    int64_t size = -42;
    std::vector<Block> _block_list;
    std::set<InstrPtr> _sweeps;
    std::set<bh_base *> _news;
    std::set<bh_base *> _frees;
    /* Bases released after the kernel without being temporaries of this loop: frees that a
     * transformer lifted out of the nested loop that held them (a base made in one nested loop and
     * read in a sibling must stay an array). */
    std::set<bh_base *> _released;
    bool _reshapable = false;

    LoopB() = default;
    LoopB(int rank, int64_t size, std::vector<Block> block_list = {});

    const std::set<InstrPtr> &getSweeps() const { return _sweeps; }
    bool isInnermost() const;
    bool isSystemOnly() const;
    std::vector<const LoopB *> getLocalSubBlocks() const;
    std::set<const bh_base *> getAllBases() const;
    void getAllNews(std::set<bh_base *> &out) const;
    std::set<bh_base *> getAllNews() const;
    void getAllFrees(std::set<bh_base *> &out) const;
    std::set<bh_base *> getAllFrees() const;
    /* Temporaries this loop declares: created here and freed inside, or freed here and created inside. */
    std::set<bh_base *> getLocalTemps() const;
    std::set<bh_base *> getAllTemps() const;
    std::set<bh_base *> getAllNonTemps() const;
    /* Number of parallel iterations of this loop alone (0 when it sweeps). */
    uint64_t localThreading() const;
    int replaceInstr(const InstrPtr &subject, const bh_instruction &replacement);
    void metadataUpdate();
    bool validation() const;
    std::string pprint(const char *newline = "\n") const;
};

struct InstrB {
    InstrPtr instr;
    int rank;
};

class Block {
    std::variant<LoopB, InstrB> _var;

public:
    Block() : _var(LoopB()) {}
    explicit Block(const LoopB &l) : _var(l) {}
    explicit Block(LoopB &&l) : _var(std::move(l)) {}
    Block(const bh_instruction &instr, int rank) : _var(InstrB{std::make_shared<bh_instruction>(instr), rank}) {}
    Block(InstrPtr instr, int rank) : _var(InstrB{std::move(instr), rank}) {}

    bool isInstr() const { return _var.index() == 1; }
    LoopB &getLoop() { return std::get<LoopB>(_var); }
    const LoopB &getLoop() const { return std::get<LoopB>(_var); }
    const InstrPtr &getInstr() const { return std::get<InstrB>(_var).instr; }
    void setInstr(const bh_instruction &instr);
    int rank() const { return isInstr() ? std::get<InstrB>(_var).rank : getLoop().rank; }
    bool isSystemOnly() const;
    bool isReshapable() const;
    bool validation() const;
    std::set<const bh_base *> getAllBases() const;
    bool dependOn(const Block &other) const;
    std::string pprint(const char *newline = "\n") const;
};

/* Iteration helpers (pre-order over every instruction, or only the direct children). */
void allInstr(const LoopB &loop, std::vector<InstrPtr> &out);
std::vector<InstrPtr> allInstr(const LoopB &loop);
std::vector<InstrPtr> allInstr(const Block &block);
std::vector<InstrPtr> allLocalInstr(const LoopB &loop);
std::set<const bh_base *> allBases(const bh_instruction &instr);

LoopB merge(const LoopB &a, const LoopB &b);
Block create_nested_block(const std::vector<InstrPtr> &instr_list, int rank = 0, std::set<bh_base *> frees = {});
Block create_nested_block(const std::vector<InstrPtr> &instr_list, int rank, int64_t size_of_rank_dim);
/* Leading ranks without sweeps (each the single child of the previous) and their total threading. */
std::pair<uint64_t, uint64_t> parallel_ranks(const LoopB &block, unsigned int max_depth = 16);
std::vector<const LoopB *> get_first_loop_blocks(const LoopB &block);
bool mergeable(const Block &b1, const Block &b2, bool avoid_rank0_sweep);
Block reshape_and_merge(const LoopB &l1, const LoopB &l2);
InstrPtr reshape_rank(const InstrPtr &instr, int rank, int64_t size_of_rank_dim);

/* Identity value that starts a sweep (reference core/jitk/instruction.cpp:546-578). */
bh_constant sweep_identity(bh_opcode opcode, bh_type dtype);

} // namespace jitk
} // namespace bohrium
