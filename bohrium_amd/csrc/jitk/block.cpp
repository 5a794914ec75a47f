/* Block IR operations.  The rules (nesting, temporaries = news x frees, parallel
 * ranks, data-parallel compatibility, merge/reshape) restate the reference's
 * core/jitk/block.cpp:36-693 and core/jitk/instruction.cpp:546-619. */
#include <bohrium/jitk/block.hpp>

#include <algorithm>
#include <cassert>
#include <sstream>
#include <stdexcept>

namespace bohrium {
namespace jitk {

// ---------- iteration helpers ----------

void allInstr(const LoopB &loop, std::vector<InstrPtr> &out) {
    for (const Block &b : loop._block_list) {
        if (b.isInstr()) {
            if (b.getInstr() != nullptr) out.push_back(b.getInstr());
        } else {
            allInstr(b.getLoop(), out);
        }
    }
}

std::vector<InstrPtr> allInstr(const LoopB &loop) {
    std::vector<InstrPtr> ret;
    allInstr(loop, ret);
    return ret;
}

std::vector<InstrPtr> allInstr(const Block &block) {
    if (block.isInstr()) return {block.getInstr()};
    return allInstr(block.getLoop());
}

std::vector<InstrPtr> allLocalInstr(const LoopB &loop) {
    std::vector<InstrPtr> ret;
    for (const Block &b : loop._block_list)
        if (b.isInstr() && b.getInstr() != nullptr) ret.push_back(b.getInstr());
    return ret;
}

std::set<const bh_base *> allBases(const bh_instruction &instr) {
    std::set<const bh_base *> ret;
    for (const bh_view &v : instr.operand)
        if (!v.isConstant()) ret.insert(v.base);
    return ret;
}

namespace {
bool is_reshapeable(const std::vector<InstrPtr> &instr_list) {
    if (instr_list.empty()) return true;
    const int64_t nd = instr_list.front()->ndim();
    for (const InstrPtr &i : instr_list)
        if (!i->reshapable() || i->ndim() != nd) return false;
    return true;
}
} // namespace

// ---------- LoopB ----------

LoopB::LoopB(int rank, int64_t size, std::vector<Block> block_list)
    : rank(rank), size(size), _block_list(std::move(block_list)) {}

bool LoopB::isInnermost() const {
    for (const Block &b : _block_list)
        if (!b.isInstr()) return false;
    return true;
}

bool LoopB::isSystemOnly() const {
    for (const Block &b : _block_list)
        if (!b.isSystemOnly()) return false;
    return true;
}

std::vector<const LoopB *> LoopB::getLocalSubBlocks() const {
    std::vector<const LoopB *> ret;
    for (const Block &b : _block_list)
        if (!b.isInstr()) ret.push_back(&b.getLoop());
    return ret;
}

std::set<const bh_base *> LoopB::getAllBases() const {
    std::set<const bh_base *> ret;
    for (const InstrPtr &i : allInstr(*this)) {
        auto t = allBases(*i);
        ret.insert(t.begin(), t.end());
    }
    return ret;
}

void LoopB::getAllNews(std::set<bh_base *> &out) const {
    out.insert(_news.begin(), _news.end());
    for (const Block &b : _block_list)
        if (!b.isInstr()) b.getLoop().getAllNews(out);
}

std::set<bh_base *> LoopB::getAllNews() const {
    std::set<bh_base *> r;
    getAllNews(r);
    return r;
}

void LoopB::getAllFrees(std::set<bh_base *> &out) const {
    out.insert(_frees.begin(), _frees.end());
    out.insert(_released.begin(), _released.end());
    for (const Block &b : _block_list)
        if (!b.isInstr()) b.getLoop().getAllFrees(out);
}

std::set<bh_base *> LoopB::getAllFrees() const {
    std::set<bh_base *> r;
    getAllFrees(r);
    return r;
}

std::set<bh_base *> LoopB::getLocalTemps() const {
    std::set<bh_base *> out;
    const std::set<bh_base *> frees = getAllFrees();
    std::set_intersection(_news.begin(), _news.end(), frees.begin(), frees.end(), std::inserter(out, out.begin()));
    const std::set<bh_base *> news = getAllNews();
    std::set_intersection(_frees.begin(), _frees.end(), news.begin(), news.end(), std::inserter(out, out.begin()));
    return out;
}

std::set<bh_base *> LoopB::getAllTemps() const {
    std::set<bh_base *> out = getLocalTemps();
    for (const Block &b : _block_list) {
        if (!b.isInstr()) {
            auto t = b.getLoop().getAllTemps();
            out.insert(t.begin(), t.end());
        }
    }
    return out;
}

std::set<bh_base *> LoopB::getAllNonTemps() const {
    std::set<bh_base *> out;
    const auto tmps = getAllTemps();
    for (const bh_base *b : getAllBases()) {
        auto *nb = const_cast<bh_base *>(b);
        if (tmps.count(nb) == 0) out.insert(nb);
    }
    return out;
}

uint64_t LoopB::localThreading() const {
    if (_sweeps.empty() && !isSystemOnly()) return static_cast<uint64_t>(size);
    return 0;
}

int LoopB::replaceInstr(const InstrPtr &subject, const bh_instruction &replacement) {
    int hits = 0;
    for (Block &b : _block_list) {
        if (b.isInstr()) {
            if (*b.getInstr() == *subject) {
                b.setInstr(replacement);
                ++hits;
            }
        } else {
            hits += b.getLoop().replaceInstr(subject, replacement);
        }
    }
    return hits;
}

void LoopB::metadataUpdate() {
    _news.clear();
    _sweeps.clear();
    for (const InstrPtr &i : allLocalInstr(*this))
        if (i->constructor) _news.insert(i->operand[0].base);
    for (const InstrPtr &i : allInstr(*this))
        if (i->sweep_axis() == rank) _sweeps.insert(i);
    _reshapable = is_reshapeable(allInstr(*this));
}

bool LoopB::validation() const {
    if (size < 0 || rank < 0) return false;
    for (const InstrPtr &i : allInstr(*this)) {
        if (bh_opcode_is_system(i->opcode)) return false;
        if (i->ndim() <= rank) return false;
        if (i->shape()[rank] != size) return false;
    }
    for (const Block &b : _block_list)
        if (!b.validation()) return false;
    for (const InstrPtr &i : allLocalInstr(*this))
        if (i->ndim() != rank + 1) return false;
    return true;
}

std::string LoopB::pprint(const char *newline) const {
    std::ostringstream ss;
    ss << std::string(std::max(rank, 0) * 4, ' ') << "rank: " << rank << ", size: " << size;
    if (!_sweeps.empty()) ss << ", sweeps: " << _sweeps.size();
    if (_reshapable) ss << ", reshapable";
    ss << ", block list:" << newline;
    for (const Block &b : _block_list) ss << b.pprint(newline);
    return ss.str();
}

// ---------- Block ----------

void Block::setInstr(const bh_instruction &instr) {
    std::get<InstrB>(_var).rank = static_cast<int>(instr.ndim());
    std::get<InstrB>(_var).instr = std::make_shared<bh_instruction>(instr);
}

bool Block::isSystemOnly() const {
    if (isInstr()) return bh_opcode_is_system(getInstr()->opcode);
    return getLoop().isSystemOnly();
}

bool Block::isReshapable() const { return isInstr() ? getInstr()->reshapable() : getLoop()._reshapable; }

bool Block::validation() const {
    if (isInstr()) return getInstr()->ndim() == rank();
    return getLoop().validation();
}

std::set<const bh_base *> Block::getAllBases() const {
    std::set<const bh_base *> ret;
    for (const InstrPtr &i : allInstr(*this)) {
        auto t = allBases(*i);
        ret.insert(t.begin(), t.end());
    }
    return ret;
}

bool Block::dependOn(const Block &other) const {
    const auto mine = allInstr(*this);
    const auto theirs = allInstr(other);
    for (const InstrPtr &a : mine)
        for (const InstrPtr &b : theirs)
            if (bh_instr_dependency(a.get(), b.get())) return true;
    return false;
}

std::string Block::pprint(const char *newline) const {
    if (isInstr()) {
        std::ostringstream ss;
        ss << std::string(rank() * 4, ' ') << getInstr()->pprint() << newline;
        return ss.str();
    }
    return getLoop().pprint(newline);
}

// ---------- construction ----------

InstrPtr reshape_rank(const InstrPtr &instr, int rank, int64_t size_of_rank_dim) {
    std::vector<int64_t> shape(rank + 1);
    for (int r = 0; r < rank; ++r) shape[r] = instr->operand[0].shape[r];
    int64_t size = 1;
    for (int64_t r = rank; r < instr->operand[0].ndim; ++r) size *= instr->operand[0].shape[r];
    shape[rank] = size_of_rank_dim;
    if (size != size_of_rank_dim) {
        if (size % size_of_rank_dim != 0) throw std::runtime_error("reshape_rank(): not divisible");
        shape.push_back(size / size_of_rank_dim);
    }
    bh_instruction ret(*instr);
    ret.reshape(shape);
    return std::make_shared<bh_instruction>(ret);
}

namespace {
void add_instr_to_block(LoopB &block, InstrPtr instr, int rank, int64_t size_of_rank_dim) {
    if (instr->ndim() <= rank) throw std::runtime_error("add_instr_to_block(): ndim <= rank");
    if (instr->reshapable() && instr->operand[0].shape[rank] != size_of_rank_dim)
        instr = reshape_rank(instr, rank, size_of_rank_dim);
    const BhIntVec shape = instr->shape();
    if (shape[rank] != size_of_rank_dim) throw std::runtime_error("add_instr_to_block(): shape mismatch");
    if (instr->ndim() > rank + 1) {
        block._block_list.push_back(create_nested_block({instr}, rank + 1, shape[rank + 1]));
    } else if (instr->opcode == BH_FREE) {
        block._frees.insert(instr->operand[0].base);
    } else {
        block._block_list.emplace_back(instr, rank + 1);
    }
    block.metadataUpdate();
}
} // namespace

Block create_nested_block(const std::vector<InstrPtr> &instr_list, int rank, std::set<bh_base *> frees) {
    if (instr_list.empty()) throw std::runtime_error("create_nested_block: empty instruction list");
    const BhIntVec shape = instr_list[0]->shape();
    const int ndim = static_cast<int>(shape.size());
    LoopB ret;
    ret.rank = rank;
    ret.size = shape[rank];
    if (rank == ndim - 1) {
        ret._frees = std::move(frees);
        for (const InstrPtr &i : instr_list) {
            if (i->opcode == BH_FREE) {
                ret._frees.insert(i->operand[0].base);
            } else {
                ret._block_list.emplace_back(i, ndim);
            }
        }
    } else {
        ret._block_list.push_back(create_nested_block(instr_list, rank + 1, std::move(frees)));
    }
    ret.metadataUpdate();
    return Block(std::move(ret));
}

Block create_nested_block(const std::vector<InstrPtr> &instr_list, int rank, int64_t size_of_rank_dim) {
    if (instr_list.empty()) throw std::runtime_error("create_nested_block: empty instruction list");
    LoopB ret(rank, size_of_rank_dim);
    for (const InstrPtr &i : instr_list) add_instr_to_block(ret, i, rank, size_of_rank_dim);
    return Block(std::move(ret));
}

LoopB merge(const LoopB &l1, const LoopB &l2) {
    LoopB ret(l1);
    ret._block_list.insert(ret._block_list.end(), l2._block_list.begin(), l2._block_list.end());
    ret._sweeps.insert(l2._sweeps.begin(), l2._sweeps.end());
    ret._news.insert(l2._news.begin(), l2._news.end());
    ret._frees.insert(l2._frees.begin(), l2._frees.end());
    ret._released.insert(l2._released.begin(), l2._released.end());
    ret._reshapable = is_reshapeable(allInstr(ret));
    return ret;
}

std::pair<uint64_t, uint64_t> parallel_ranks(const LoopB &block, unsigned int max_depth) {
    std::pair<uint64_t, uint64_t> ret{0, 0};
    const uint64_t thds = block.localThreading();
    --max_depth;
    if (thds > 0) {
        if (max_depth > 0) {
            const size_t nblocks = block.getLocalSubBlocks().size();
            const bool no_instr = allLocalInstr(block).empty();
            if (nblocks == 1 && no_instr) {
                auto sub = parallel_ranks(block._block_list[0].getLoop(), max_depth);
                ret.first += sub.first;
                ret.second += sub.second;
            }
        }
        ret.second += thds;
        ++ret.first;
    }
    return ret;
}

std::vector<const LoopB *> get_first_loop_blocks(const LoopB &block) {
    std::vector<const LoopB *> out;
    const LoopB *cur = &block;
    while (true) {
        out.push_back(cur);
        if (cur->_block_list.empty() || cur->_block_list[0].isInstr()) break;
        cur = &cur->_block_list[0].getLoop();
    }
    return out;
}

// ---------- merge rules ----------

namespace {
bool dp_compatible(const bh_view &writer, const bh_view &reader) {
    if (writer.isConstant() || reader.isConstant() || writer.base != reader.base) return true;
    if (writer.start != reader.start) return false;
    if (writer.ndim == reader.ndim) return writer.shape == reader.shape && writer.stride == reader.stride;
    return writer.shape.prod() == reader.shape.prod() && writer.isContiguous() && reader.isContiguous();
}

bool writes_arbitrarily_conflict(const bh_instruction &a, const bh_instruction &b) {
    if (b.opcode == BH_GATHER && a.operand[0].base == b.operand[1].base) return true;
    if (a.opcode == BH_SCATTER || a.opcode == BH_COND_SCATTER) {
        for (const bh_view &v : b.operand)
            if (!v.isConstant() && a.operand[0].base == v.base) return true;
    } else if (b.opcode == BH_SCATTER || b.opcode == BH_COND_SCATTER) {
        for (const bh_view &v : a.operand)
            if (!v.isConstant() && b.operand[0].base == v.base) return true;
    }
    return false;
}

bool dp_compatible(const InstrPtr &a, const InstrPtr &b) {
    if (bh_opcode_is_system(a->opcode) || bh_opcode_is_system(b->opcode)) return true;
    if (writes_arbitrarily_conflict(*a, *b)) return false;
    for (const bh_view &v : b->operand)
        if (!dp_compatible(a->operand[0], v)) return false;
    for (const bh_view &v : a->operand)
        if (!dp_compatible(b->operand[0], v)) return false;
    return true;
}

bool dp_compatible(const LoopB &b1, const LoopB &b2) {
    const auto l1 = allInstr(b1);
    const auto l2 = allInstr(b2);
    for (const InstrPtr &i1 : l1)
        for (const InstrPtr &i2 : l2)
            if (i1.get() != i2.get() && !dp_compatible(i1, i2)) return false;
    return true;
}

bool sweeps_accessed_by_block(const std::set<InstrPtr> &sweeps, const LoopB &loop) {
    if (sweeps.empty()) return false;
    const auto bases = loop.getAllBases();
    for (const InstrPtr &i : sweeps)
        if (bases.count(i->operand[0].base)) return true;
    return false;
}

Block reshape_loop(const LoopB &l1, int64_t size_of_rank_dim) {
    std::vector<InstrPtr> list;
    for (const InstrPtr &i : allInstr(l1)) list.push_back(reshape_rank(i, l1.rank, size_of_rank_dim));
    if (!list.empty()) return create_nested_block(list, l1.rank, l1.getAllFrees());
    LoopB ret = l1;
    ret.size = size_of_rank_dim;
    return Block(std::move(ret));
}
} // namespace

Block reshape_and_merge(const LoopB &l1, const LoopB &l2) {
    if (l1.size == l2.size) return Block(merge(l1, l2));
    if (l2._reshapable && l2.size % l1.size == 0) return Block(merge(l1, reshape_loop(l2, l1.size).getLoop()));
    if (l1._reshapable && l1.size % l2.size == 0) return Block(merge(reshape_loop(l1, l2.size).getLoop(), l2));
    // A block of frees only: its frees are released with the other block, never made temporaries
    // of it (the freed bases may come from loops nested in it).
    if (allInstr(l1).empty()) {
        LoopB ret = l2;
        auto f = l1.getAllFrees();
        ret._released.insert(f.begin(), f.end());
        return Block(std::move(ret));
    }
    if (allInstr(l2).empty()) {
        LoopB ret = l1;
        auto f = l2.getAllFrees();
        ret._released.insert(f.begin(), f.end());
        return Block(std::move(ret));
    }
    throw std::runtime_error("reshape_and_merge: the blocks are not mergeable");
}

bool mergeable(const Block &b1, const Block &b2, bool avoid_rank0_sweep) {
    if (b1.isInstr() || b2.isInstr()) return false;
    const LoopB &l1 = b1.getLoop();
    const LoopB &l2 = b2.getLoop();
    if (l2.isSystemOnly()) return true;
    if (avoid_rank0_sweep && l1.rank == 0 && l2.rank == 0 && (l1._sweeps.empty() != l2._sweeps.empty())) return false;
    if (sweeps_accessed_by_block(l1._sweeps, l2)) return false;
    if (l1.size == l2.size || (l2._reshapable && l2.size % l1.size == 0) ||
        (l1._reshapable && l1.size % l2.size == 0)) {
        return dp_compatible(l1, l2);
    }
    return false;
}

// ---------- identities ----------

bh_constant sweep_identity(bh_opcode opcode, bh_type dtype) {
    switch (opcode) {
        case BH_ADD_REDUCE:
        case BH_BITWISE_OR_REDUCE:
        case BH_BITWISE_XOR_REDUCE:
        case BH_LOGICAL_OR_REDUCE:
        case BH_LOGICAL_XOR_REDUCE:
        case BH_ADD_ACCUMULATE: return bh_constant::make(0, dtype);
        case BH_MULTIPLY_REDUCE:
        case BH_MULTIPLY_ACCUMULATE: return bh_constant::make(1, dtype);
        case BH_BITWISE_AND_REDUCE:
        case BH_LOGICAL_AND_REDUCE: return bh_constant::make(~0u, dtype);  // 32-bit all-ones, as the reference
        case BH_MAXIMUM_REDUCE:
            return dtype == bh_type::BOOL ? bh_constant::make(1, dtype) : bh_constant::get_min(dtype);
        case BH_MINIMUM_REDUCE:
            return dtype == bh_type::BOOL ? bh_constant::make(1, dtype) : bh_constant::get_max(dtype);
        default: throw std::runtime_error("sweep_identity(): unsupported operation");
    }
}

} // namespace jitk
} // namespace bohrium
