/* Fusion pipeline: simplification, pre-fusion, DAG fusers, transformers and
 * kernel splitting.  Restates the reference algorithms named in fuser.hpp with
 * its own graph representation (index-based adjacency sets instead of
 * Boost.Graph); partition results can differ where the reference breaks ties
 * on Boost's edge order. */
#include <bohrium/jitk/fuser.hpp>
#include <bohrium/flat_ptr_map.hpp>

#include <algorithm>
#include <cassert>
#include <deque>
#include <map>
#include <stdexcept>
#include <unordered_set>

namespace bohrium {
namespace jitk {

// ---------------------------------------------------------------- simplify

void simplify_instr(bh_instruction &instr) {
    if (instr.operand.empty()) return;
    {
        const BhIntVec shape = instr.shape();
        const int sa = instr.sweep_axis();
        size_t ndim_left = bh_opcode_is_reduction(instr.opcode) ? shape.size() - 1 : shape.size();
        for (int64_t i = static_cast<int64_t>(shape.size()) - 1; i >= 0 && ndim_left > 1; --i) {
            if (sa != i && shape[i] == 1) {
                instr.remove_axis(i);
                --ndim_left;
            }
        }
    }
    if (instr.ndim() > 1 && instr.reshapable()) {
        instr.reshape({instr.shape().prod()});
    }
}

namespace {
bool fully_dp_compatible(const bh_view &w, const bh_view &r) {
    if (w.isConstant() || r.isConstant() || w.base != r.base) return true;
    return w.start == r.start && w.ndim == r.ndim && w.shape == r.shape && w.stride == r.stride;
}

bool fully_dp_compatible(const InstrPtr &a, const InstrPtr &b) {
    if (bh_opcode_is_system(a->opcode) || bh_opcode_is_system(b->opcode)) return true;
    if (b->opcode == BH_GATHER && a->operand[0].base == b->operand[1].base) return false;
    if (a->opcode == BH_SCATTER || a->opcode == BH_COND_SCATTER) {
        for (const bh_view &v : b->operand)
            if (!v.isConstant() && a->operand[0].base == v.base) return false;
    } else if (b->opcode == BH_SCATTER || b->opcode == BH_COND_SCATTER) {
        for (const bh_view &v : a->operand)
            if (!v.isConstant() && b->operand[0].base == v.base) return false;
    }
    for (const bh_view &v : b->operand)
        if (!fully_dp_compatible(a->operand[0], v)) return false;
    for (const bh_view &v : a->operand)
        if (!fully_dp_compatible(b->operand[0], v)) return false;
    return true;
}

bool fully_fusible(const std::vector<InstrPtr> &list, const InstrPtr &instr) {
    if (list.empty() || bh_opcode_is_system(instr->opcode)) return true;
    const auto dshape = instr->shape();
    for (const InstrPtr &i : list) {
        if (bh_opcode_is_system(i->opcode)) continue;
        if (i->shape() != dshape || !fully_dp_compatible(i, instr)) return false;
    }
    return true;
}
} // namespace

// Moves every BH_FREE right after the last instruction touching its base and
// simplifies the shapes (reference fuser.cpp:167-221).
std::vector<InstrPtr> simplify_instr_list(const std::vector<bh_instruction *> &instr_list) {
    std::map<const bh_instruction *, std::vector<bh_base *> > last_access;
    std::map<bh_base *, const bh_instruction *> free_of;
    std::set<const bh_instruction *> free_instrs;
    {
        std::set<bh_base *> pending;
        for (auto it = instr_list.rbegin(); it != instr_list.rend(); ++it) {
            const bh_instruction *instr = *it;
            if (instr->opcode == BH_FREE) {
                free_instrs.insert(instr);
                free_of[instr->operand[0].base] = instr;
                pending.insert(instr->operand[0].base);
            } else {
                std::vector<bh_base *> hits;
                for (const bh_view &v : instr->operand)
                    if (!v.isConstant() && pending.count(v.base)) {
                        if (std::find(hits.begin(), hits.end(), v.base) == hits.end()) hits.push_back(v.base);
                    }
                if (!hits.empty()) {
                    std::sort(hits.begin(), hits.end());
                    last_access[instr] = hits;
                    for (bh_base *b : hits) pending.erase(b);
                }
            }
        }
    }
    std::vector<InstrPtr> ret;
    for (const bh_instruction *instr : instr_list) {
        if (instr->operand.empty() || free_instrs.count(instr)) continue;
        bh_instruction s(*instr);
        simplify_instr(s);
        ret.push_back(std::make_shared<const bh_instruction>(s));
        auto la = last_access.find(instr);
        if (la != last_access.end()) {
            for (bh_base *b : la->second) ret.push_back(std::make_shared<const bh_instruction>(*free_of.at(b)));
        }
    }
    return ret;
}

namespace {
// Whether `i` reads or writes an array that an instruction of `g` reads or writes.
bool shares_base(const std::vector<InstrPtr> &g, const InstrPtr &i) {
    for (const bh_view &v : i->operand) {
        if (v.isConstant()) continue;
        for (const InstrPtr &j : g)
            for (const bh_view &w : j->operand)
                if (!w.isConstant() && w.base == v.base) return true;
    }
    return false;
}
} // namespace

std::vector<Block> pre_fuser_lossy(const std::vector<bh_instruction *> &instr_list, bool accumulate_alone,
                                   bool one_sweep_axis) {
    const std::vector<InstrPtr> list = simplify_instr_list(instr_list);
    std::vector<std::vector<InstrPtr> > groups;
    for (auto it = list.begin(); it != list.end();) {
        groups.push_back({*it});
        std::vector<InstrPtr> &g = groups.back();
        if (bh_opcode_is_system((*it)->opcode)) {
            // A group never starts with a system instruction; it is shape-less.
            ++it;
            continue;
        }
        bool has_acc = bh_opcode_is_accumulate((*it)->opcode) != 0;
        bool has_sweep = bh_opcode_is_sweep((*it)->opcode) != 0;
        int axis = has_sweep ? (*it)->sweep_axis() : -1;
        ++it;
        for (; it != list.end(); ++it) {
            const bool acc = bh_opcode_is_accumulate((*it)->opcode) != 0;
            const bool sweep = bh_opcode_is_sweep((*it)->opcode) != 0;
            // Sweeps along two axes of one shape cannot both be inner reductions of a loop nest.
            if (one_sweep_axis && sweep && axis >= 0 && (*it)->sweep_axis() != axis) break;
            if (accumulate_alone && ((has_acc && sweep) || (acc && has_sweep))) {
                if (acc) {
                    // The element-wise instructions at the end of the group that only feed this
                    // accumulate start its group instead, so their results need no array.
                    std::set<const bh_base *> feeds;
                    for (size_t o = 1; o < (*it)->operand.size(); ++o)
                        if (!(*it)->operand[o].isConstant()) feeds.insert((*it)->operand[o].base);
                    size_t keep = g.size();
                    while (keep > 1) {
                        const InstrPtr &c = g[keep - 1];
                        if (bh_opcode_is_system(c->opcode)) {
                            --keep;
                            continue;
                        }
                        if (bh_opcode_is_sweep(c->opcode) || c->operand.empty() || !feeds.count(c->operand[0].base)) break;
                        for (size_t o = 1; o < c->operand.size(); ++o)
                            if (!c->operand[o].isConstant()) feeds.insert(c->operand[o].base);
                        --keep;
                    }
                    while (keep < g.size() && bh_opcode_is_system(g[keep]->opcode)) ++keep;  // frees stay put
                    if (keep < g.size()) {
                        std::vector<InstrPtr> moved(g.begin() + static_cast<std::ptrdiff_t>(keep), g.end());
                        g.resize(keep);
                        groups.push_back(moved);
                        std::vector<InstrPtr> &n = groups.back();
                        n.push_back(*it);
                        ++it;
                        for (; it != list.end(); ++it) {
                            if (bh_opcode_is_sweep((*it)->opcode) || !fully_fusible(n, *it)) break;
                            if (!bh_opcode_is_system((*it)->opcode) && !shares_base(n, *it)) break;  // as below
                            n.push_back(*it);
                        }
                    }
                }
                break;
            }
            // A group with an accumulate (a GPU scan) takes no instruction that shares no array with
            // it: that work would only wait on the scan's tiles.
            if (accumulate_alone && has_acc && !acc && !bh_opcode_is_system((*it)->opcode) && !shares_base(g, *it))
                break;
            if (fully_fusible(g, *it)) {
                g.push_back(*it);
                has_acc |= acc;
                has_sweep |= sweep;
                if (sweep && axis < 0) axis = (*it)->sweep_axis();
            } else {
                break;
            }
        }
    }
    std::vector<Block> ret;
    for (const auto &g : groups) ret.push_back(create_nested_block(g));
    return ret;
}

std::vector<Block> fuser_singleton(const std::vector<bh_instruction *> &instr_list) {
    const std::vector<InstrPtr> list = simplify_instr_list(instr_list);
    std::vector<Block> ret;
    for (const InstrPtr &i : list) ret.push_back(create_nested_block({i}, 0, i->shape()[0]));
    return ret;
}

void fuser_serial(std::vector<Block> &block_list, const FusionConfig &config) {
    std::vector<Block> ret;
    for (auto it = block_list.begin(); it != block_list.end();) {
        ret.push_back(*it);
        Block &cur = ret.back();
        ++it;
        if (cur.isInstr()) continue;
        for (; it != block_list.end(); ++it) {
            if (!it->isInstr() && config.can_merge(cur, *it)) {
                cur = reshape_and_merge(cur.getLoop(), it->getLoop());
            } else {
                break;
            }
        }
        fuser_serial(cur.getLoop()._block_list, config);
    }
    block_list = ret;
}

// ---------------------------------------------------------------- costs

uint64_t weight(const Block &b1, const Block &b2) {
    if (b1.isInstr() || b2.isInstr()) return 0;
    const auto news = b1.getLoop().getAllNews();
    const auto frees = b2.getLoop().getAllFrees();
    uint64_t total = 0;
    for (bh_base *b : news)
        if (frees.count(b)) total += static_cast<uint64_t>(b->nbytes());
    return total;
}

uint64_t block_cost(const Block &block) {
    std::vector<const bh_base *> non_temps;
    const std::set<bh_base *> temps = block.isInstr() ? std::set<bh_base *>() : block.getLoop().getAllTemps();
    for (const InstrPtr &i : allInstr(block)) {
        for (const bh_view &v : i->operand) {
            if (v.isConstant() || temps.count(v.base)) continue;
            if (std::find(non_temps.begin(), non_temps.end(), v.base) == non_temps.end()) non_temps.push_back(v.base);
        }
    }
    uint64_t total = 0;
    for (const bh_base *b : non_temps) total += static_cast<uint64_t>(b->nbytes());
    return total;
}

// ---------------------------------------------------------------- DAG

namespace {
struct Dag {
    std::vector<Block> node;
    std::vector<std::set<int> > child, parent;
    std::vector<bool> alive;

    int add(const Block &b) {
        node.push_back(b);
        child.emplace_back();
        parent.emplace_back();
        alive.push_back(true);
        return static_cast<int>(node.size()) - 1;
    }
    void add_edge(int a, int b) {
        if (a == b) return;
        child[a].insert(b);
        parent[b].insert(a);
    }
    void remove_edge(int a, int b) {
        child[a].erase(b);
        parent[b].erase(a);
    }
    size_t num_edges() const {
        size_t n = 0;
        for (size_t i = 0; i < child.size(); ++i)
            if (alive[i]) n += child[i].size();
        return n;
    }
    // Is there a path a -> ... -> b of length >= 2?
    bool long_path(int a, int b) const {
        std::vector<char> seen(node.size(), 0);
        std::deque<int> q;
        for (int c : child[a])
            if (c != b) {
                q.push_back(c);
                seen[c] = 1;
            }
        while (!q.empty()) {
            const int u = q.front();
            q.pop_front();
            if (u == b) return true;
            for (int c : child[u])
                if (!seen[c]) {
                    seen[c] = 1;
                    q.push_back(c);
                }
        }
        return false;
    }
    void merge(int a, int b) {
        node[a] = reshape_and_merge(node[a].getLoop(), node[b].getLoop());
        for (int c : std::vector<int>(child[b].begin(), child[b].end())) {
            remove_edge(b, c);
            add_edge(a, c);
        }
        for (int p : std::vector<int>(parent[b].begin(), parent[b].end())) {
            remove_edge(p, b);
            if (p != a) add_edge(p, a);
        }
        alive[b] = false;
    }
    // Kahn's algorithm, smallest original position first.
    std::vector<Block> topo() const {
        std::vector<int> indeg(node.size(), 0);
        for (size_t i = 0; i < node.size(); ++i)
            if (alive[i])
                for (int c : child[i]) ++indeg[c];
        std::set<int> ready;
        for (size_t i = 0; i < node.size(); ++i)
            if (alive[i] && indeg[i] == 0) ready.insert(static_cast<int>(i));
        std::vector<Block> ret;
        while (!ready.empty()) {
            const int v = *ready.begin();
            ready.erase(ready.begin());
            ret.push_back(node[v]);
            for (int c : child[v])
                if (--indeg[c] == 0) ready.insert(c);
        }
        return ret;
    }
};

Dag from_block_list(const std::vector<Block> &block_list) {
    Dag dag;
    std::map<const bh_base *, std::set<int> > base2v;
    for (const Block &block : block_list) {
        const int v = dag.add(block);
        std::set<int> conn;
        for (const bh_base *b : block.getAllBases()) {
            auto &vs = base2v[b];
            conn.insert(vs.begin(), vs.end());
            vs.insert(v);
        }
        for (int u : conn)
            if (u != v && block.dependOn(dag.node[u])) dag.add_edge(u, v);
        if (!block.isInstr()) {
            std::set<int> conn_freed;
            for (const bh_base *b : block.getLoop().getAllFrees()) {
                auto &vs = base2v[b];
                conn_freed.insert(vs.begin(), vs.end());
                vs.insert(v);
            }
            for (int u : conn_freed)
                if (u != v) dag.add_edge(u, v);
        }
    }
    return dag;
}

std::vector<Block> topological_fuse(Dag &dag, const FusionConfig &config, bool reshapable_first) {
    std::vector<int> indeg(dag.node.size(), 0);
    for (size_t i = 0; i < dag.node.size(); ++i)
        for (int c : dag.child[i]) ++indeg[c];
    std::deque<int> roots;
    for (size_t i = 0; i < dag.node.size(); ++i)
        if (indeg[i] == 0) roots.push_back(static_cast<int>(i));
    auto pop = [&](std::deque<int> &q) {
        if (reshapable_first) {
            for (auto it = q.rbegin(); it != q.rend(); ++it) {
                if (dag.node[*it].isReshapable()) {
                    const int v = *it;
                    q.erase(std::next(it).base());
                    return v;
                }
            }
        }
        const int v = q.front();
        q.pop_front();
        return v;
    };
    auto release = [&](int v, std::deque<int> &q) {
        for (int c : dag.child[v])
            if (--indeg[c] == 0) q.push_back(c);
    };
    std::vector<Block> ret;
    while (!roots.empty()) {
        const int v = pop(roots);
        ret.push_back(dag.node[v]);
        Block &block = ret.back();
        release(v, roots);
        if (block.isInstr()) continue;
        std::deque<int> nonfusible;
        while (!roots.empty()) {
            const int u = pop(roots);
            if (!dag.node[u].isInstr() && config.can_merge(block, dag.node[u])) {
                block = reshape_and_merge(block.getLoop(), dag.node[u].getLoop());
                release(u, roots);
            } else {
                nonfusible.push_back(u);
            }
        }
        roots = std::move(nonfusible);
    }
    return ret;
}
} // namespace

void fuser_greedy(const FusionConfig &config, std::vector<Block> &block_list, const MergeWeight &weight_fn) {
    Dag dag = from_block_list(block_list);
    if (dag.num_edges() > config.greedy_threshold) {
        block_list = topological_fuse(dag, config, true);
        for (Block &b : block_list)
            if (!b.isInstr()) fuser_greedy(config, b.getLoop()._block_list, weight_fn);
        return;
    }
    while (true) {
        std::vector<std::pair<int, int> > fusibles;
        for (size_t a = 0; a < dag.node.size(); ++a) {
            if (!dag.alive[a]) continue;
            for (int b : std::vector<int>(dag.child[a].begin(), dag.child[a].end())) {
                if (dag.long_path(static_cast<int>(a), b)) {
                    dag.remove_edge(static_cast<int>(a), b);
                } else if (config.can_merge(dag.node[a], dag.node[b])) {
                    fusibles.emplace_back(static_cast<int>(a), b);
                }
            }
        }
        if (fusibles.empty()) break;
        std::pair<int, int> best = fusibles.front();
        double best_w = weight_fn(dag.node[best.first], dag.node[best.second]);
        for (const auto &e : fusibles) {
            const double w = weight_fn(dag.node[e.first], dag.node[e.second]);
            if (w > best_w) {
                best = e;
                best_w = w;
            }
        }
        dag.merge(best.first, best.second);
    }
    block_list = dag.topo();
    for (Block &b : block_list)
        if (!b.isInstr()) fuser_greedy(config, b.getLoop()._block_list, weight_fn);
}

bool fuser_greedy_learned(const FusionConfig &config, std::vector<Block> &block_list) {
    if (!config.learned_cost) return false;
    Dag dag = from_block_list(block_list);
    if (dag.num_edges() > config.greedy_threshold) return false;
    std::vector<double> node_cost(dag.node.size(), -1.0);
    bool first = true;
    while (true) {
        std::vector<std::pair<int, int> > fusibles;
        for (size_t a = 0; a < dag.node.size(); ++a) {
            if (!dag.alive[a]) continue;
            for (int b : std::vector<int>(dag.child[a].begin(), dag.child[a].end())) {
                if (dag.long_path(static_cast<int>(a), b)) {
                    dag.remove_edge(static_cast<int>(a), b);
                } else if (config.can_merge(dag.node[a], dag.node[b])) {
                    fusibles.emplace_back(static_cast<int>(a), b);
                }
            }
        }
        if (fusibles.empty()) break;
        // One cost query per round: every node without a cached cost, then every candidate union.
        std::vector<std::vector<InstrPtr> > sets;
        std::vector<int> node_of_set;
        std::set<int> need;
        for (const auto &e : fusibles) {
            if (node_cost[e.first] < 0) need.insert(e.first);
            if (node_cost[e.second] < 0) need.insert(e.second);
        }
        for (int v : need) {
            sets.push_back(allInstr(dag.node[v]));
            node_of_set.push_back(v);
        }
        for (const auto &e : fusibles) {
            std::vector<InstrPtr> u = allInstr(dag.node[e.first]);
            const std::vector<InstrPtr> b = allInstr(dag.node[e.second]);
            u.insert(u.end(), b.begin(), b.end());
            sets.push_back(std::move(u));
        }
        std::vector<double> cost;
        if (!config.learned_cost(sets, cost) || cost.size() != sets.size()) {
            if (first) return false;
            throw std::runtime_error("learned fuser: the cost model failed");
        }
        first = false;
        for (size_t k = 0; k < node_of_set.size(); ++k) node_cost[node_of_set[k]] = cost[k];
        double best_w = 0.0;
        int best = -1;
        for (size_t k = 0; k < fusibles.size(); ++k) {
            const auto &e = fusibles[k];
            const double w = node_cost[e.first] + node_cost[e.second] - cost[node_of_set.size() + k];
            if (w > best_w) {
                best_w = w;
                best = static_cast<int>(k);
            }
        }
        if (best < 0) break;  // no merge is predicted to save time
        dag.merge(fusibles[best].first, fusibles[best].second);
        node_cost[fusibles[best].first] = -1.0;
    }
    block_list = dag.topo();
    const MergeWeight analytic = [](const Block &a, const Block &b) { return static_cast<double>(weight(a, b)); };
    for (Block &b : block_list)
        if (!b.isInstr()) fuser_greedy(config, b.getLoop()._block_list, analytic);
    return true;
}

// ---------------------------------------------------------------- transformers

namespace {
std::vector<InstrPtr> swap_axis(const std::vector<InstrPtr> &list, int64_t a1, int64_t a2) {
    std::vector<InstrPtr> ret;
    for (const InstrPtr &i : list) {
        bh_instruction t(*i);
        t.transpose(a1, a2);
        ret.push_back(std::make_shared<bh_instruction>(t));
    }
    return ret;
}

// The parent is split into one loop per sub-block; the swapped child is rebuilt from its
// instructions, one nested loop per instruction of higher rank.  The child's own frees stay
// frees of the rebuilt loop (their bases are made and used at the child's level); the frees of
// its nested loops are kept only as releases, since the instructions that made and read those
// bases may now sit in different nested loops.  The parent's own frees go to the last loop of
// the split (after every use inside the parent).
std::vector<Block> swap_blocks(const LoopB &parent, const LoopB *child) {
    std::vector<Block> ret;
    for (const Block &b : parent._block_list) {
        LoopB loop;
        loop.rank = parent.rank;
        if (b.isInstr() || &b.getLoop() != child) {
            loop.size = parent.size;
            loop._block_list.push_back(b);
        } else {
            loop.size = child->size;
            const auto t = swap_axis(allInstr(*child), parent.rank, child->rank);
            loop._block_list.push_back(create_nested_block(t, child->rank, parent.size));
            loop._frees = child->_frees;
            for (bh_base *f : child->getAllFrees())
                if (!child->_frees.count(f)) loop._released.insert(f);
        }
        ret.emplace_back(std::move(loop));
    }
    if (!ret.empty()) ret.back().getLoop()._frees.insert(parent._frees.begin(), parent._frees.end());
    for (Block &b : ret) b.getLoop().metadataUpdate();
    return ret;
}

const LoopB *find_swappable_sub_block(const LoopB &parent) {
    for (const InstrPtr &sweep : parent._sweeps) {
        for (const Block &b : parent._block_list) {
            if (b.isInstr()) continue;
            for (const Block &ib : b.getLoop()._block_list)
                if (ib.isInstr() && *ib.getInstr() == *sweep) return &b.getLoop();
        }
    }
    return nullptr;
}

bool collapse_instr_axes(LoopB &loop, int axis) {
    for (Block &block : loop._block_list) {
        if (block.isInstr()) {
            bh_instruction instr(*block.getInstr());
            const int sa = instr.sweep_axis();
            if (sa == axis || sa == axis + 1) return false;
            for (size_t i = 0; i < instr.operand.size(); ++i) {
                bh_view &v = instr.operand[i];
                if (v.isConstant()) continue;
                int ax = axis;
                if (i == 0 && bh_opcode_is_reduction(instr.opcode)) ax = sa < ax ? ax - 1 : ax;
                if (v.ndim <= ax + 1) return false;
                if (v.shape[ax + 1] * v.stride[ax + 1] != v.stride[ax]) return false;
                v.shape[ax] *= v.shape[ax + 1];
                v.stride[ax] = v.stride[ax + 1];
            }
            instr.remove_axis(axis + 1);
            block.setInstr(instr);
        } else {
            --block.getLoop().rank;
            if (!collapse_instr_axes(block.getLoop(), axis)) return false;
        }
    }
    loop.metadataUpdate();
    return true;
}

bool collapse_loop_with_child(LoopB &loop) {
    if (loop._sweeps.empty() && loop._block_list.size() == 1 && loop._reshapable) {
        Block &child = loop._block_list[0];
        if (!child.isInstr() && child.getLoop()._sweeps.empty() && child.getLoop()._reshapable) {
            loop.size *= child.getLoop().size;
            loop._frees.insert(child.getLoop()._frees.begin(), child.getLoop()._frees.end());
            loop._released.insert(child.getLoop()._released.begin(), child.getLoop()._released.end());
            auto tmp = std::move(child.getLoop()._block_list);
            loop._block_list = std::move(tmp);
            return collapse_instr_axes(loop, loop.rank);
        }
    }
    return false;
}
} // namespace

void push_reductions_inwards(std::vector<Block> &block_list) {
    std::vector<Block> list2(block_list);
    for (Block &b : list2)
        if (!b.isInstr()) push_reductions_inwards(b.getLoop()._block_list);
    std::vector<Block> ret;
    for (const Block &b : list2) {
        const LoopB *sw = nullptr;
        if (!b.isInstr() && (sw = find_swappable_sub_block(b.getLoop())) != nullptr) {
            const auto t = swap_blocks(b.getLoop(), sw);
            ret.insert(ret.end(), t.begin(), t.end());
        } else {
            ret.push_back(b);
        }
    }
    block_list = ret;
}

void collapse_redundant_axes(std::vector<Block> &block_list) {
    for (Block &b : block_list)
        if (!b.isInstr()) collapse_redundant_axes(b.getLoop()._block_list);
    std::vector<Block> ret;
    for (const Block &block : block_list) {
        if (!block.isInstr()) {
            Block b(block);
            if (collapse_loop_with_child(b.getLoop())) {
                ret.push_back(std::move(b));
                continue;
            }
        }
        ret.push_back(block);
    }
    block_list = ret;
}

// ---------------------------------------------------------------- kernels

namespace {
std::vector<InstrPtr> order_sweeps(const std::set<InstrPtr> &sweeps) {
    std::vector<InstrPtr> ret(sweeps.begin(), sweeps.end());
    std::sort(ret.begin(), ret.end(), [](const InstrPtr &a, const InstrPtr &b) { return a->origin_id > b->origin_id; });
    return ret;
}

bh_instruction identity_for(const InstrPtr &sweep, int64_t &origin_count) {
    bh_instruction id(BH_IDENTITY, {sweep->operand[0]});
    id.operand.resize(2);
    id.operand[1].base = nullptr;
    id.constant = sweep_identity(sweep->opcode, sweep->operand[0].base->dtype());
    id.origin_id = origin_count++;
    id.constructor = sweep->constructor;
    if (bh_opcode_is_accumulate(sweep->opcode)) id.operand[0].shape[sweep->sweep_axis()] = 1;
    return id;
}

// Insert identity writers before every nested sweeping loop (reference apply_fusion.cpp:87-131).
void add_identity_block(LoopB &loop, int64_t &origin_count) {
    std::vector<Block> ret;
    for (Block &block : loop._block_list) {
        if (block.isInstr()) {
            ret.push_back(block);
            continue;
        }
        add_identity_block(block.getLoop(), origin_count);
        for (const InstrPtr &sw : order_sweeps(block.getLoop().getSweeps())) {
            bh_instruction id = identity_for(sw, origin_count);
            if (loop.rank == -1 && sw->operand[0].is_scalar()) {
                ret.emplace_back(id, 0);
            } else if (loop.rank == sw->operand[0].ndim - 1) {
                ret.emplace_back(id, static_cast<int>(sw->operand[0].ndim));
            } else {
                ret.push_back(create_nested_block({std::make_shared<const bh_instruction>(id)}, loop.rank + 1));
            }
            bh_instruction upd(*sw);
            upd.constructor = false;
            block.getLoop().replaceInstr(sw, upd);
            block.getLoop().metadataUpdate();
        }
        ret.push_back(block);
    }
    loop._block_list = ret;
    loop.metadataUpdate();
}
} // namespace

std::vector<LoopB> kernels_from_blocks(std::vector<Block> block_list, int64_t &origin_count) {
    std::vector<LoopB> ret;
    for (Block &block : block_list) {
        if (block.isInstr()) {
            ret.emplace_back(-1, 1, std::vector<Block>{block});
            continue;
        }
        add_identity_block(block.getLoop(), origin_count);
        LoopB kernel(-1, 1);
        for (const InstrPtr &sw : order_sweeps(block.getLoop().getSweeps())) {
            bh_instruction id = identity_for(sw, origin_count);
            if (sw->operand[0].is_scalar()) {
                kernel._block_list.emplace_back(id, 0);
            } else {
                kernel._block_list.push_back(create_nested_block({std::make_shared<const bh_instruction>(id)}, 0));
            }
            bh_instruction upd(*sw);
            upd.constructor = false;
            block.getLoop().replaceInstr(sw, upd);
            block.getLoop().metadataUpdate();
        }
        kernel._block_list.push_back(block);
        kernel.metadataUpdate();
        ret.push_back(std::move(kernel));
    }
    return ret;
}

std::vector<Block> get_block_list(const std::vector<bh_instruction *> &instr_list, const FusionConfig &config) {
    std::vector<Block> block_list;
    if (config.pre_fuser == "none" || config.pre_fuser == "singleton") {
        block_list = fuser_singleton(instr_list);
    } else if (config.pre_fuser == "lossy" || config.pre_fuser == "pre_fuser_lossy") {
        block_list = pre_fuser_lossy(instr_list, config.accumulate_alone, config.one_sweep_axis);
    } else {
        throw std::runtime_error("Unknown pre-fuser: " + config.pre_fuser);
    }
    const MergeWeight analytic = [](const Block &a, const Block &b) { return static_cast<double>(weight(a, b)); };
    for (const std::string &t : config.fuser_list) {
        if (t == "greedy") {
            fuser_greedy(config, block_list, analytic);
        } else if (t == "learned") {
            // The model scores whole partitions: the analytic greedy one, the one its own greedy
            // merging builds and the fully merged (serial) one.  Another partition replaces the
            // analytic one only when its predicted total is at least 10 % lower (the model's
            // per-kernel error is about 30 %; smaller predicted gains were noise when measured).
            // Without a model: the analytic greedy fuser.
            std::vector<Block> greedy = block_list;
            fuser_greedy(config, greedy, config.learned_weight ? config.learned_weight : analytic);
            std::vector<Block> learned = block_list;
            if (!fuser_greedy_learned(config, learned)) {
                block_list = std::move(greedy);
                continue;
            }
            std::vector<Block> serial = block_list;
            fuser_serial(serial, config);
            std::vector<std::vector<Block> *> cands{&greedy, &learned, &serial};
            std::vector<std::vector<InstrPtr> > sets;
            std::vector<size_t> owner;
            for (size_t c = 0; c < cands.size(); ++c)
                for (const Block &b : *cands[c]) {
                    std::vector<InstrPtr> in = b.isInstr() ? std::vector<InstrPtr>{b.getInstr()} : allInstr(b.getLoop());
                    std::vector<InstrPtr> work;
                    for (const InstrPtr &i : in)
                        if (!bh_opcode_is_system(i->opcode)) work.push_back(i);
                    if (work.empty()) continue;
                    sets.push_back(std::move(work));
                    owner.push_back(c);
                }
            std::vector<double> cost;
            size_t pick = 0;
            if (config.learned_cost(sets, cost) && cost.size() == sets.size()) {
                std::vector<double> total(cands.size(), 0.0);
                for (size_t k = 0; k < sets.size(); ++k) total[owner[k]] += cost[k];
                double best = 0.9 * total[0];
                for (size_t c = 1; c < cands.size(); ++c)
                    if (total[c] < best) {
                        best = total[c];
                        pick = c;
                    }
            }
            block_list = std::move(*cands[pick]);
        } else if (t == "serial") {
            fuser_serial(block_list, config);
        } else if (t == "breadth_first") {
            Dag dag = from_block_list(block_list);
            block_list = topological_fuse(dag, config, false);
        } else if (t == "reshapable_first") {
            Dag dag = from_block_list(block_list);
            block_list = topological_fuse(dag, config, true);
        } else if (t == "push_reductions_inwards") {
            push_reductions_inwards(block_list);
        } else if (t == "collapse_redundant_axes") {
            collapse_redundant_axes(block_list);
        } else if (t == "split_for_threading") {
            // Not needed: reductions are cooperative on the device, so no block is split for threading.
        } else {
            throw std::runtime_error("Unknown transformer: " + t);
        }
    }
    return block_list;
}

std::vector<LoopB> get_kernel_list(const std::vector<bh_instruction *> &instr_list, const FusionConfig &config) {
    int64_t origin_count = 0;
    for (bh_instruction *i : instr_list) i->origin_id = origin_count++;
    return kernels_from_blocks(get_block_list(instr_list, config), origin_count);
}

std::vector<bh_instruction *> remove_non_computed_system_instr(std::vector<bh_instruction> &instr_list,
                                                              std::set<bh_base *> &frees) {
    std::vector<bh_instruction *> ret;
    ret.reserve(instr_list.size());
    // (flat hash sets: flushes of tens of thousands of instructions pass through here)
    FlatPtrMap computes(instr_list.size());
    for (bh_instruction &instr : instr_list) {
        if (instr.opcode == BH_FREE && computes.find(instr.operand[0].base) == nullptr) {
            frees.insert(instr.operand[0].base);
        } else if (!(instr.opcode == BH_NONE || instr.opcode == BH_TALLY)) {
            for (const bh_view &v : instr.operand)
                if (!v.isConstant()) computes.insert(v.base, 0);
            ret.push_back(&instr);
        }
    }
    return ret;
}

void set_constructor_flag(std::vector<bh_instruction *> &instr_list) {
    FlatPtrMap constructed(instr_list.size());
    for (bh_instruction *instr : instr_list) {
        instr->constructor = false;
        for (size_t o = 0; o < instr->operand.size(); ++o) {
            const bh_view &v = instr->operand[o];
            if (v.isConstant()) continue;
            if (constructed.insert(v.base, 0).second && o == 0) instr->constructor = true;
        }
    }
}

} // namespace jitk
} // namespace bohrium
