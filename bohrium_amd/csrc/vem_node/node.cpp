/* The `node` VEM: a pass-through manager between filters and the engine.
 * Like the reference (vem/node/main.cpp:56-95) it forwards every call to its
 * child; with BH_MEM_WARN set it reports bases that are freed without ever
 * having been computed on. */
#include <bohrium/bh_component.hpp>

#include <cstdlib>
#include <iostream>
#include <set>

namespace {
class Node : public bohrium::component::ComponentImpl {
    std::set<bh_base *> seen;
    const bool mem_warn = std::getenv("BH_MEM_WARN") != nullptr;

public:
    explicit Node(int stack_level) : ComponentImpl(stack_level) {}
    void execute(BhIR *bhir) override {
        if (mem_warn) {
            for (const bh_instruction &i : bhir->instr_list) {
                if (i.opcode == BH_FREE) {
                    if (!seen.erase(i.operand[0].base))
                        std::cerr << "[NODE-VEM] freeing a base that was never used\n";
                } else {
                    for (const bh_view &v : i.operand)
                        if (!v.isConstant()) seen.insert(v.base);
                }
            }
        }
        child.execute(bhir);
    }
};
} // namespace

extern "C" bohrium::component::ComponentImpl *create(int stack_level) { return new Node(stack_level); }
extern "C" void destroy(bohrium::component::ComponentImpl *self) { delete self; }
