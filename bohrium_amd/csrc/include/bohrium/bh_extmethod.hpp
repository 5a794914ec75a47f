/* Extension-method plugin interface.
 *
 * Mirrors include/bohrium/bh_extmethod.hpp:57-125 and core/bh_extmethod.cpp:33-90
 * of the reference: a library in the engine's `libs` list exports
 * `<name>_create()` / `<name>_destroy()`, and the created object's
 * `execute(instr, arg)` runs one extension instruction (`arg` is the engine).
 */
#pragma once

#include <string>
#include <vector>
#include <bohrium/bh_instruction.hpp>

namespace bohrium {
namespace extmethod {

class ExtmethodImpl {
public:
    virtual ~ExtmethodImpl() = default;
    virtual void execute(bh_instruction *instr, void *arg) = 0;
};

class ExtmethodFace {
    void *_lib_handle = nullptr;
    ExtmethodImpl *_implementation = nullptr;
    void (*_destroy)(ExtmethodImpl *) = nullptr;
    std::string _name;

public:
    /* Search `libs` for `<name>_create`; throws std::out_of_range when not found. */
    ExtmethodFace(const std::vector<std::string> &libs, const std::string &name);
    ExtmethodFace(const ExtmethodFace &) = delete;
    ExtmethodFace(ExtmethodFace &&other) noexcept;
    ~ExtmethodFace();
    void execute(bh_instruction *instr, void *arg);
    const std::string &name() const { return _name; }
};

} // namespace extmethod
} // namespace bohrium
