/* The component plugin interface: a stack of shared libraries, each exporting
 * `extern "C" ComponentImpl* create(int stack_level)` and
 * `extern "C" void destroy(ComponentImpl*)`.
 *
 * The virtual interface (execute, extmethod, message, getMemoryPointer,
 * setMemoryPointer, memCopy, get/setDeviceContext, userKernel) and the
 * pass-to-child defaults follow include/bohrium/bh_component.hpp:51-264 of the
 * reference; the loader follows core/bh_component.cpp:30-61.
 */
#pragma once

#include <map>
#include <set>
#include <string>
#include <vector>
#include <bohrium/bh_config_parser.hpp>
#include <bohrium/bh_extmethod.hpp>
#include <bohrium/bh_ir.hpp>

namespace bohrium {
namespace component {

class ComponentImpl;

class ComponentFace {
    void *_lib_handle = nullptr;
    ComponentImpl *_implementation = nullptr;
    ComponentImpl *(*_create)(int) = nullptr;
    void (*_destroy)(ComponentImpl *) = nullptr;

public:
    ComponentFace(const std::string &lib_path, int stack_level);
    ComponentFace() = default;
    ComponentFace(const ComponentFace &other) = delete;
    ComponentFace(ComponentFace &&other) = delete;
    ~ComponentFace();
    ComponentFace &operator=(ComponentFace &&other) noexcept;

    bool initiated() const { return _implementation != nullptr; }
    void execute(BhIR *bhir);
    void extmethod(const std::string &name, bh_opcode opcode);
    std::string message(const std::string &msg);
    void *getMemoryPointer(bh_base &base, bool copy2host, bool force_alloc, bool nullify);
    void setMemoryPointer(bh_base *base, bool host_ptr, void *mem);
    void memCopy(bh_view &src, bh_view &dst, const std::string &param);
    void *getDeviceContext();
    void setDeviceContext(void *device_context);
    std::string userKernel(const std::string &kernel, std::vector<bh_view> &operand_list,
                           const std::string &compile_cmd, const std::string &tag, const std::string &param);
};

class ComponentImpl {
protected:
    bool disabled = false;

public:
    const int stack_level;
    const ConfigParser config;
    ComponentFace child;

    explicit ComponentImpl(int stack_level, bool initiate_child = true) : stack_level(stack_level), config(stack_level) {
        const std::string child_lib_path = config.getChildLibraryPath();
        if (initiate_child && !child_lib_path.empty()) {
            child = ComponentFace{child_lib_path, stack_level + 1};
        }
    }
    virtual ~ComponentImpl() = default;

    virtual void execute(BhIR *bhir) { child.execute(bhir); }
    virtual void extmethod(const std::string &name, bh_opcode opcode) { child.extmethod(name, opcode); }
    virtual std::string message(const std::string &msg) { return child.message(msg); }
    virtual void *getMemoryPointer(bh_base &base, bool copy2host, bool force_alloc, bool nullify) {
        return child.getMemoryPointer(base, copy2host, force_alloc, nullify);
    }
    virtual void setMemoryPointer(bh_base *base, bool host_ptr, void *mem) {
        child.setMemoryPointer(base, host_ptr, mem);
    }
    virtual void memCopy(bh_view &src, bh_view &dst, const std::string &param) { child.memCopy(src, dst, param); }
    virtual void *getDeviceContext() { return child.getDeviceContext(); }
    virtual void setDeviceContext(void *device_context) { child.setDeviceContext(device_context); }
    virtual std::string userKernel(const std::string &kernel, std::vector<bh_view> &operand_list,
                                   const std::string &compile_cmd, const std::string &tag, const std::string &param) {
        return child.userKernel(kernel, operand_list, compile_cmd, tag, param);
    }
};

class ComponentVE : public ComponentImpl {
public:
    std::map<bh_opcode, extmethod::ExtmethodFace> extmethods;
    std::set<bh_opcode> child_extmethods;
    explicit ComponentVE(int stack_level, bool initiate_child = true) : ComponentImpl(stack_level, initiate_child) {}
};

} // namespace component
} // namespace bohrium
