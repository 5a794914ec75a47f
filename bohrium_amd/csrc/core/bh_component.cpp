#include <bohrium/bh_component.hpp>

#include <dlfcn.h>
#include <iostream>

namespace bohrium {
namespace component {

ComponentFace::ComponentFace(const std::string &lib_path, int stack_level) {
    _lib_handle = dlopen(lib_path.c_str(), RTLD_NOW);
    if (_lib_handle == nullptr) {
        throw std::runtime_error(std::string("ComponentFace: cannot load library: ") + dlerror());
    }
    dlerror();
    *reinterpret_cast<void **>(&_create) = dlsym(_lib_handle, "create");
    if (const char *err = dlerror()) throw std::runtime_error(std::string("ComponentFace: no 'create': ") + err);
    *reinterpret_cast<void **>(&_destroy) = dlsym(_lib_handle, "destroy");
    if (const char *err = dlerror()) throw std::runtime_error(std::string("ComponentFace: no 'destroy': ") + err);
    _implementation = _create(stack_level);
}

ComponentFace::~ComponentFace() {
    if (initiated()) {
        _destroy(_implementation);
        _implementation = nullptr;
        // The library stays mapped: kernels and device state it owns may outlive the stack.
    }
}

ComponentFace &ComponentFace::operator=(ComponentFace &&other) noexcept {
    _lib_handle = other._lib_handle;
    _implementation = other._implementation;
    _create = other._create;
    _destroy = other._destroy;
    other._lib_handle = nullptr;
    other._implementation = nullptr;
    other._create = nullptr;
    other._destroy = nullptr;
    return *this;
}

namespace {
[[noreturn]] void uninitiated() { throw std::runtime_error("uninitiated component interface"); }
} // namespace

void ComponentFace::execute(BhIR *bhir) {
    if (!initiated()) uninitiated();
    _implementation->execute(bhir);
}
void ComponentFace::extmethod(const std::string &name, bh_opcode opcode) {
    if (!initiated()) uninitiated();
    _implementation->extmethod(name, opcode);
}
std::string ComponentFace::message(const std::string &msg) {
    if (!initiated()) uninitiated();
    return _implementation->message(msg);
}
void *ComponentFace::getMemoryPointer(bh_base &base, bool copy2host, bool force_alloc, bool nullify) {
    if (!initiated()) uninitiated();
    return _implementation->getMemoryPointer(base, copy2host, force_alloc, nullify);
}
void ComponentFace::setMemoryPointer(bh_base *base, bool host_ptr, void *mem) {
    if (!initiated()) uninitiated();
    _implementation->setMemoryPointer(base, host_ptr, mem);
}
void ComponentFace::memCopy(bh_view &src, bh_view &dst, const std::string &param) {
    if (!initiated()) uninitiated();
    _implementation->memCopy(src, dst, param);
}
void *ComponentFace::getDeviceContext() {
    if (!initiated()) uninitiated();
    return _implementation->getDeviceContext();
}
void ComponentFace::setDeviceContext(void *device_context) {
    if (!initiated()) uninitiated();
    _implementation->setDeviceContext(device_context);
}
std::string ComponentFace::userKernel(const std::string &kernel, std::vector<bh_view> &operand_list,
                                      const std::string &compile_cmd, const std::string &tag,
                                      const std::string &param) {
    if (!initiated()) uninitiated();
    return _implementation->userKernel(kernel, operand_list, compile_cmd, tag, param);
}

} // namespace component

namespace extmethod {

ExtmethodFace::ExtmethodFace(const std::vector<std::string> &libs, const std::string &name) : _name(name) {
    for (const std::string &lib : libs) {
        void *h = dlopen(lib.c_str(), RTLD_NOW);
        if (h == nullptr) continue;
        dlerror();
        ExtmethodImpl *(*create)();
        *reinterpret_cast<void **>(&create) = dlsym(h, (name + "_create").c_str());
        if (dlerror() != nullptr) continue;
        *reinterpret_cast<void **>(&_destroy) = dlsym(h, (name + "_destroy").c_str());
        if (dlerror() != nullptr) continue;
        _lib_handle = h;
        _implementation = create();
        return;
    }
    throw std::out_of_range("Extension method '" + name + "' not found");
}

ExtmethodFace::ExtmethodFace(ExtmethodFace &&o) noexcept
    : _lib_handle(o._lib_handle), _implementation(o._implementation), _destroy(o._destroy), _name(o._name) {
    o._lib_handle = nullptr;
    o._implementation = nullptr;
    o._destroy = nullptr;
}

ExtmethodFace::~ExtmethodFace() {
    if (_implementation != nullptr && _destroy != nullptr) _destroy(_implementation);
}

void ExtmethodFace::execute(bh_instruction *instr, void *arg) { _implementation->execute(instr, arg); }

} // namespace extmethod
} // namespace bohrium
