/* bh_ve_hip — the drop-in boundary of the hip vector engine.
 *
 * libbh_ve_hip.so is a Bohrium component: the parent component (normally the
 * `node` VEM, or the bridge when the engine is first in the stack) loads it
 * with dlopen and resolves exactly these two symbols, as the reference loader
 * does (core/bh_component.cpp:30-61, include/bohrium/bh_component.hpp:51-116).
 * Select it with a config stack whose engine entry is `hip`, e.g.
 *     [stacks]  hip = hip        (or: bcexp_gpu, bccon, node, hip)
 *     [hip]     impl = <path>/libbh_ve_hip.so
 *
 * The returned object implements the ComponentImpl virtual interface
 * (execute, extmethod, message, getMemoryPointer, setMemoryPointer, memCopy,
 * get/setDeviceContext, userKernel) declared in
 * bohrium_amd/csrc/include/bohrium/bh_component.hpp, which restates
 * include/bohrium/bh_component.hpp:120-264 of the reference without Boost.
 */
#ifndef BH_VE_HIP_H
#define BH_VE_HIP_H

#ifdef __cplusplus
namespace bohrium {
namespace component {
class ComponentImpl;
}
}
typedef bohrium::component::ComponentImpl bh_component_impl;
extern "C" {
#else
typedef struct bh_component_impl bh_component_impl;
#endif

/* What the engine hands an extension method as the `arg` of ExtmethodImpl::execute(instr, arg).
 * The reference GPU engines pass the engine object itself (core engine_gpu.hpp:177); this engine
 * passes a plain C view of what a device method needs.  Every operand of the instruction is
 * resident on the device when execute() is called, and operand 0 counts as written afterwards. */
typedef struct bh_hip_ext_context {
    int version;     /* 1 */
    int dry_run;     /* compile-only engine: check the operands, launch nothing */
    void *engine;
    /* Device address of a bh_base's buffer (element 0 of the base). */
    void *(*device_pointer)(void *engine, const void *base);
    void *stream;    /* hipStream_t every engine launch is ordered on */
} bh_hip_ext_context;

/* Create the engine at `stack_level` (reads its [hip] config section). */
bh_component_impl *create(int stack_level);
/* Destroy an engine created by create(); device memory is released. */
void destroy(bh_component_impl *self);

#ifdef __cplusplus
}
#endif

#endif
