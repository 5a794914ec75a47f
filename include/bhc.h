/* bhc — C-ABI bridge that records bytecode and flushes it down the component
 * stack (the bridge side of the boundary; the engine side is bh_ve_hip.h).
 *
 * It is this build's counterpart of the reference's C bridge and its
 * generic py_api entry points, which call the same bhxx runtime:
 *   bhc_new            <- bhc_new_A<type> / _bh_api.new   (bridge/c/gen_array_operations.py:26-106,
 *                                                          bridge/py_api/src/_bh_api.c:617-630)
 *   bhc_op             <- bhc_<op>_A<type>... / _bh_api.op (the generic opcode entry; one call per instruction)
 *   bhc_random123      <- bhc_random123_A<type>           (bridge/cxx/src/Runtime.cpp:44-56)
 *   bhc_free           <- base deleter -> BH_FREE          (bridge/cxx/src/Runtime.cpp:58-70)
 *   bhc_sync           <- bhc_sync                         (bridge/cxx Runtime::sync)
 *   bhc_flush          <- bhc_flush                        (bridge/cxx/src/Runtime.cpp:74-101)
 *   bhc_flush_and_repeat <- bhc_flush_and_repeat           (Runtime::flushAndRepeat, bh_ir.hpp:38-52)
 *   bhc_data_get       <- bhc_data_get                     (getMemoryPointer, bh_component.hpp:188-199)
 *   bhc_data_set       <- bhc_data_set                     (setMemoryPointer, bh_component.hpp:201-211)
 *   bhc_message        <- bhc_message                      (message, bh_component.hpp:180-186)
 *   bhc_extmethod      <- bhc_extmethod                    (Runtime::enqueueExtmethod, Runtime.hpp:243-261)
 *   bhc_user_kernel    <- bhc_user_kernel                  (bridge/c/gen_specials.py:286-304,
 *                                                          Runtime::userKernel, bridge/cxx/src/Runtime.cpp:131-138)
 *
 * Errors: functions return 0 (or a non-null handle) on success; on failure
 * they return non-zero (or NULL) and bhc_last_error() holds the message of the
 * C++ exception raised below the boundary.  Not thread-safe (as the reference
 * runtime, bridge/cxx/include/bhxx/Runtime.hpp:30-35).
 */
#ifndef BHC_H
#define BHC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BHC_MAXDIM 16

typedef void *bhc_base_t;  /* an opaque bh_base* owned by the bridge */

typedef struct {
    bhc_base_t base;       /* NULL: this operand is the instruction's constant */
    int64_t start;
    int64_t ndim;
    int64_t shape[BHC_MAXDIM];
    int64_t stride[BHC_MAXDIM];
} bhc_view_t;

/* Load the component stack named by BH_STACK from the config file (BH_CONFIG,
 * ~/.bohrium/config.ini, /etc/bohrium, else `fallback_config`). */
int bhc_init(const char *fallback_config);
int bhc_shutdown(void);
const char *bhc_last_error(void);

bhc_base_t bhc_new(int dtype, int64_t nelem);
/* Queue one instruction; a constant operand (base == NULL) takes its value
 * from `const_value` interpreted as `const_type` (a bh_type id). */
int bhc_op(int64_t opcode, int nops, const bhc_view_t *views, int const_type, const void *const_value);
int bhc_random123(const bhc_view_t *out, uint64_t start, uint64_t key);
/* Make operand `operand` of the instruction queued last slide between the
 * repeats of bhc_flush_and_repeat: its start moves by offset_change * stride
 * elements per step (wrapping inside stride * shape) and dimension `rank` grows
 * by shape_change, every `step_delay` repeats.  Same fields as the reference's
 * bh_slide_dim (include/bohrium/bh_slide.hpp) set by _bh.slide_view
 * (bridge/npbackend/bohrium/loop.py:522-548). */
int bhc_op_slide(int operand, int64_t rank, int64_t offset_change, int64_t shape_change, int64_t stride,
                 int64_t shape, int64_t step_delay);
/* Queue BH_FREE; the base handle stays valid until the next flush has run. */
int bhc_free(bhc_base_t base);
int bhc_sync(bhc_base_t base);
int bhc_flush(void);
int bhc_flush_and_repeat(uint64_t nrepeats, bhc_base_t condition);
void *bhc_data_get(bhc_base_t base, int copy2host, int force_alloc, int nullify);
int bhc_data_set(bhc_base_t base, int host_ptr, void *data);
/* Give `base` a new host buffer holding a copy of `src` (nbytes == the base's size). */
int bhc_data_copy_in(bhc_base_t base, const void *src, int64_t nbytes);
/* Copy the base's current contents to `dst` (flushes nothing; call bhc_sync + bhc_flush first). */
int bhc_data_copy_out(bhc_base_t base, void *dst, int64_t nbytes);
const char *bhc_message(const char *msg);
/* Register extension method `name` and queue one instruction of it. */
int bhc_extmethod(const char *name, int nops, const bhc_view_t *views);
/* Compile and run a user kernel on the operands (`tag` selects the component, "hip" for this engine).
 * Returns "" on success, else the error message; the string stays valid until the next call.  Returns
 * NULL only when the call itself failed (bhc_last_error()).  The caller flushes first, as the reference
 * front ends do (bridge/bh107/bh107/user_kernel.py:88-90). */
const char *bhc_user_kernel(const char *kernel, int nops, const bhc_view_t *operands, const char *compile_cmd,
                            const char *tag, const char *param);
/* Recorded flushes (no reference C-ABI counterpart; the format restates what the reference's
 * BhIR serializer writes, core/bh_ir.cpp:37-154, without Boost; see core/bh_ir.cpp here).
 * bhc_record_flushes: flush what is queued, then append every later flush to `path` (a new file)
 * as a BhIR archive plus the data of the bases it introduces that hold host data; NULL or "" stops.
 * bhc_replay_file: execute the flushes recorded in `path` through the stack, in order (what this
 * process queued is flushed first).  bhc_replay_base: the base that stands in for the one the
 * recording process had at `recorded_address` (NULL when none, or when it was freed); it belongs
 * to the replay: read it (bhc_sync + bhc_flush, bhc_data_copy_out), never bhc_free it. */
int bhc_record_flushes(const char *path);
int bhc_replay_file(const char *path);
bhc_base_t bhc_replay_base(uint64_t recorded_address);
/* Auto-flush after this many queued instructions (default 1000 like the reference; <= 0 disables). */
int bhc_set_flush_threshold(int64_t n);
int64_t bhc_base_nelem(bhc_base_t base);
int bhc_base_dtype(bhc_base_t base);

#ifdef __cplusplus
}
#endif

#endif
