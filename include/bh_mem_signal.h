/* bh_mem_signal — call a function when the program touches a registered memory segment.
 *
 * The same C interface as the reference's memory-signal library
 * (include/bohrium/bh_mem_signal.h, core/bh_mem_signal.cpp:121-196), which the
 * NumPy bridge uses to copy an array back from Bohrium when NumPy code touches
 * its protected host pages (bridge/npbackend/src/memory.c:148-283).  The
 * reference dispatches through libsigsegv; this one installs a POSIX SIGSEGV
 * handler (sigaction, SA_SIGINFO) and chains faults outside every registered
 * segment to the handler that was installed before it.  Exported by libbh.so.
 *
 * The hip engine uses it itself to keep arrays resident on the device after a
 * sync: the host copy is handed out read-only, and the first host write to it
 * makes the host copy the current one (see host_write_tracking in config.ini).
 *
 * The caller protects the memory (mprotect); this library only dispatches.
 * A callback runs in signal context on the faulting thread and must return
 * non-zero when it made the access legal (the access is then retried).
 */
#ifndef BH_MEM_SIGNAL_H
#define BH_MEM_SIGNAL_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Called with the faulting address and the `idx` given at attach time (reference bh_mem_signal.h:34). */
typedef int (*bh_mem_signal_callback_t)(void *fault_address, void *segment_idx);

/* Install the SIGSEGV handler (idempotent; reference bh_mem_signal_init). */
void bh_mem_signal_init(void);
/* Remove the handler, restoring the one installed before it (reference bh_mem_signal_shutdown). */
void bh_mem_signal_shutdown(void);
/* Register [addr, addr + size) with its callback.  Overlapping an attached segment is an error:
 * returns -1 (the reference throws std::runtime_error); 0 on success. */
int bh_mem_signal_attach(void *idx, void *addr, uint64_t size, bh_mem_signal_callback_t callback);
/* Unregister the segment that contains `addr` (no-op when there is none; reference bh_mem_signal_detach). */
void bh_mem_signal_detach(const void *addr);
/* Non-zero when `addr` lies in an attached segment (reference bh_mem_signal_exist). */
int bh_mem_signal_exist(const void *addr);
/* Print the attached segments to stdout (reference bh_mem_signal_pprint_db). */
void bh_mem_signal_pprint_db(void);

#ifdef __cplusplus
}
#endif

#endif
