/* Host memory of base arrays.
 *
 * bh_data_malloc / bh_data_free follow the reference's ownership rule
 * (core/bh_main_memory.cpp:110-131): the data pointer of a base is allocated on
 * demand, owned by the runtime, and released when the engine executes BH_FREE.
 * Buffers are page aligned (mmap) and recycled through a size-keyed cache, as
 * the reference's MallocCache does (include/bohrium/bh_malloc_cache.hpp:33-257).
 */
#pragma once

#include <cstdint>
#include <bohrium/bh_base.hpp>

/* Allocate the host buffer of `base` if it has none. */
void bh_data_malloc(bh_base *base);
/* Release the host buffer of `base` (if any) and reset its data pointer. */
void bh_data_free(bh_base *base);
/* Raw page-aligned allocation through the same cache. */
void *bh_memory_malloc(uint64_t nbytes);
void bh_memory_free(void *mem, uint64_t nbytes);
/* True when `mem` is a live bh_memory_malloc allocation of at least `nbytes` bytes: its pages
 * (a mapping of their own, page aligned) belong to it alone. */
bool bh_memory_owns(const void *mem, uint64_t nbytes);
/* Limit (bytes) of memory kept in the reuse cache. */
void bh_set_malloc_cache_limit(uint64_t nbytes);
