/*
 * kmc_internal.h -- helpers shared by the host-side modules of libkmc.
 */
#pragma once
#include <stddef.h>

namespace kmc {

/* Caching device allocator: blocks return to a free list and are reused. */
void* device_alloc(size_t bytes);
void device_free(void* p);
void device_pool_trim();

}  // namespace kmc
