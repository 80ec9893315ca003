/*
 * kmc_guard.h -- no C++ exception crosses the C ABI: each entry point runs its
 * body through abi_guard, which turns a failed host allocation of memory
 * (std::bad_alloc) or of a thread (std::system_error) into
 * KMC_ERR_OUT_OF_MEMORY and any other exception into KMC_ERR_INVALID_ARGUMENT.
 */
#pragma once

#include <new>
#include <system_error>

#include "kmc.h"

namespace kmc {

template <typename F>
kmc_status abi_guard(F&& body) noexcept
{
    try {
        return body();
    } catch (const std::bad_alloc&) {
        return KMC_ERR_OUT_OF_MEMORY;
    } catch (const std::system_error&) {
        return KMC_ERR_OUT_OF_MEMORY;
    } catch (...) {
        return KMC_ERR_INVALID_ARGUMENT;
    }
}

}  // namespace kmc
