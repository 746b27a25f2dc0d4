// Host-side heap object behind the opaque mmc_heap handle of the C ABI.
#pragma once

#include "mallocmc_hip.h"

#include "mallocMC/detail/scatter_heap.hpp"

#include <hip/hip_runtime.h>

#include <string>

struct mmc_heap
{
    mmc_config cfg{};
    int device = 0;
    void* rawPool = nullptr; // as returned by hipMalloc (nullptr for a caller-owned pool)
    bool ownsPool = true;
    std::uint64_t rawBytes = 0;
    char* pool = nullptr; // after Shrink::alignPool
    std::uint64_t memsize = 0;
    mallocMC::detail::ScatterScalars* scalars = nullptr;
    std::uint32_t* scratch = nullptr; // slot-count staging
    std::uint32_t* masks = nullptr;   // side segment masks (geom.sideMaskWords per page)
    mallocMC::detail::ScatterGeometry geom{};
    mallocMC::detail::ScatterHeap view{};
};

namespace mmc
{
    int fail(int code, const std::string& msg);
    int hipFail(hipError_t e, const char* what);
    unsigned gridFor(std::uint64_t n, unsigned threads);
} // namespace mmc
