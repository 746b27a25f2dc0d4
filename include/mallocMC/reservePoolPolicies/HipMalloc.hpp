/*
 * Reserve-pool policy: the heap's memory comes from hipMalloc on the chosen
 * device (the reference's AlpakaBuf, src/include/mallocMC/reservePoolPolicies/AlpakaBuf.hpp:37-61,
 * allocates an alpaka buffer, which is hipMalloc on its HIP backend).
 * AlpakaBuf<Acc> is kept as a name so existing policy lists compile.
 */
#pragma once

#include "../mallocMC_utils.hpp"

#include <hip/hip_runtime.h>

#include <cstddef>
#include <stdexcept>
#include <string>

namespace mallocMC
{
    namespace ReservePoolPolicies
    {
        struct HipMalloc
        {
            void* setMemPool(int device, std::size_t memsize)
            {
                if(hipSetDevice(device) != hipSuccess)
                    throw std::runtime_error("hipSetDevice failed");
                void* p = nullptr;
                if(hipMalloc(&p, memsize) != hipSuccess)
                    throw std::runtime_error("hipMalloc of the heap pool failed");
                pool = p;
                return p;
            }

            void resetMemPool(void* p)
            {
                if(p)
                    (void) hipFree(p);
                pool = nullptr;
            }

            static std::string classname()
            {
                return "HipMalloc";
            }

        private:
            void* pool = nullptr;
        };

        template<class T_Acc = AccGfx950>
        struct AlpakaBuf : HipMalloc
        {
            static std::string classname()
            {
                return "AlpakaBuf";
            }
        };
    } // namespace ReservePoolPolicies
} // namespace mallocMC
