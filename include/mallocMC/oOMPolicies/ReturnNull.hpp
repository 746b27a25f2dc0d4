/*
 * OOMPolicies::ReturnNull (src/include/mallocMC/oOMPolicies/ReturnNull.hpp:47-56).
 */
#pragma once

#include "../mallocMC_utils.hpp"

#include <string>

namespace mallocMC
{
    namespace OOMPolicies
    {
        class ReturnNull
        {
        public:
            MALLOCMC_HD static void* handleOOM(void*)
            {
                return nullptr;
            }

            static std::string classname()
            {
                return "ReturnNull";
            }
        };
    } // namespace OOMPolicies
} // namespace mallocMC
