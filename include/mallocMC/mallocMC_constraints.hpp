/*
 * Policy constraints (src/include/mallocMC/mallocMC_constraints.hpp:40-88):
 * Scatter and XMallocSIMD must agree on the pagesize.
 */
#pragma once

#include "creationPolicies/Scatter.hpp"
#include "distributionPolicies/XMallocSIMD.hpp"

namespace mallocMC
{
    template<typename Creation, typename Distribution>
    struct PolicyCheck2
    {
    };

    template<typename X, typename Y, typename Z>
    struct PolicyCheck2<CreationPolicies::Scatter<X, Y>, DistributionPolicies::XMallocSIMD<Z>>
    {
        static_assert(
            CreationPolicies::Scatter<X, Y>::pagesize == DistributionPolicies::XMallocSIMD<Z>::pagesize,
            "Pagesize must be the same when combining Scatter and XMallocSIMD");
    };

    template<typename Creation, typename Distribution, typename OOM, typename ReservePool, typename Alignment>
    struct PolicyConstraints : PolicyCheck2<Creation, Distribution>
    {
    };
} // namespace mallocMC
