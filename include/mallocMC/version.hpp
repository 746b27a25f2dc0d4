/*
 * Version of this implementation and of the mallocMC API it provides
 * (the reference's src/include/mallocMC/version.hpp:40-47 names 2.6.0).
 */
#pragma once

#define MALLOCMC_HIP_VERSION_MAJOR 0
#define MALLOCMC_HIP_VERSION_MINOR 1
#define MALLOCMC_HIP_VERSION_PATCH 0

// mallocMC API level this header set is a drop-in for
#define MALLOCMC_VERSION_MAJOR 2
#define MALLOCMC_VERSION_MINOR 6
#define MALLOCMC_VERSION_PATCH 0
