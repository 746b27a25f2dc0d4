/*
 * Umbrella header of the MI355X-native ScatterAlloc heap
 * (the reference's src/include/mallocMC/mallocMC.hpp:39-55).  Device code:
 * handle.malloc(acc, n) / free(acc, p) / getAvailableSlots(acc, n) inside any
 * HIP kernel; host code: mallocMC::Allocator<...>.  Link with
 * libmallocmc_hip.so.
 */
#pragma once

#include "version.hpp"

#include "alignmentPolicies/Noop.hpp"
#include "alignmentPolicies/Shrink.hpp"
#include "allocator.hpp"
#include "creationPolicies/Scatter.hpp"
#include "device_allocator.hpp"
#include "distributionPolicies/Noop.hpp"
#include "distributionPolicies/XMallocSIMD.hpp"
#include "mallocMC_allocator_handle.hpp"
#include "mallocMC_constraints.hpp"
#include "mallocMC_traits.hpp"
#include "oOMPolicies/ReturnNull.hpp"
#include "reservePoolPolicies/HipMalloc.hpp"
