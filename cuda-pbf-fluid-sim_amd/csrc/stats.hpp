// stats.hpp — density-constraint statistics over a cell-sorted state.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pbf_params.h"
#include "sorted_path.hpp"

namespace pbfx {

// out5 = {mean(rho/rho0 - 1), max(rho/rho0 - 1), mean in-range neighbours (capped
// like the lambda pass), mean |v|, mean(max(rho/rho0 - 1, 0))}.  Synchronises `s`.
bool density_stats_sorted(const float4* Ps, const float4* V, const uint32_t* idxS, const uint32_t* segOf,
                          const SpanTable& st, const sim::DeviceParams& dp, uint32_t N, double* out5,
                          hipStream_t s);

}  // namespace pbfx
