// g4s_build.h -- host connectivity builder interfaces.
#pragma once
#include "g4s_math.h"

namespace g4s {

int host_threads(int requested);

// Outgoing table, row-major [n][nsyn] of {target, weight bits}, rows sorted.
void build_table(uint64_t seed, int64_t n, int64_t n_exc, int32_t nsyn, uint2 *tab, int threads);

}  // namespace g4s
