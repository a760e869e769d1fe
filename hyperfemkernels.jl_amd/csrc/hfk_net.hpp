// Internal interface of the learned coarse correction (hfk_net.hip) for the p-multigrid host code.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/hfk.h"

namespace hfk {
// z = (s/h) N(r/s) on device vectors of the coarse level (s: RMS of r, h: cell size)
int net_apply_dev(hfk_net* nt, const double* r, double* z, hipStream_t s);
int64_t net_ncoarse(const hfk_net* nt);
}  // namespace hfk
