// Internal interface of the learned coarse correction (hfk_net.hip) for the p-multigrid host code.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/hfk.h"

namespace hfk {
// z = (s/h) N(r/s) on device vectors of the coarse level (s: RMS of r, h: cell size)
int net_apply_dev(hfk_net* nt, const double* r, double* z, hipStream_t s);
int64_t net_ncoarse(const hfk_net* nt);
double net_h(const hfk_net* nt);
// w = the network's output for r, in its own units (z = (s/h) w); the input scale s is kept for the backward
int net_forward_vec(hfk_net* nt, const double* r, double* w, hipStream_t s);
// accumulate the weight gradient of a loss whose derivative with respect to w is sc * gw
int net_backward_vec(hfk_net* nt, const double* gw, double sc, hipStream_t s);
// one Adam step on the gradient accumulated over `batch` samples (cleared afterwards)
int net_adam(hfk_net* nt, double lr, int batch, hipStream_t s);
// clear the accumulated weight gradient; copy it to host memory (FP32, num_params values)
int net_zero_grad(hfk_net* nt, hipStream_t s);
int net_copy_grad(hfk_net* nt, float* out, hipStream_t s);
}  // namespace hfk
