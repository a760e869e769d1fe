// Internal declarations shared by the host side of libhfk (hfk_capi.hip) and the kernel launchers
// (hfk_kernels.hip).  Not part of the C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

namespace hfk {

// 1D tables of one (Q, P) pair: Lagrange values/derivatives on Q equispaced nodes of [0,1] at the P Gauss
// points (TPTable.N1/dN1, src/tensorproduct.jl:13-47), with the points and weights.  N[a*Q + i].
struct Tab {
  double N[16];
  double dN[16];
  double xq[4];
  double wq[4];
};

// Cells per work-group batch: the point phase maps one thread to one (cell, point) pair of the batch.
constexpr int kThreads = 256;
constexpr int cells_per_batch(int P) { return kThreads / (P * P * P); }

// Geometry of the launch of one colour (or of all batches when the kernel has no scatter).
struct BatchRange {
  int first;  // first batch
  int count;  // number of batches
};

struct OpDev {
  int Q, P, C, nn, np, nd;
  int law;
  double lam, mu;
  Tab tab;
  const int32_t* ids;     // [batch][cell][c][t] tensor order, 0 for padding cells
  const double* coords;   // [batch][cell][v][i]
  const double* mu_pts;   // [batch][cell*np + q] or null
  const double* dvals;    // Dirichlet values
  const int32_t* loc_of;  // tensor index c*nn + t -> table-local dof (assembly)
  const int32_t* nzpos;   // [cell (batch order)][ii + nd*jj] table-local, 1-based, or null
};

// Launchers (hfk_kernels.hip).  All are asynchronous on `s`.
hipError_t launch_prepare(const OpDev& op, int storage, void* tangents, const double* x, BatchRange r, hipStream_t s);
// Single-launch mode of the product: colfc[2c] = first batch and colfc[2c+1] = batch count of colour c (device),
// gx = neighbourhoods per XCD, grid = 8 * gx * ncol work-groups.
struct ApplyMap {
  const int* colfc;
  int ncol;
  int gx;
  int grid;
};

// y[id] += alpha * (J dx)[id] over the batches r of one colour (m == null), or over every batch in one launch
// with atomic adds (m != null); the caller zeroes y first for a plain product
hipError_t launch_apply(const OpDev& op, int storage, const void* tangents, const double* dx, double* y, BatchRange r,
                        hipStream_t s, double alpha = 1.0, const ApplyMap* m = nullptr);
// nodal 3x3 blocks of the tangent, accumulated into B[blk*9 + ci + 3cj] over the batches r of one colour
hipError_t launch_blocks(const OpDev& op, int storage, const void* tangents, const int32_t* block_of, double* B,
                         BatchRange r, hipStream_t s);
hipError_t launch_residual(const OpDev& op, const double* x, double* b, BatchRange r, hipStream_t s);
// pointwise laws compiled at run time: the point state (F, iJt, dV, mu: kStateEntries entries per point,
// entry-major per batch) and the residual from point coefficients (9 per point, entry-major per batch)
constexpr int kStateEntries = 20;
hipError_t launch_point_state(const OpDev& op, const double* x, double* st, BatchRange r, hipStream_t s);
hipError_t launch_residual_coefs(const OpDev& op, const double* coef, double* b, BatchRange r, hipStream_t s);
hipError_t launch_assemble(const OpDev& op, const double* tangents64, double* nzval, BatchRange r, hipStream_t s);
hipError_t launch_fill(double* y, int64_t n, double v, hipStream_t s);
hipError_t launch_expand_tangents(const OpDev& op, int storage, const void* tangents, double* D, int64_t nbatches,
                                  hipStream_t s);

// number of stored entries and bytes per entry of a storage mode
inline int store_entries(int storage) { return (storage == 1 || storage == 3) ? 45 : 81; }
inline int store_bytes(int storage) { return storage >= 2 ? 4 : 8; }

}  // namespace hfk
