// Plain data shared by the host side of libhfk and its device code, including the kernels compiled at run time
// with hiprtc (this header and hfk_point.hpp are embedded in the library and compiled again there).
#pragma once
#include <stdint.h>

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
constexpr int cells_per_batch(int P) { return P == 3 ? 8 : kThreads / (P * P * P); }

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
  double* cellgeo;        // [slot][8] per-cell geometry of the affine factorised storage (iJt'iJt packed, det J)
};

}  // namespace hfk
