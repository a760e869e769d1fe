// Device kernels of libhfk for gfx950 (CDNA4, wave64).
//
// Data layout (HBM).  Cells are processed in batches of C = 256 / P^3 cells (9 at Q2/P=3, 4 at Q3/P=4); the
// batches of one colour are contiguous, so one launch per colour is race-free (src/coloring.jl:1-3).  Per batch:
//   ids      [cell][c][t]          int32, tensor node order t = i + Q j + Q^2 k, 0 on padding cells
//   coords   [cell][v][i]          the 8 vertices of the trilinear map
//   tangents [entry][cell*P^3 + q] entry-major, so the 256 threads of the point phase read each entry of their
//                                  point with one fully coalesced access per wave (64 contiguous doubles)
// The stored tangent is D = dV iJt' T iJt (reference coordinates): the point phase needs neither the geometry
// nor dV, and the apply reads nothing per point but D (src/matrixfree.jl:366-375 multiplies T, iJt and dV at
// every point).
//
// Work-group structure (256 threads, one batch): gather -> 3 interpolation sweeps (LDS) -> point phase (one
// thread per (cell, point)) -> 3 contraction sweeps (LDS) -> scatter.  The sweeps are the 3D sum factorisation
// of src/tensorproduct.jl:125-159 (tp_interpolate!) and :189-224 (tp_contract!), one thread per line of the
// Q x Q x Q (or P x P x P) cube of one (cell, component).
#include <hip/hip_runtime.h>

#include <type_traits>

#include <cmath>

#include "hfk_internal.hpp"

namespace hfk {
namespace {

template <int Q, int P>
struct Dims {
  static constexpr int NN = Q * Q * Q;
  static constexpr int NP = P * P * P;
  static constexpr int C = kThreads / NP;
  static constexpr int NSLOT = C * NP;
  static constexpr int SA = (3 * NN > 9 * P * P * Q) ? 3 * NN : 9 * P * P * Q;
  static constexpr int SB = (6 * P * Q * Q > 9 * NP) ? 6 * P * Q * Q : 9 * NP;
  static constexpr int LDS = C * (SA + SB);
};

// ---------------------------------------------------------------------------------------------------------
// 3x3 helpers, column-major a[i + 3 j]
__device__ __forceinline__ double det3(const double* a) {
  return a[0] * (a[4] * a[8] - a[7] * a[5]) - a[3] * (a[1] * a[8] - a[7] * a[2]) + a[6] * (a[1] * a[5] - a[4] * a[2]);
}
__device__ __forceinline__ void inv3(const double* a, double* r, double d) {
  const double id = 1.0 / d;
  r[0] = (a[4] * a[8] - a[7] * a[5]) * id;
  r[3] = -(a[3] * a[8] - a[6] * a[5]) * id;
  r[6] = (a[3] * a[7] - a[6] * a[4]) * id;
  r[1] = -(a[1] * a[8] - a[7] * a[2]) * id;
  r[4] = (a[0] * a[8] - a[6] * a[2]) * id;
  r[7] = -(a[0] * a[7] - a[6] * a[1]) * id;
  r[2] = (a[1] * a[5] - a[4] * a[2]) * id;
  r[5] = -(a[0] * a[5] - a[3] * a[2]) * id;
  r[8] = (a[0] * a[4] - a[3] * a[1]) * id;
}

// ---------------------------------------------------------------------------------------------------------
// Pointwise laws.  F[c + 3J] = delta_cJ + du_c/dx_J (Gridap: F = I + ∇u', src/kernels.jl:203-205).
// Per point the law is set up once (`Law` holds what its stress and tangent need), then evaluated.
//   stress:  P_cJ                                 (residual coefficients, src/kernels.jl:319-356)
//   tangent: T[(J,c),(L,k)] = dP_cJ / dF_kL       (point_tangent, src/factorforms.jl:317-322)
struct LawSVK {  // St Venant-Kirchhoff (test/common.jl:2-8): P = F S, S = lam tr(E) I + 2 mu E
  double F[9], S[9], b[9], lam, mu;
  __device__ void setup(const double* Fin, double l, double m) {
    lam = l;
    mu = m;
#pragma unroll
    for (int i = 0; i < 9; ++i) F[i] = Fin[i];
    double E[9];
#pragma unroll
    for (int I = 0; I < 3; ++I)
#pragma unroll
      for (int J = 0; J < 3; ++J) {
        double s = 0;
#pragma unroll
        for (int k = 0; k < 3; ++k) s += F[k + 3 * I] * F[k + 3 * J];
        E[I + 3 * J] = 0.5 * (s - (I == J ? 1.0 : 0.0));
      }
    const double tr = E[0] + E[4] + E[8];
#pragma unroll
    for (int I = 0; I < 3; ++I)
#pragma unroll
      for (int J = 0; J < 3; ++J) S[I + 3 * J] = (I == J ? lam * tr : 0.0) + 2 * mu * E[I + 3 * J];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        double s = 0;
#pragma unroll
        for (int I = 0; I < 3; ++I) s += F[c + 3 * I] * F[k + 3 * I];
        b[c + 3 * k] = s;
      }
  }
  __device__ double stress(int c, int J) const {
    double s = 0;
#pragma unroll
    for (int I = 0; I < 3; ++I) s += F[c + 3 * I] * S[I + 3 * J];
    return s;
  }
  // dP_cJ/dF_kL = delta_ck S_LJ + lam F_cJ F_kL + mu (F_cL F_kJ + (F F')_ck delta_JL)
  __device__ double tangent(int c, int J, int k, int L) const {
    return (c == k ? S[L + 3 * J] : 0.0) + lam * F[c + 3 * J] * F[k + 3 * L] +
           mu * (F[c + 3 * L] * F[k + 3 * J] + (J == L ? b[c + 3 * k] : 0.0));
  }
};

struct LawNeoHooke {  // W = mu/2 (I1 - 3) - mu lnJ + lam/2 lnJ^2:  P = mu (F - F^-T) + lam lnJ F^-T
  double F[9], A[9], lam, mu, lnJ;  // A = F^-1, A[J + 3c] = (F^-1)_Jc
  __device__ void setup(const double* Fin, double l, double m) {
    lam = l;
    mu = m;
#pragma unroll
    for (int i = 0; i < 9; ++i) F[i] = Fin[i];
    const double d = det3(F);
    inv3(F, A, d);
    lnJ = log(d);
  }
  __device__ double stress(int c, int J) const { return mu * F[c + 3 * J] + (lam * lnJ - mu) * A[J + 3 * c]; }
  // mu d_ck d_JL + (mu - lam lnJ) A_Jk A_Lc + lam A_Jc A_Lk
  __device__ double tangent(int c, int J, int k, int L) const {
    return (c == k && J == L ? mu : 0.0) + (mu - lam * lnJ) * A[J + 3 * k] * A[L + 3 * c] + lam * A[J + 3 * c] * A[L + 3 * k];
  }
};

struct LawLinear {  // sigma = lam tr(G) I + mu (G + G'), G = F - I
  double G[9], lam, mu, tr;
  __device__ void setup(const double* Fin, double l, double m) {
    lam = l;
    mu = m;
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int J = 0; J < 3; ++J) G[c + 3 * J] = Fin[c + 3 * J] - (c == J ? 1.0 : 0.0);
    tr = G[0] + G[4] + G[8];
  }
  __device__ double stress(int c, int J) const {
    return (c == J ? lam * tr : 0.0) + mu * (G[c + 3 * J] + G[J + 3 * c]);
  }
  __device__ double tangent(int c, int J, int k, int L) const {
    return (c == J && k == L ? lam : 0.0) + (c == k && J == L ? mu : 0.0) + (J == k && c == L ? mu : 0.0);
  }
};

// ---------------------------------------------------------------------------------------------------------
// Sum-factorised interpolation of the 3 components of the C cells of a batch.
// In : A[cell*SA + c*NN + t]              nodal values, tensor order
// Out: B[cell*SB + q*9 + K + 3c]          d u_c / d xi_K at tensor point q = a + P b + P^2 cc
template <int Q, int P, class R = double>
__device__ __forceinline__ void interp_sweeps(R* __restrict__ A, R* __restrict__ B, const Tab& tb, int tid) {
  using Dm = Dims<Q, P>;
  constexpr int C = Dm::C, SA = Dm::SA, SB = Dm::SB;
  // sweep 1 (x): items (cell, c, jk); value and x-derivative at (a, j, k)
  for (int it = tid; it < C * 3 * Q * Q; it += kThreads) {
    const int cell = it / (3 * Q * Q), r = it - cell * 3 * Q * Q, c = r / (Q * Q), jk = r - c * Q * Q;
    const R* u = A + cell * SA + c * Dm::NN + jk * Q;
    R ui[Q];
#pragma unroll
    for (int i = 0; i < Q; ++i) ui[i] = u[i];
    R* o = B + cell * SB + c * 2 * P * Q * Q;
#pragma unroll
    for (int a = 0; a < P; ++a) {
      R s = 0, ds = 0;
#pragma unroll
      for (int i = 0; i < Q; ++i) {
        s += (R)tb.N[a * Q + i] * ui[i];
        ds += (R)tb.dN[a * Q + i] * ui[i];
      }
      o[a + P * jk] = s;
      o[P * Q * Q + a + P * jk] = ds;
    }
  }
  __syncthreads();
  // sweep 2 (y): items (cell, c, a, k); value, x- and y-derivatives at (a, b, k)
  for (int it = tid; it < C * 3 * P * Q; it += kThreads) {
    const int cell = it / (3 * P * Q), r = it - cell * 3 * P * Q, c = r / (P * Q), ak = r - c * P * Q;
    const int a = ak % P, k = ak / P;
    const R* in = B + cell * SB + c * 2 * P * Q * Q;
    R av[Q], bv[Q];
#pragma unroll
    for (int j = 0; j < Q; ++j) {
      av[j] = in[a + P * (j + Q * k)];
      bv[j] = in[P * Q * Q + a + P * (j + Q * k)];
    }
    R* o = A + cell * SA + c * 3 * P * P * Q;
#pragma unroll
    for (int b = 0; b < P; ++b) {
      R v = 0, g1 = 0, g2 = 0;
#pragma unroll
      for (int j = 0; j < Q; ++j) {
        v += (R)tb.N[b * Q + j] * av[j];
        g1 += (R)tb.N[b * Q + j] * bv[j];
        g2 += (R)tb.dN[b * Q + j] * av[j];
      }
      const int idx = a + P * b + P * P * k;
      o[idx] = v;
      o[P * P * Q + idx] = g1;
      o[2 * P * P * Q + idx] = g2;
    }
  }
  __syncthreads();
  // sweep 3 (z): items (cell, c, ab); the 3 reference derivatives at (a, b, cc)
  for (int it = tid; it < C * 3 * P * P; it += kThreads) {
    const int cell = it / (3 * P * P), r = it - cell * 3 * P * P, c = r / (P * P), ab = r - c * P * P;
    const R* in = A + cell * SA + c * 3 * P * P * Q;
    R vv[Q], g1v[Q], g2v[Q];
#pragma unroll
    for (int k = 0; k < Q; ++k) {
      vv[k] = in[ab + P * P * k];
      g1v[k] = in[P * P * Q + ab + P * P * k];
      g2v[k] = in[2 * P * P * Q + ab + P * P * k];
    }
    R* o = B + cell * SB;
#pragma unroll
    for (int cc = 0; cc < P; ++cc) {
      R x = 0, y = 0, z = 0;
#pragma unroll
      for (int k = 0; k < Q; ++k) {
        x += (R)tb.N[cc * Q + k] * g1v[k];
        y += (R)tb.N[cc * Q + k] * g2v[k];
        z += (R)tb.dN[cc * Q + k] * vv[k];
      }
      const int q = ab + P * P * cc;
      o[q * 9 + 3 * c + 0] = x;
      o[q * 9 + 3 * c + 1] = y;
      o[q * 9 + 3 * c + 2] = z;
    }
  }
  __syncthreads();
}

// Sum-factorised contraction (transpose of the interpolation, gradient coefficients only).
// In : B[cell*SB + q*9 + K + 3c]   coefficient of d/dxi_K of the test function of component c at point q
// Out: emit(cell, c, t, r)          r = sum_q sum_K dN_t/dxi_K(q) coef   for every tensor node t
template <int Q, int P, class R = double, class Emit>
__device__ __forceinline__ void contract_sweeps(R* __restrict__ A, R* __restrict__ B, const Tab& tb, int tid,
                                                Emit&& emit) {
  using Dm = Dims<Q, P>;
  constexpr int C = Dm::C, SA = Dm::SA, SB = Dm::SB;
  __syncthreads();
  // sweep 3T (z): items (cell, c, ab)
  for (int it = tid; it < C * 3 * P * P; it += kThreads) {
    const int cell = it / (3 * P * P), r = it - cell * 3 * P * P, c = r / (P * P), ab = r - c * P * P;
    const R* in = B + cell * SB;
    R cx[P], cy[P], cz[P];
#pragma unroll
    for (int cc = 0; cc < P; ++cc) {
      const int q = ab + P * P * cc;
      cx[cc] = in[q * 9 + 3 * c + 0];
      cy[cc] = in[q * 9 + 3 * c + 1];
      cz[cc] = in[q * 9 + 3 * c + 2];
    }
    R* o = A + cell * SA + c * 3 * P * P * Q;
#pragma unroll
    for (int k = 0; k < Q; ++k) {
      R t0 = 0, t1 = 0, t2 = 0;
#pragma unroll
      for (int cc = 0; cc < P; ++cc) {
        t0 += (R)tb.dN[cc * Q + k] * cz[cc];
        t1 += (R)tb.N[cc * Q + k] * cx[cc];
        t2 += (R)tb.N[cc * Q + k] * cy[cc];
      }
      o[ab + P * P * k] = t0;
      o[P * P * Q + ab + P * P * k] = t1;
      o[2 * P * P * Q + ab + P * P * k] = t2;
    }
  }
  __syncthreads();
  // sweep 2T (y): items (cell, c, a, k)
  for (int it = tid; it < C * 3 * P * Q; it += kThreads) {
    const int cell = it / (3 * P * Q), r = it - cell * 3 * P * Q, c = r / (P * Q), ak = r - c * P * Q;
    const int a = ak % P, k = ak / P;
    const R* in = A + cell * SA + c * 3 * P * P * Q;
    R t0[P], t1[P], t2[P];
#pragma unroll
    for (int b = 0; b < P; ++b) {
      t0[b] = in[a + P * b + P * P * k];
      t1[b] = in[P * P * Q + a + P * b + P * P * k];
      t2[b] = in[2 * P * P * Q + a + P * b + P * P * k];
    }
    R* o = B + cell * SB + c * 2 * P * Q * Q;
#pragma unroll
    for (int j = 0; j < Q; ++j) {
      R s0 = 0, s1 = 0;
#pragma unroll
      for (int b = 0; b < P; ++b) {
        s0 += (R)tb.N[b * Q + j] * t0[b] + (R)tb.dN[b * Q + j] * t2[b];
        s1 += (R)tb.N[b * Q + j] * t1[b];
      }
      o[a + P * (j + Q * k)] = s0;
      o[P * Q * Q + a + P * (j + Q * k)] = s1;
    }
  }
  __syncthreads();
  // sweep 1T (x): items (cell, c, jk)
  for (int it = tid; it < C * 3 * Q * Q; it += kThreads) {
    const int cell = it / (3 * Q * Q), r = it - cell * 3 * Q * Q, c = r / (Q * Q), jk = r - c * Q * Q;
    const R* in = B + cell * SB + c * 2 * P * Q * Q;
    R s0[P], s1[P];
#pragma unroll
    for (int a = 0; a < P; ++a) {
      s0[a] = in[a + P * jk];
      s1[a] = in[P * Q * Q + a + P * jk];
    }
    R rl[Q];
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      R s = 0;
#pragma unroll
      for (int a = 0; a < P; ++a) s += (R)tb.N[a * Q + i] * s0[a] + (R)tb.dN[a * Q + i] * s1[a];
      rl[i] = s;
    }
    emit(cell, c, Q * jk, rl);  // the Q nodes t = Q jk + i of one x-line
  }
}

// y[id] += alpha * r for the Q nodes of one x-line of a cell: every load first, then every store (the ids of a
// line are distinct, so the read-modify-writes are independent and their latencies overlap).
template <int Q, class R = double>
__device__ __forceinline__ void scatter_line(const int32_t* __restrict__ id_line, const R* r, double alpha,
                                             double* __restrict__ y) {
  int id[Q];
  double yv[Q];
#pragma unroll
  for (int i = 0; i < Q; ++i) id[i] = id_line[i];
#pragma unroll
  for (int i = 0; i < Q; ++i) yv[i] = id[i] > 0 ? y[id[i] - 1] : 0.0;
#pragma unroll
  for (int i = 0; i < Q; ++i)
    if (id[i] > 0) y[id[i] - 1] = yv[i] + alpha * static_cast<double>(r[i]);
}

// Geometry of the trilinear map at tensor point q (point_jacobian/point_geometry, src/assembly.jl:1073-1111):
// Jt[K + 3i] = sum_v dM_v/dxi_K x_v[i]; iJt = Jt^-1; dV = w_q |det Jt|.
template <int P>
__device__ __forceinline__ void point_geometry(const Tab& tb, const double* __restrict__ X, int q, double* iJt,
                                               double& dV) {
  const int a = q % P, b = (q / P) % P, cc = q / (P * P);
  const double xi[3] = {tb.xq[a], tb.xq[b], tb.xq[cc]};
  double Jt[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) Jt[i] = 0;
#pragma unroll
  for (int v = 0; v < 8; ++v) {
    const int bx = v & 1, by = (v >> 1) & 1, bz = (v >> 2) & 1;
    const double m0 = bx ? xi[0] : 1 - xi[0], m1 = by ? xi[1] : 1 - xi[1], m2 = bz ? xi[2] : 1 - xi[2];
    const double d0 = bx ? 1.0 : -1.0, d1 = by ? 1.0 : -1.0, d2 = bz ? 1.0 : -1.0;
    const double dM[3] = {d0 * m1 * m2, m0 * d1 * m2, m0 * m1 * d2};
#pragma unroll
    for (int K = 0; K < 3; ++K)
#pragma unroll
      for (int i = 0; i < 3; ++i) Jt[K + 3 * i] += dM[K] * X[v * 3 + i];
  }
  const double d = det3(Jt);
  inv3(Jt, iJt, d);
  dV = tb.wq[a] * tb.wq[b] * tb.wq[cc] * fabs(d);
}

// Index of the packed upper triangle (i <= j) of a symmetric 9x9 matrix.
__host__ __device__ constexpr int sym_index(int i, int j) { return j * (j + 1) / 2 + i; }

// the tangents are written once per prepare! and read by the next products, never from cache.  Non-temporal
// stores measured faster for the packed symmetric storages (-12 % sym64, -20 % sym32 at 64^3 Q2) and slower for
// the full ones (+15 % full64), so they are used for the former only
template <class TS, bool NT>
__device__ __forceinline__ void store_entry(void* base, size_t idx, double v) {
  if (NT) __builtin_nontemporal_store(static_cast<TS>(v), static_cast<TS*>(base) + idx);
  else static_cast<TS*>(base)[idx] = static_cast<TS>(v);
}

// ---------------------------------------------------------------------------------------------------------
// prepare!: per point, the reference-coordinate tangent D = dV iJt' T iJt (src/matrixfree.jl:199-228).
template <int Q, int P, class Law, class TS, bool SYM>
__global__ __launch_bounds__(kThreads) void k_prepare(OpDev op, void* __restrict__ tangents,
                                                      const double* __restrict__ x, int batch0) {
  using Dm = Dims<Q, P>;
  constexpr int C = Dm::C, SA = Dm::SA, SB = Dm::SB, NSLOT = Dm::NSLOT, NE = SYM ? 45 : 81;
  __shared__ double lds[Dm::LDS];
  double* A = lds;
  double* B = lds + C * SA;
  const int tid = threadIdx.x;
  const int batch = batch0 + blockIdx.x;
  const int32_t* ids = op.ids + (size_t)batch * C * 3 * Dm::NN;
  for (int it = tid; it < C * 3 * Dm::NN; it += kThreads) {
    const int id = ids[it];
    const int cell = it / (3 * Dm::NN);
    A[cell * SA + (it - cell * 3 * Dm::NN)] = id > 0 ? x[id - 1] : (id < 0 ? op.dvals[-id - 1] : 0.0);
  }
  __syncthreads();
  interp_sweeps<Q, P>(A, B, op.tab, tid);
  if (tid < NSLOT) {
    const int cell = tid / Dm::NP, q = tid - cell * Dm::NP;
    const double* X = op.coords + ((size_t)batch * C + cell) * 24;
    double iJt[9], dV;
    point_geometry<P>(op.tab, X, q, iJt, dV);
    const double* su = B + cell * SB + q * 9;
    double F[9];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int J = 0; J < 3; ++J) {
        double g = 0;
#pragma unroll
        for (int K = 0; K < 3; ++K) g += iJt[J + 3 * K] * su[3 * c + K];
        F[c + 3 * J] = (c == J ? 1.0 : 0.0) + g;
      }
    const double mu = op.mu_pts ? op.mu_pts[(size_t)batch * NSLOT + tid] : op.mu;
    Law law;
    law.setup(F, op.lam, mu);
    const size_t base = (size_t)batch * NE * NSLOT + tid;
    // D[(K,c),(M,k)] = dV sum_J sum_L iJt[J,K] T[(J,c),(L,k)] iJt[L,M], one (c, k) block at a time
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        if (SYM && c > k) continue;
        double Tb[9];  // Tb[J + 3L]
#pragma unroll
        for (int L = 0; L < 3; ++L)
#pragma unroll
          for (int J = 0; J < 3; ++J) Tb[J + 3 * L] = law.tangent(c, J, k, L);
        double W[9];  // W[K + 3L] = sum_J iJt[J,K] Tb[J,L]
#pragma unroll
        for (int L = 0; L < 3; ++L)
#pragma unroll
          for (int K = 0; K < 3; ++K) {
            double s = 0;
#pragma unroll
            for (int J = 0; J < 3; ++J) s += iJt[J + 3 * K] * Tb[J + 3 * L];
            W[K + 3 * L] = s;
          }
#pragma unroll
        for (int M = 0; M < 3; ++M)
#pragma unroll
          for (int K = 0; K < 3; ++K) {
            double s = 0;
#pragma unroll
            for (int L = 0; L < 3; ++L) s += W[K + 3 * L] * iJt[L + 3 * M];
            const int i = K + 3 * c, j = M + 3 * k;
            if (!SYM) {
              store_entry<TS, false>(tangents, base + (size_t)(i + 9 * j) * NSLOT, dV * s);
            } else if (c < k || K <= M) {
              // c < k: the whole block lies above the diagonal; c == k: its upper triangle.  The tangents of
              // the built-in laws are symmetric, so the upper triangle determines D.
              store_entry<TS, true>(tangents, base + (size_t)sym_index(i, j) * NSLOT, dV * s);
            }
          }
      }
  }
}

// ---------------------------------------------------------------------------------------------------------
// Batch of a work-group in the single-launch (atomic) mode.  The hardware hands work-group b to XCD b % 8;
// work-groups j = b / 8 of one XCD take the colours of one neighbourhood in turn: colour c = j % ncol of
// neighbourhood g = xcd * gx + j / ncol (the g-th batch of every colour covers neighbouring cells, which share
// nodes), so the dx entries they gather and the y entries they add to stay in that XCD's L2.  -1: idle.
__device__ __forceinline__ int interleaved_batch(const int* __restrict__ colfc, int ncol, int gx) {
  const int b = blockIdx.x, xcd = b & 7, j = b >> 3;
  const int c = j % ncol, g = xcd * gx + j / ncol;
  return g < colfc[2 * c + 1] ? colfc[2 * c] + g : -1;
}

// mul!(y, mf, dx) (src/matrixfree.jl:341-396): gather dx (0 at Dirichlet dofs), interpolate, sv = D su,
// contract, add to the free rows of y.  Deterministic mode (ATOMIC = false): the batches of one colour per
// launch, plain read-modify-writes, the reference's colour order of the sums.  ATOMIC: every batch in one
// launch, FP64 atomic adds (the order of the additions into one y entry varies from run to run).
// the stored tangents are read once per product and never reused within it: non-temporal loads keep them from
// evicting the gathered dx and the scattered y from the caches
template <class T>
__device__ __forceinline__ T ld_stream(const T* p) {
  return __builtin_nontemporal_load(p);
}

template <int Q, int P, class TS, bool SYM, bool ATOMIC>
__global__ __launch_bounds__(kThreads) void k_apply(OpDev op, const void* __restrict__ tangents,
                                                    const double* __restrict__ dx, double* __restrict__ y,
                                                    int batch0, double alpha, const int* __restrict__ colfc,
                                                    int ncol, int gx) {
  using Dm = Dims<Q, P>;
  // FP32 tangents (the smoothers' storages): the sweeps and the point product run in FP32 too, which halves the
  // LDS and register footprint of the work-group; y accumulates in FP64
  using CT = typename std::conditional<std::is_same<TS, float>::value, float, double>::type;
  constexpr int C = Dm::C, SA = Dm::SA, SB = Dm::SB, NSLOT = Dm::NSLOT, NE = SYM ? 45 : 81;
  __shared__ CT lds[Dm::LDS];
  CT* A = lds;
  CT* B = lds + C * SA;
  const int tid = threadIdx.x;
  const int batch = ATOMIC ? interleaved_batch(colfc, ncol, gx) : batch0 + blockIdx.x;
  if (ATOMIC && batch < 0) return;  // uniform over the work-group
  const int32_t* ids = op.ids + (size_t)batch * C * 3 * Dm::NN;
  for (int it = tid; it < C * 3 * Dm::NN; it += kThreads) {
    const int id = ids[it];
    const int cell = it / (3 * Dm::NN);
    A[cell * SA + (it - cell * 3 * Dm::NN)] = id > 0 ? static_cast<CT>(dx[id - 1]) : CT(0);
  }
  __syncthreads();
  interp_sweeps<Q, P, CT>(A, B, op.tab, tid);
  if (tid < NSLOT) {
    const int cell = tid / Dm::NP, q = tid - cell * Dm::NP;
    CT* s = B + cell * SB + q * 9;
    CT su[9], sv[9];
#pragma unroll
    for (int j = 0; j < 9; ++j) {
      su[j] = s[j];
      sv[j] = 0.0;
    }
    const TS* Dp = static_cast<const TS*>(tangents) + (size_t)batch * NE * NSLOT + tid;
    if (!SYM) {
#pragma unroll
      for (int j = 0; j < 9; ++j)
#pragma unroll
        for (int i = 0; i < 9; ++i) sv[i] += static_cast<CT>(ld_stream(Dp + (size_t)(i + 9 * j) * NSLOT)) * su[j];
    } else {
#pragma unroll
      for (int j = 0; j < 9; ++j)
#pragma unroll
        for (int i = 0; i <= j; ++i) {
          const CT d = static_cast<CT>(ld_stream(Dp + (size_t)sym_index(i, j) * NSLOT));
          sv[i] += d * su[j];
          if (i != j) sv[j] += d * su[i];
        }
    }
#pragma unroll
    for (int i = 0; i < 9; ++i) s[i] = sv[i];
  }
  contract_sweeps<Q, P, CT>(A, B, op.tab, tid, [&](int cell, int c, int t0, const CT* r) {
    const int32_t* idl = ids + cell * 3 * Dm::NN + c * Dm::NN + t0;
    if (ATOMIC) {
#pragma unroll
      for (int i = 0; i < Q; ++i) {
        const int id = idl[i];
        if (id > 0) unsafeAtomicAdd(y + id - 1, alpha * static_cast<double>(r[i]));  // hardware FP64 add in L2
      }
    } else {
      scatter_line<Q>(idl, r, alpha, y);  // race-free: the cells of one colour share no free dof
    }
  });
}

// ---------------------------------------------------------------------------------------------------------
// BlockJacobi from the stored tangents (_blocks_from_tangents!, src/preconditioners.jl:234-263), the cells of
// one colour: B_n[ci,cj] += sum_q sum_K sum_M dN_n/dxi_K D_q[(K,ci),(M,cj)] dN_n/dxi_M.  One thread per
// (cell, node); the tangents of the current point of the batch's cells are staged in LDS.
template <int Q, int P, class TS, bool SYM>
__global__ __launch_bounds__(kThreads) void k_blocks(OpDev op, const void* __restrict__ tangents,
                                                     const int32_t* __restrict__ block_of, double* __restrict__ Bk,
                                                     int batch0) {
  using Dm = Dims<Q, P>;
  constexpr int C = Dm::C, NN = Dm::NN, NP = Dm::NP, NSLOT = Dm::NSLOT, NE = SYM ? 45 : 81;
  __shared__ double Dl[C * 81];
  const int tid = threadIdx.x;
  const int batch = batch0 + blockIdx.x;
  const bool active = tid < C * NN;
  const int cell = tid / NN, t = tid - cell * NN;
  const int i = t % Q, j = (t / Q) % Q, k = t / (Q * Q);
  const TS* Dg = static_cast<const TS*>(tangents) + (size_t)batch * NE * NSLOT;
  double acc[9];
#pragma unroll
  for (int e = 0; e < 9; ++e) acc[e] = 0.0;
  for (int q = 0; q < NP; ++q) {
    __syncthreads();
    for (int it = tid; it < C * 81; it += kThreads) {
      const int c2 = it / 81, e = it - c2 * 81, ii = e % 9, jj = e / 9;
      const int se = SYM ? (ii <= jj ? sym_index(ii, jj) : sym_index(jj, ii)) : e;
      Dl[it] = static_cast<double>(ld_stream(Dg + (size_t)se * NSLOT + c2 * NP + q));
    }
    __syncthreads();
    if (active) {
      const int a = q % P, b = (q / P) % P, cc = q / (P * P);
      const double na = op.tab.N[a * Q + i], nb = op.tab.N[b * Q + j], nc = op.tab.N[cc * Q + k];
      const double G[3] = {op.tab.dN[a * Q + i] * nb * nc, na * op.tab.dN[b * Q + j] * nc, na * nb * op.tab.dN[cc * Q + k]};
      const double* D = Dl + cell * 81;
#pragma unroll
      for (int cj = 0; cj < 3; ++cj)
#pragma unroll
        for (int ci = 0; ci < 3; ++ci) {
          double s = 0;
#pragma unroll
          for (int M = 0; M < 3; ++M) {
            double u = 0;
#pragma unroll
            for (int K = 0; K < 3; ++K) u += G[K] * D[(K + 3 * ci) + 9 * (M + 3 * cj)];
            s += u * G[M];
          }
          acc[ci + 3 * cj] += s;
        }
    }
  }
  if (active) {
    const int blk = block_of[((size_t)batch * C + cell) * NN + t];
    if (blk >= 0) {
#pragma unroll
      for (int e = 0; e < 9; ++e) Bk[(size_t)blk * 9 + e] += acc[e];  // race-free within a colour
    }
  }
}

// ---------------------------------------------------------------------------------------------------------
// residual!: b += sum over the batches of one colour of the element residuals  ∫ ∇(v)' ⊙ P(F) dΩ
// (element_residual_mm!, src/assembly.jl:1219-1241; residual_coefs!, src/kernels.jl:340-356).
template <int Q, int P, class Law>
__global__ __launch_bounds__(kThreads) void k_residual(OpDev op, const double* __restrict__ x, double* __restrict__ b,
                                                       int batch0) {
  using Dm = Dims<Q, P>;
  constexpr int C = Dm::C, SA = Dm::SA, SB = Dm::SB, NSLOT = Dm::NSLOT;
  __shared__ double lds[Dm::LDS];
  double* A = lds;
  double* B = lds + C * SA;
  const int tid = threadIdx.x;
  const int batch = batch0 + blockIdx.x;
  const int32_t* ids = op.ids + (size_t)batch * C * 3 * Dm::NN;
  for (int it = tid; it < C * 3 * Dm::NN; it += kThreads) {
    const int id = ids[it];
    const int cell = it / (3 * Dm::NN);
    A[cell * SA + (it - cell * 3 * Dm::NN)] = id > 0 ? x[id - 1] : (id < 0 ? op.dvals[-id - 1] : 0.0);
  }
  __syncthreads();
  interp_sweeps<Q, P>(A, B, op.tab, tid);
  if (tid < NSLOT) {
    const int cell = tid / Dm::NP, q = tid - cell * Dm::NP;
    const double* X = op.coords + ((size_t)batch * C + cell) * 24;
    double iJt[9], dV;
    point_geometry<P>(op.tab, X, q, iJt, dV);
    double* s = B + cell * SB + q * 9;
    double F[9];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int J = 0; J < 3; ++J) {
        double g = 0;
#pragma unroll
        for (int K = 0; K < 3; ++K) g += iJt[J + 3 * K] * s[3 * c + K];
        F[c + 3 * J] = (c == J ? 1.0 : 0.0) + g;
      }
    const double mu = op.mu_pts ? op.mu_pts[(size_t)batch * NSLOT + tid] : op.mu;
    Law law;
    law.setup(F, op.lam, mu);
    // coefficient of d/dxi_K of the test function of component c: dV sum_J iJt[J,K] P_cJ
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      double Pc[3];
#pragma unroll
      for (int J = 0; J < 3; ++J) Pc[J] = law.stress(c, J);
#pragma unroll
      for (int K = 0; K < 3; ++K) {
        double v = 0;
#pragma unroll
        for (int J = 0; J < 3; ++J) v += iJt[J + 3 * K] * Pc[J];
        s[3 * c + K] = dV * v;
      }
    }
  }
  contract_sweeps<Q, P>(A, B, op.tab, tid, [&](int cell, int c, int t0, const double* r) {
    scatter_line<Q>(ids + cell * 3 * Dm::NN + c * Dm::NN + t0, r, 1.0, b);
  });
}

// ---------------------------------------------------------------------------------------------------------
// Point state for a pointwise law compiled at run time: F (9), iJt (9), dV and mu of every point, entry-major
// st[(batch * kStateEntries + e) * NSLOT + slot] (the gather, interpolation and geometry of k_prepare).
template <int Q, int P>
__global__ __launch_bounds__(kThreads) void k_point_state(OpDev op, const double* __restrict__ x, double* __restrict__ st,
                                                          int batch0) {
  using Dm = Dims<Q, P>;
  constexpr int C = Dm::C, SA = Dm::SA, SB = Dm::SB, NSLOT = Dm::NSLOT;
  __shared__ double lds[Dm::LDS];
  double* A = lds;
  double* B = lds + C * SA;
  const int tid = threadIdx.x;
  const int batch = batch0 + blockIdx.x;
  const int32_t* ids = op.ids + (size_t)batch * C * 3 * Dm::NN;
  for (int it = tid; it < C * 3 * Dm::NN; it += kThreads) {
    const int id = ids[it];
    const int cell = it / (3 * Dm::NN);
    A[cell * SA + (it - cell * 3 * Dm::NN)] = id > 0 ? x[id - 1] : (id < 0 ? op.dvals[-id - 1] : 0.0);
  }
  __syncthreads();
  interp_sweeps<Q, P>(A, B, op.tab, tid);
  if (tid < NSLOT) {
    const int cell = tid / Dm::NP, q = tid - cell * Dm::NP;
    const double* X = op.coords + ((size_t)batch * C + cell) * 24;
    double iJt[9], dV;
    point_geometry<P>(op.tab, X, q, iJt, dV);
    const double* su = B + cell * SB + q * 9;
    double* o = st + (size_t)batch * kStateEntries * NSLOT + tid;
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int J = 0; J < 3; ++J) {
        double g = 0;
#pragma unroll
        for (int K = 0; K < 3; ++K) g += iJt[J + 3 * K] * su[3 * c + K];
        o[(size_t)(c + 3 * J) * NSLOT] = (c == J ? 1.0 : 0.0) + g;
      }
#pragma unroll
    for (int e = 0; e < 9; ++e) o[(size_t)(9 + e) * NSLOT] = iJt[e];
    o[(size_t)18 * NSLOT] = dV;
    o[(size_t)19 * NSLOT] = op.mu_pts ? op.mu_pts[(size_t)batch * NSLOT + tid] : op.mu;
  }
}

// residual! from point coefficients coef[(batch * 9 + K + 3c) * NSLOT + slot] (dV sum_J iJt[J,K] P_cJ, written
// by a law compiled at run time): the contraction and scatter of k_residual, colour by colour
template <int Q, int P>
__global__ __launch_bounds__(kThreads) void k_residual_coefs(OpDev op, const double* __restrict__ coef,
                                                             double* __restrict__ b, int batch0) {
  using Dm = Dims<Q, P>;
  constexpr int C = Dm::C, SA = Dm::SA, SB = Dm::SB, NSLOT = Dm::NSLOT;
  __shared__ double lds[Dm::LDS];
  double* A = lds;
  double* B = lds + C * SA;
  const int tid = threadIdx.x;
  const int batch = batch0 + blockIdx.x;
  const int32_t* ids = op.ids + (size_t)batch * C * 3 * Dm::NN;
  if (tid < NSLOT) {
    const int cell = tid / Dm::NP, q = tid - cell * Dm::NP;
    double* s = B + cell * SB + q * 9;
#pragma unroll
    for (int e = 0; e < 9; ++e) s[e] = coef[((size_t)batch * 9 + e) * NSLOT + tid];
  }
  contract_sweeps<Q, P>(A, B, op.tab, tid, [&](int cell, int c, int t0, const double* r) {
    scatter_line<Q>(ids + cell * 3 * Dm::NN + c * Dm::NN + t0, r, 1.0, b);
  });
}

// ---------------------------------------------------------------------------------------------------------
// jacobian!: element matrices from the prepared FP64 tangents, one work-group per cell of a colour, summed
// into the nonzeros through nzpos (element_jacobian_mm!/scatter_jacobian!, src/assembly.jl:1316-1408).
// Ke[(n,c),(m,k)] = sum_q sum_K sum_M dN_n/dxi_K(q) D_q[(K,c),(M,k)] dN_m/dxi_M(q), in two steps per point:
//   H[K][c][(m,k)] = sum_M D_q[(K,c),(M,k)] dN_m/dxi_M(q);   Ke[(n,c),(m,k)] += sum_K dN_n/dxi_K(q) H[K][c][(m,k)]
template <int Q, int P, int NT>
__global__ __launch_bounds__(NT) void k_assemble(OpDev op, const double* __restrict__ D, double* __restrict__ nzval,
                                                 int batch0) {
  using Dm = Dims<Q, P>;
  constexpr int C = Dm::C, NN = Dm::NN, ND = 3 * NN, NSLOT = Dm::NSLOT;
  constexpr int NE = (ND * ND + NT - 1) / NT;
  __shared__ double H[9 * ND];
  __shared__ double G[3 * NN];  // dN_t/dxi_K at the current point
  __shared__ double Dq[81];
  const int tid = threadIdx.x;
  const int batch = batch0 + blockIdx.x / C, cell = blockIdx.x % C;
  const int32_t* ids = op.ids + ((size_t)batch * C + cell) * ND;
  bool pad = true;
  for (int j = 0; j < ND; ++j) pad = pad && ids[j] == 0;
  if (pad) return;  // padding cell of the last batch of a colour (uniform over the work-group)
  double acc[NE];
#pragma unroll
  for (int e = 0; e < NE; ++e) acc[e] = 0.0;
  for (int q = 0; q < Dm::NP; ++q) {
    const int a = q % P, b = (q / P) % P, cc = q / (P * P);
    for (int it = tid; it < 3 * NN; it += NT) {
      const int t = it / 3, K = it - 3 * t;
      const int i = t % Q, j = (t / Q) % Q, k = t / (Q * Q);
      const double na = op.tab.N[a * Q + i], nb = op.tab.N[b * Q + j], nc = op.tab.N[cc * Q + k];
      G[it] = K == 0 ? op.tab.dN[a * Q + i] * nb * nc : K == 1 ? na * op.tab.dN[b * Q + j] * nc : na * nb * op.tab.dN[cc * Q + k];
    }
    for (int it = tid; it < 81; it += NT) Dq[it] = D[(size_t)batch * 81 * NSLOT + (size_t)it * NSLOT + cell * Dm::NP + q];
    __syncthreads();
    for (int it = tid; it < 9 * ND; it += NT) {
      const int Kc = it % 9, mk = it / 9;  // H[Kc + 9 mk], Kc = K + 3c, mk = m + NN k (tensor order)
      const int k = mk / NN, m = mk - k * NN;
      double s = 0;
#pragma unroll
      for (int M = 0; M < 3; ++M) s += Dq[Kc + 9 * (M + 3 * k)] * G[3 * m + M];
      H[it] = s;
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < NE; ++e) {
      const int idx = tid + e * NT;
      if (idx < ND * ND) {
        const int nc_ = idx % ND, mk = idx / ND;  // row (n, c) and column (m, k), tensor order c*NN + n
        const int c = nc_ / NN, n = nc_ - c * NN;
        double s = 0;
#pragma unroll
        for (int K = 0; K < 3; ++K) s += G[3 * n + K] * H[(K + 3 * c) + 9 * mk];
        acc[e] += s;
      }
    }
    __syncthreads();
  }
  const int32_t* pos = op.nzpos + ((size_t)batch * C + cell) * ND * ND;
#pragma unroll
  for (int e = 0; e < NE; ++e) {
    const int idx = tid + e * NT;
    if (idx < ND * ND) {
      const int ii = op.loc_of[idx % ND], jj = op.loc_of[idx / ND];
      const int p = pos[ii + ND * jj];
      if (p > 0) nzval[p - 1] += acc[e];  // race-free: the cells of one colour share no free test row
    }
  }
}

__global__ void k_fill(double* __restrict__ y, int64_t n, double v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) y[i] = v;
}

// Expand the stored tangents of every batch to D[(batch*C + cell)*np + q][81] (FP64), for inspection.
template <int P, class TS, bool SYM>
__global__ void k_expand(const void* __restrict__ tangents, double* __restrict__ out, int64_t nbatches) {
  constexpr int NP = P * P * P, C = kThreads / NP, NSLOT = C * NP, NE = SYM ? 45 : 81;
  const int64_t n = nbatches * NSLOT;
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < n; g += (int64_t)gridDim.x * blockDim.x) {
    const int64_t batch = g / NSLOT;
    const int slot = (int)(g - batch * NSLOT);
    const TS* Dp = static_cast<const TS*>(tangents) + batch * NE * NSLOT + slot;
    for (int j = 0; j < 9; ++j)
      for (int i = 0; i < 9; ++i) {
        const int e = SYM ? (i <= j ? sym_index(i, j) : sym_index(j, i)) : i + 9 * j;
        out[g * 81 + i + 9 * j] = static_cast<double>(Dp[(size_t)e * NSLOT]);
      }
  }
}

// ---------------------------------------------------------------------------------------------------------
// dispatch helpers
template <int Q, int P, class Law>
hipError_t prepare_qp(const OpDev& op, int storage, void* T, const double* x, BatchRange r, hipStream_t s) {
  const dim3 g(r.count), b(kThreads);
  switch (storage) {
    case 0: hipLaunchKernelGGL((k_prepare<Q, P, Law, double, false>), g, b, 0, s, op, T, x, r.first); break;
    case 1: hipLaunchKernelGGL((k_prepare<Q, P, Law, double, true>), g, b, 0, s, op, T, x, r.first); break;
    case 2: hipLaunchKernelGGL((k_prepare<Q, P, Law, float, false>), g, b, 0, s, op, T, x, r.first); break;
    default: hipLaunchKernelGGL((k_prepare<Q, P, Law, float, true>), g, b, 0, s, op, T, x, r.first); break;
  }
  return hipGetLastError();
}

template <int Q, int P>
hipError_t prepare_law(const OpDev& op, int storage, void* T, const double* x, BatchRange r, hipStream_t s) {
  switch (op.law) {
    case 0: return prepare_qp<Q, P, LawSVK>(op, storage, T, x, r, s);
    case 1: return prepare_qp<Q, P, LawNeoHooke>(op, storage, T, x, r, s);
    default: return prepare_qp<Q, P, LawLinear>(op, storage, T, x, r, s);
  }
}

template <int Q, int P, bool ATOMIC>
hipError_t apply_qpa(const OpDev& op, int storage, const void* T, const double* dx, double* y, BatchRange r,
                     hipStream_t s, double alpha, const ApplyMap* m) {
  const dim3 g(ATOMIC ? m->grid : r.count), b(kThreads);
  const int* fc = ATOMIC ? m->colfc : nullptr;
  const int nc = ATOMIC ? m->ncol : 0, gx = ATOMIC ? m->gx : 0;
  switch (storage) {
    case 0: hipLaunchKernelGGL((k_apply<Q, P, double, false, ATOMIC>), g, b, 0, s, op, T, dx, y, r.first, alpha, fc, nc, gx); break;
    case 1: hipLaunchKernelGGL((k_apply<Q, P, double, true, ATOMIC>), g, b, 0, s, op, T, dx, y, r.first, alpha, fc, nc, gx); break;
    case 2: hipLaunchKernelGGL((k_apply<Q, P, float, false, ATOMIC>), g, b, 0, s, op, T, dx, y, r.first, alpha, fc, nc, gx); break;
    default: hipLaunchKernelGGL((k_apply<Q, P, float, true, ATOMIC>), g, b, 0, s, op, T, dx, y, r.first, alpha, fc, nc, gx); break;
  }
  return hipGetLastError();
}

template <int Q, int P>
hipError_t apply_qp(const OpDev& op, int storage, const void* T, const double* dx, double* y, BatchRange r,
                    hipStream_t s, double alpha, const ApplyMap* m) {
  return m ? apply_qpa<Q, P, true>(op, storage, T, dx, y, r, s, alpha, m)
           : apply_qpa<Q, P, false>(op, storage, T, dx, y, r, s, alpha, m);
}

template <int Q, int P>
hipError_t blocks_qp(const OpDev& op, int storage, const void* T, const int32_t* block_of, double* B, BatchRange r,
                     hipStream_t s) {
  const dim3 g(r.count), b(kThreads);
  switch (storage) {
    case 0: hipLaunchKernelGGL((k_blocks<Q, P, double, false>), g, b, 0, s, op, T, block_of, B, r.first); break;
    case 1: hipLaunchKernelGGL((k_blocks<Q, P, double, true>), g, b, 0, s, op, T, block_of, B, r.first); break;
    case 2: hipLaunchKernelGGL((k_blocks<Q, P, float, false>), g, b, 0, s, op, T, block_of, B, r.first); break;
    default: hipLaunchKernelGGL((k_blocks<Q, P, float, true>), g, b, 0, s, op, T, block_of, B, r.first); break;
  }
  return hipGetLastError();
}

template <int Q, int P>
hipError_t residual_law(const OpDev& op, const double* x, double* bvec, BatchRange r, hipStream_t s) {
  const dim3 g(r.count), b(kThreads);
  switch (op.law) {
    case 0: hipLaunchKernelGGL((k_residual<Q, P, LawSVK>), g, b, 0, s, op, x, bvec, r.first); break;
    case 1: hipLaunchKernelGGL((k_residual<Q, P, LawNeoHooke>), g, b, 0, s, op, x, bvec, r.first); break;
    default: hipLaunchKernelGGL((k_residual<Q, P, LawLinear>), g, b, 0, s, op, x, bvec, r.first); break;
  }
  return hipGetLastError();
}

template <int Q, int P>
hipError_t point_state_qp(const OpDev& op, const double* x, double* st, BatchRange r, hipStream_t s) {
  hipLaunchKernelGGL((k_point_state<Q, P>), dim3(r.count), dim3(kThreads), 0, s, op, x, st, r.first);
  return hipGetLastError();
}

template <int Q, int P>
hipError_t residual_coefs_qp(const OpDev& op, const double* coef, double* b, BatchRange r, hipStream_t s) {
  hipLaunchKernelGGL((k_residual_coefs<Q, P>), dim3(r.count), dim3(kThreads), 0, s, op, coef, b, r.first);
  return hipGetLastError();
}

template <int Q, int P>
hipError_t assemble_qp(const OpDev& op, const double* D, double* nzval, BatchRange r, hipStream_t s) {
  constexpr int C = Dims<Q, P>::C;
  constexpr int NT = Q >= 4 ? 1024 : 256;
  hipLaunchKernelGGL((k_assemble<Q, P, NT>), dim3(r.count * C), dim3(NT), 0, s, op, D, nzval, r.first);
  return hipGetLastError();
}

// The (Q, P) pairs with kernels: Q1, Q2, Q3 with P = Q (degree 2 order) and P = Q + 1; Q1 with the points of
// the finer levels (the coarse level of p-multigrid keeps the fine quadrature, src/pmultigrid.jl:248-251).
#define HFK_DISPATCH_QP(Q_, P_, CALL)                                   \
  do {                                                                  \
    const int q__ = (Q_), p__ = (P_);                                   \
    if (q__ == 2 && p__ == 2) return CALL(2, 2);                        \
    if (q__ == 2 && p__ == 3) return CALL(2, 3);                        \
    if (q__ == 2 && p__ == 4) return CALL(2, 4);                        \
    if (q__ == 3 && p__ == 3) return CALL(3, 3);                        \
    if (q__ == 3 && p__ == 4) return CALL(3, 4);                        \
    if (q__ == 4 && p__ == 4) return CALL(4, 4);                        \
    return hipErrorInvalidValue;                                        \
  } while (0)

}  // namespace

bool supported_qp(int Q, int P) {
  return (Q == 2 && (P == 2 || P == 3 || P == 4)) || (Q == 3 && (P == 3 || P == 4)) || (Q == 4 && P == 4);
}

hipError_t launch_prepare(const OpDev& op, int storage, void* T, const double* x, BatchRange r, hipStream_t s) {
  if (r.count <= 0) return hipSuccess;
#define CALL(q, p) prepare_law<q, p>(op, storage, T, x, r, s)
  HFK_DISPATCH_QP(op.Q, op.P, CALL);
#undef CALL
}

hipError_t launch_apply(const OpDev& op, int storage, const void* T, const double* dx, double* y, BatchRange r,
                        hipStream_t s, double alpha, const ApplyMap* m) {
  if (!m && r.count <= 0) return hipSuccess;
#define CALL(q, p) apply_qp<q, p>(op, storage, T, dx, y, r, s, alpha, m)
  HFK_DISPATCH_QP(op.Q, op.P, CALL);
#undef CALL
}

hipError_t launch_blocks(const OpDev& op, int storage, const void* T, const int32_t* block_of, double* B, BatchRange r,
                         hipStream_t s) {
  if (r.count <= 0) return hipSuccess;
#define CALL(q, p) blocks_qp<q, p>(op, storage, T, block_of, B, r, s)
  HFK_DISPATCH_QP(op.Q, op.P, CALL);
#undef CALL
}

hipError_t launch_residual(const OpDev& op, const double* x, double* b, BatchRange r, hipStream_t s) {
  if (r.count <= 0) return hipSuccess;
#define CALL(q, p) residual_law<q, p>(op, x, b, r, s)
  HFK_DISPATCH_QP(op.Q, op.P, CALL);
#undef CALL
}

hipError_t launch_point_state(const OpDev& op, const double* x, double* st, BatchRange r, hipStream_t s) {
  if (r.count <= 0) return hipSuccess;
#define CALL(q, p) point_state_qp<q, p>(op, x, st, r, s)
  HFK_DISPATCH_QP(op.Q, op.P, CALL);
#undef CALL
}

hipError_t launch_residual_coefs(const OpDev& op, const double* coef, double* b, BatchRange r, hipStream_t s) {
  if (r.count <= 0) return hipSuccess;
#define CALL(q, p) residual_coefs_qp<q, p>(op, coef, b, r, s)
  HFK_DISPATCH_QP(op.Q, op.P, CALL);
#undef CALL
}

hipError_t launch_assemble(const OpDev& op, const double* D, double* nzval, BatchRange r, hipStream_t s) {
  if (r.count <= 0) return hipSuccess;
#define CALL(q, p) assemble_qp<q, p>(op, D, nzval, r, s)
  HFK_DISPATCH_QP(op.Q, op.P, CALL);
#undef CALL
}

hipError_t launch_fill(double* y, int64_t n, double v, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  int64_t blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_fill, dim3((unsigned)blocks), dim3(256), 0, s, y, n, v);
  return hipGetLastError();
}

hipError_t launch_expand_tangents(const OpDev& op, int storage, const void* T, double* D, int64_t nbatches,
                                  hipStream_t s) {
  const dim3 g(1024), b(256);
#define EXP(P_)                                                                                           \
  switch (storage) {                                                                                      \
    case 0: hipLaunchKernelGGL((k_expand<P_, double, false>), g, b, 0, s, T, D, nbatches); break;          \
    case 1: hipLaunchKernelGGL((k_expand<P_, double, true>), g, b, 0, s, T, D, nbatches); break;           \
    case 2: hipLaunchKernelGGL((k_expand<P_, float, false>), g, b, 0, s, T, D, nbatches); break;           \
    default: hipLaunchKernelGGL((k_expand<P_, float, true>), g, b, 0, s, T, D, nbatches); break;           \
  }
  if (op.P == 2) {
    EXP(2)
  } else if (op.P == 3) {
    EXP(3)
  } else if (op.P == 4) {
    EXP(4)
  } else {
    return hipErrorInvalidValue;
  }
#undef EXP
  return hipGetLastError();
}

}  // namespace hfk
