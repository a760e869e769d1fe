// Device code of the pointwise phase of libhfk: batch dimensions, 3x3 helpers, the built-in constitutive laws,
// the sum-factorised interpolation and contraction sweeps, the trilinear geometry, and the prepare! and
// residual! kernels templated on the law.  Compiled by hipcc into libhfk (hfk_kernels.hip instantiates the
// built-in laws) and, embedded in the library as text, by hiprtc for laws given as source at run time.  Uses
// nothing beyond the HIP device runtime, so both compilers accept it.
#pragma once
#include "hfk_types.hpp"

namespace hfk {

template <int Q, int P>
struct Dims {
  static constexpr int NN = Q * Q * Q;
  static constexpr int NP = P * P * P;
  static constexpr int C = cells_per_batch(P);
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
  // W = lam/2 tr(E)^2 + mu E:E
  __device__ double energy() const {
    double tr = 0, ee = 0;
#pragma unroll
    for (int I = 0; I < 3; ++I)
#pragma unroll
      for (int J = 0; J < 3; ++J) {
        double s = 0;
#pragma unroll
        for (int k = 0; k < 3; ++k) s += F[k + 3 * I] * F[k + 3 * J];
        const double e = 0.5 * (s - (I == J ? 1.0 : 0.0));
        ee += e * e;
        if (I == J) tr += e;
      }
    return 0.5 * lam * tr * tr + mu * ee;
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
  __device__ double energy() const {
    double I1 = 0;
#pragma unroll
    for (int i = 0; i < 9; ++i) I1 += F[i] * F[i];
    return 0.5 * mu * (I1 - 3.0) - mu * lnJ + 0.5 * lam * lnJ * lnJ;
  }
  // mu d_ck d_JL + (mu - lam lnJ) A_Jk A_Lc + lam A_Jc A_Lk
  __device__ double tangent(int c, int J, int k, int L) const {
    return (c == k && J == L ? mu : 0.0) + (mu - lam * lnJ) * A[J + 3 * k] * A[L + 3 * c] + lam * A[J + 3 * c] * A[L + 3 * k];
  }
  // the tangent in the isotropic factorised form mu d_ck d_JL + m Ao_Jk Ao_Lc + lam Ao_Jc Ao_Lk
  __device__ void factor(double* Ao, double& m) const {
#pragma unroll
    for (int i = 0; i < 9; ++i) Ao[i] = A[i];
    m = mu - lam * lnJ;
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
  // W = lam/2 tr(eps)^2 + mu eps:eps, eps = (G + G')/2
  __device__ double energy() const {
    double ee = 0;
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int J = 0; J < 3; ++J) {
        const double e = 0.5 * (G[c + 3 * J] + G[J + 3 * c]);
        ee += e * e;
      }
    return 0.5 * lam * tr * tr + mu * ee;
  }
  __device__ double tangent(int c, int J, int k, int L) const {
    return (c == J && k == L ? lam : 0.0) + (c == k && J == L ? mu : 0.0) + (J == k && c == L ? mu : 0.0);
  }
  // the factorised form of the tangent (see LawNeoHooke::factor) with Ao = I and m = mu
  __device__ void factor(double* Ao, double& m) const {
#pragma unroll
    for (int i = 0; i < 9; ++i) Ao[i] = (i % 4 == 0) ? 1.0 : 0.0;
    m = mu;
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

// Physical coordinates of tensor point q (P points per direction) under the trilinear map of the vertices X[v][i]:
// the points at which a function of x is materialised (src/coordfunctions.jl:186-213).
__device__ __forceinline__ void point_coordinates(const Tab& tb, const double* __restrict__ X, int q, int P, double* x) {
  const int a = q % P, b = (q / P) % P, cc = q / (P * P);
  const double xi[3] = {tb.xq[a], tb.xq[b], tb.xq[cc]};
  x[0] = x[1] = x[2] = 0.0;
#pragma unroll
  for (int v = 0; v < 8; ++v) {
    const int bx = v & 1, by = (v >> 1) & 1, bz = (v >> 2) & 1;
    const double m = (bx ? xi[0] : 1 - xi[0]) * (by ? xi[1] : 1 - xi[1]) * (bz ? xi[2] : 1 - xi[2]);
#pragma unroll
    for (int i = 0; i < 3; ++i) x[i] += m * X[v * 3 + i];
  }
}

// Index of the packed upper triangle (i <= j) of a symmetric 9x9 matrix.
__host__ __device__ constexpr int sym_index(int i, int j) { return j * (j + 1) / 2 + i; }

// Layouts of the stored tangent D[(K,c),(M,k)] of one point (i = K + 3c, j = M + 3k):
//   kLayFull  81 entries D[i + 9j];
//   kLaySym   45 entries, the upper triangle of a symmetric D at sym_index(i, j);
//   kLayFac   17 factors of an isotropic tangent of the form T = mu d_ck d_JL + m A_Jk A_Lc + lam A_Jc A_Lk (the
//             built-in neo-Hookean law, A = F^-1, m = mu - lam lnJ; the linear law, A = I, m = mu).  With
//             H = iJt' A (H[K + 3c] = sum_J iJt[J,K] A_Jc) and the metric g = iJt' iJt:
//               D[(K,c),(M,k)] = G_KM d_ck + b H_Kk H_Mc + l H_Kc H_Mk,
//             stored as G = dV mu g (entries 0..5, packed upper triangle sym_index(K, M)), H (6..14),
//             b = dV m (15), l = dV lam (16).
//   kLayAfac  the same tangent on affine cells (parallelepipeds: iJt and det J constant per cell, dV = w_q det J),
//             11 entries per point: H (0..8), b = dV m (9), g = dV mu (10); per cell (OpDev::cellgeo) the metric
//             iJt' iJt (packed, 0..5) and det J (6), so that G = g iJt' iJt and l = lam w_q det J.
constexpr int kLayFull = 0, kLaySym = 1, kLayFac = 2, kLayAfac = 3;
__host__ __device__ constexpr int lay_entries(int lay) {
  return lay == kLayFull ? 81 : (lay == kLaySym ? 45 : (lay == kLayFac ? 17 : 11));
}

// the quadrature weight of tensor point q = a + P b + P^2 c
template <int P>
__device__ __forceinline__ double point_weight(const Tab& tb, int q) {
  return tb.wq[q % P] * tb.wq[(q / P) % P] * tb.wq[q / (P * P)];
}

// entry D[i + 9j] of a point stored in the affine factorised layout (f[e * stride] is entry e, cg the cell's
// geometry, w the point's weight)
template <class TS>
__device__ __forceinline__ double afac_entry(const TS* f, size_t stride, const double* cg, double w, double lam, int i,
                                             int j) {
  const int K = i % 3, c = i / 3, M = j % 3, k = j / 3;
  const double hKk = (double)f[(K + 3 * k) * stride], hMc = (double)f[(M + 3 * c) * stride];
  const double hKc = (double)f[(K + 3 * c) * stride], hMk = (double)f[(M + 3 * k) * stride];
  const double g = c == k ? (double)f[10 * stride] * cg[K <= M ? sym_index(K, M) : sym_index(M, K)] : 0.0;
  return g + (double)f[9 * stride] * hKk * hMc + lam * w * cg[6] * hKc * hMk;
}

// entry D[i + 9j] of a point stored in the factorised layout; f[e * stride] is factor e
template <class TS>
__device__ __forceinline__ double fac_entry(const TS* f, size_t stride, int i, int j) {
  const int K = i % 3, c = i / 3, M = j % 3, k = j / 3;
  const double hKk = (double)f[(6 + K + 3 * k) * stride], hMc = (double)f[(6 + M + 3 * c) * stride];
  const double hKc = (double)f[(6 + K + 3 * c) * stride], hMk = (double)f[(6 + M + 3 * k) * stride];
  const double g = c == k ? (double)f[(K <= M ? sym_index(K, M) : sym_index(M, K)) * stride] : 0.0;
  return g + (double)f[15 * stride] * hKk * hMc + (double)f[16 * stride] * hKc * hMk;
}

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
template <int Q, int P, class Law, class TS, int LAY>
__global__ __launch_bounds__(kThreads) void k_prepare(OpDev op, void* __restrict__ tangents,
                                                      const double* __restrict__ x, int batch0) {
  using Dm = Dims<Q, P>;
  constexpr bool SYM = LAY == kLaySym;
  constexpr int C = Dm::C, SA = Dm::SA, SB = Dm::SB, NSLOT = Dm::NSLOT, NE = lay_entries(LAY);
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
    if constexpr (LAY == kLayAfac) {  // per point H, b, g; per cell the metric and det J (from its first point)
      double Af[9], m;
      law.factor(Af, m);
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int K = 0; K < 3; ++K) {
          double h = 0;
#pragma unroll
          for (int J = 0; J < 3; ++J) h += iJt[J + 3 * K] * Af[J + 3 * c];
          store_entry<TS, true>(tangents, base + (size_t)(K + 3 * c) * NSLOT, h);
        }
      store_entry<TS, true>(tangents, base + (size_t)9 * NSLOT, dV * m);
      store_entry<TS, true>(tangents, base + (size_t)10 * NSLOT, dV * mu);
      if (q == 0) {
        double* cg = op.cellgeo + ((size_t)batch * C + cell) * 8;
#pragma unroll
        for (int M = 0; M < 3; ++M)
#pragma unroll
          for (int K = 0; K <= M; ++K) {
            double g = 0;
#pragma unroll
            for (int J = 0; J < 3; ++J) g += iJt[J + 3 * K] * iJt[J + 3 * M];
            cg[sym_index(K, M)] = g;
          }
        cg[6] = dV / point_weight<P>(op.tab, 0);
      }
      return;
    }
    if constexpr (LAY == kLayFac) {
      double Af[9], m;
      law.factor(Af, m);
#pragma unroll
      for (int M = 0; M < 3; ++M)
#pragma unroll
        for (int K = 0; K <= M; ++K) {
          double g = 0;
#pragma unroll
          for (int J = 0; J < 3; ++J) g += iJt[J + 3 * K] * iJt[J + 3 * M];
          store_entry<TS, true>(tangents, base + (size_t)sym_index(K, M) * NSLOT, dV * mu * g);
        }
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int K = 0; K < 3; ++K) {
          double h = 0;
#pragma unroll
          for (int J = 0; J < 3; ++J) h += iJt[J + 3 * K] * Af[J + 3 * c];
          store_entry<TS, true>(tangents, base + (size_t)(6 + K + 3 * c) * NSLOT, h);
        }
      store_entry<TS, true>(tangents, base + (size_t)15 * NSLOT, dV * m);
      store_entry<TS, true>(tangents, base + (size_t)16 * NSLOT, dV * op.lam);
      return;
    }
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
// Scalar functionals of the unknown (KernelFunctional, src/functionals.jl:9-154): the strain energy
// sum_cells sum_q W(F) dV (reduce = :sum) and the extremum over the quadrature points of a pointwise function of
// the displacement gradient and the gradient of an increment (reduce = :min / :max).  Each work-group reduces its
// batch in a fixed order into part[batch]; k_reduce_partials reduces the batches in a fixed order.

// 1 when the law defines `double energy() const` (detection by partial specialisation; no standard headers)
template <class L, class = void>
struct HasEnergy {
  static constexpr int value = 0;
};
template <class L>
struct HasEnergy<L, decltype((void)static_cast<const L*>(nullptr)->energy())> {
  static constexpr int value = 1;
};
template <class L>
__device__ __forceinline__ auto law_energy(const L& l, int) -> decltype(l.energy()) {
  return l.energy();
}
template <class L>
__device__ __forceinline__ double law_energy(const L&, long) {
  return __builtin_nan("");
}

// combine for MODE 0 (sum), 1 (min), 2 (max); min and max propagate NaN
template <int MODE>
__device__ __forceinline__ double combine(double a, double b) {
  if (MODE == 0) return a + b;
  if (a != a || b != b) return __builtin_nan("");
  return MODE == 1 ? (a < b ? a : b) : (a > b ? a : b);
}
template <int MODE>
__device__ __forceinline__ double identity_of() {
  return MODE == 0 ? 0.0 : (MODE == 1 ? __builtin_inf() : -__builtin_inf());
}

// fixed-order tree over the kThreads values of a work-group; the result is returned to every thread
template <int MODE>
__device__ __forceinline__ double block_reduce(double v, double* red) {
  const int tid = threadIdx.x;
  red[tid] = v;
  __syncthreads();
#pragma unroll
  for (int st = kThreads / 2; st > 0; st >>= 1) {
    if (tid < st) red[tid] = combine<MODE>(red[tid], red[tid + st]);
    __syncthreads();
  }
  return red[0];
}

// the nodal values of the C cells of a batch into A: free values x, Dirichlet values d (null: zero), 0 on padding
template <int Q, int P>
__device__ __forceinline__ void gather_batch(const int32_t* __restrict__ ids, const double* __restrict__ x,
                                             const double* __restrict__ d, double* A, int tid) {
  using Dm = Dims<Q, P>;
  for (int it = tid; it < Dm::C * 3 * Dm::NN; it += kThreads) {
    const int id = ids[it];
    const int cell = it / (3 * Dm::NN);
    A[cell * Dm::SA + (it - cell * 3 * Dm::NN)] = id > 0 ? x[id - 1] : (id < 0 && d ? d[-id - 1] : 0.0);
  }
}

// physical gradient G[c + 3J] = du_c/dx_J from the reference gradients s[3c + K] = du_c/dxi_K
__device__ __forceinline__ void physical_gradient(const double* iJt, const double* s, double* G) {
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int J = 0; J < 3; ++J) {
      double g = 0;
#pragma unroll
      for (int K = 0; K < 3; ++K) g += iJt[J + 3 * K] * s[3 * c + K];
      G[c + 3 * J] = g;
    }
}

template <int Q, int P, class Law>
__global__ __launch_bounds__(kThreads) void k_energy(OpDev op, const double* __restrict__ x, double* __restrict__ part) {
  using Dm = Dims<Q, P>;
  constexpr int C = Dm::C, SA = Dm::SA, SB = Dm::SB, NSLOT = Dm::NSLOT;
  __shared__ double lds[Dm::LDS];
  __shared__ double red[kThreads];
  double* A = lds;
  double* B = lds + C * SA;
  const int tid = threadIdx.x;
  const int batch = blockIdx.x;
  const int32_t* ids = op.ids + (size_t)batch * C * 3 * Dm::NN;
  gather_batch<Q, P>(ids, x, op.dvals, A, tid);
  __syncthreads();
  interp_sweeps<Q, P>(A, B, op.tab, tid);
  double e = 0.0;
  if (tid < NSLOT) {
    const int cell = tid / Dm::NP, q = tid - cell * Dm::NP;
    if (ids[cell * 3 * Dm::NN] != 0) {  // not a padding cell
      double iJt[9], dV, F[9];
      point_geometry<P>(op.tab, op.coords + ((size_t)batch * C + cell) * 24, q, iJt, dV);
      physical_gradient(iJt, B + cell * SB + q * 9, F);
#pragma unroll
      for (int c = 0; c < 3; ++c) F[c + 3 * c] += 1.0;
      const double mu = op.mu_pts ? op.mu_pts[(size_t)batch * NSLOT + tid] : op.mu;
      Law law;
      law.setup(F, op.lam, mu);
      e = dV * law_energy(law, 0);
    }
  }
  e = block_reduce<0>(e, red);
  if (tid == 0) part[batch] = e;
}

// Fn::eval(Gu, Gdu, p): u the free values x with the Dirichlet values of the operator, du the free values dx with
// the Dirichlet values ddir (null: zero); MODE 1 min, 2 max over the points, no weights
template <int Q, int P, class Fn, int MODE>
__global__ __launch_bounds__(kThreads) void k_extremum(OpDev op, const double* __restrict__ x,
                                                       const double* __restrict__ dx, const double* __restrict__ ddir,
                                                       const double* __restrict__ prm, double* __restrict__ part) {
  using Dm = Dims<Q, P>;
  constexpr int C = Dm::C, SA = Dm::SA, SB = Dm::SB, NSLOT = Dm::NSLOT;
  __shared__ double lds[Dm::LDS];
  __shared__ double red[kThreads];
  double* A = lds;
  double* B = lds + C * SA;
  const int tid = threadIdx.x;
  const int batch = blockIdx.x;
  const int32_t* ids = op.ids + (size_t)batch * C * 3 * Dm::NN;
  const int cell = tid / Dm::NP, q = tid - cell * Dm::NP;
  const bool real = tid < NSLOT && ids[cell * 3 * Dm::NN] != 0;
  double iJt[9], dV, Gu[9];
  gather_batch<Q, P>(ids, x, op.dvals, A, tid);
  __syncthreads();
  interp_sweeps<Q, P>(A, B, op.tab, tid);
  if (real) {
    point_geometry<P>(op.tab, op.coords + ((size_t)batch * C + cell) * 24, q, iJt, dV);
    physical_gradient(iJt, B + cell * SB + q * 9, Gu);
  }
  __syncthreads();  // B read before the increment overwrites A and B
  gather_batch<Q, P>(ids, dx, ddir, A, tid);
  __syncthreads();
  interp_sweeps<Q, P>(A, B, op.tab, tid);
  double v = identity_of<MODE>();
  if (real) {
    double Gdu[9];
    physical_gradient(iJt, B + cell * SB + q * 9, Gdu);
    v = Fn::eval(Gu, Gdu, prm);
  }
  v = block_reduce<MODE>(v, red);
  if (tid == 0) part[batch] = v;
}

template <int MODE>
__global__ __launch_bounds__(kThreads) void k_reduce_partials(const double* __restrict__ part, int64_t n,
                                                              double* __restrict__ out) {
  __shared__ double red[kThreads];
  constexpr int U = 8;  // loads issued together; each thread still combines part[t], part[t + 256], ... in order
  double v = identity_of<MODE>();
  for (int64_t i0 = threadIdx.x; i0 < n; i0 += (int64_t)U * kThreads) {
    double buf[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t i = i0 + (int64_t)u * kThreads;
      buf[u] = i < n ? part[i] : identity_of<MODE>();
    }
#pragma unroll
    for (int u = 0; u < U; ++u) v = combine<MODE>(v, buf[u]);
  }
  v = block_reduce<MODE>(v, red);
  if (threadIdx.x == 0) out[0] = v;
}

}  // namespace hfk
