// Face integrals of the hexahedral operator, compiled at run time with hiprtc for the operator's (Q, P) and a point
// form (this header is embedded in the library):
// * skeleton integrals (NS = 2): interior facets with a cell on each side (plus, minus), the facet kernels of
//   residual!, the matrix-free product, jacobian! and the energy (src/skeleton.jl);
// * boundary integrals (NS = 1): faces of one cell with a captured integrand (the boundary groups of
//   src/assembly.jl:539-578), e.g. a follower pressure.
//
// A point form is a struct with three static device functions of the 12 NS slots at a face point
//   s[0..2]    u on the plus side,  s[3 + 3c + K]  du_c/dx_K on the plus side,
//   s[12..14]  u on the minus side, s[15 + 3c + K] du_c/dx_K on the minus side (NS = 2),
// the physical point x, the unit normal n (outward from the plus cell, reference configuration), h = sqrt(area of
// the face) and the parameters p[8]:
//   residual(s, x, n, h, p, R)   R[12 NS]: the coefficient of every test slot (r(v) = int R . v_slots dS)
//   jvp(s, ds, x, n, h, p, dR)   dR = (dR/ds) ds, the linearisation of R at s in the direction ds
//   energy(s, x, n, h, p)        the energy density (its gradient is R; 0 if the form has none)
// Facet geometry: P x P Gauss points on the plus face, coordinates (a, b) along its tangential directions
// (e1, e2); the minus cell's point is the same physical point, mapped through the facet orientation o:
// (s', t') = o & 1 ? (t, s) : (s, t), then s' -> 1 - s' if o & 2, t' -> 1 - t' if o & 4.  The measure is the plus
// face's (src/skeleton.jl:1-16: `w meas(Jf+ . Jt+)`).
#pragma once
#include "hfk_types.hpp"

namespace hfk {

struct SkelDev {
  const int64_t* slot;    // [f][2] batch slots of the plus and the minus cell
  const int32_t* face;    // [f][3] plus local face, minus local face, orientation
  const int32_t* order;   // the facets of this launch (one colour): order[blockIdx.x / split]
  const int64_t* colptr;  // CSC pattern, 0-based (the Jacobian's positions); null for the other kernels
  const int32_t* rowval;
  double p[8];
};

// local face lf = 2 d + side is the face xi_d = side; (e1, e2) its tangential directions in increasing order
__host__ __device__ inline void skel_face_dirs(int lf, int& d, int& side, int& e1, int& e2) {
  d = lf >> 1;
  side = lf & 1;
  e1 = d == 0 ? 1 : 0;
  e2 = d == 2 ? 1 : 2;
}

// 1D Lagrange basis on Q equispaced nodes of [0, 1]: value and derivative of function i at x
template <int Q>
__device__ __forceinline__ void skel_lag(double x, int i, double& v, double& dv) {
  constexpr double h = 1.0 / (Q - 1);
  v = 1.0;
  dv = 0.0;
#pragma unroll
  for (int m = 0; m < Q; ++m) {
    if (m == i) continue;
    const double den = (i - m) * h;
    const double f = (x - m * h) / den;
    dv = dv * f + v / den;
    v *= f;
  }
}

// dx_i/dxi_K of the trilinear map of the 8 vertices X[v][i] (vertex v = bx + 2 by + 4 bz) at xi
__device__ __forceinline__ void skel_jacobian(const double* __restrict__ X, const double* xi, double* J) {
#pragma unroll
  for (int e = 0; e < 9; ++e) J[e] = 0.0;
  for (int v = 0; v < 8; ++v) {
    double m[3], dm[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const int bit = (v >> k) & 1;
      m[k] = bit ? xi[k] : 1.0 - xi[k];
      dm[k] = bit ? 1.0 : -1.0;
    }
    const double g[3] = {dm[0] * m[1] * m[2], m[0] * dm[1] * m[2], m[0] * m[1] * dm[2]};
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int K = 0; K < 3; ++K) J[i + 3 * K] += X[v * 3 + i] * g[K];
  }
}

// the reference point of side `side` (0 plus, 1 minus) of facet point (a, b)
template <int P>
__device__ __forceinline__ void skel_xi(const Tab& tb, int lf, int orient, bool minus, int a, int b, double* xi) {
  int d, sd, e1, e2;
  skel_face_dirs(lf, d, sd, e1, e2);
  double s = tb.xq[a], t = tb.xq[b];
  if (minus) {
    if (orient & 1) {
      const double w = s;
      s = t;
      t = w;
    }
    if (orient & 2) s = 1.0 - s;
    if (orient & 4) t = 1.0 - t;
  }
  xi[d] = sd;
  xi[e1] = s;
  xi[e2] = t;
}

template <int Q, int P>
struct SkelDims {
  static constexpr int NN = Q * Q * Q, ND = 3 * NN, NQ = P * P;
};

// Shared part of every face kernel: the ids and vertices of the face's NS cells, and the geometry of every face
// point (measure, normal, point, and per side the shape values and physical shape gradients).
template <int Q, int P, int NS>
struct SkelStage {
  using Dm = SkelDims<Q, P>;
  int32_t id[NS * Dm::ND];            // [side][c][t] tensor order
  double X[NS][24];                   // vertices of the cells
  double Nv[NS][Dm::NQ][Dm::NN];      // shape values at the points, per side
  double G[NS][Dm::NQ][Dm::NN][3];    // physical shape gradients
  double dS[Dm::NQ], nrm[Dm::NQ][3], xq[Dm::NQ][3];
  double h;
};

template <int Q, int P, int NS>
__device__ void skel_geometry(const OpDev& op, const SkelDev& sk, int64_t f, SkelStage<Q, P, NS>& st, int tid,
                              int nthreads) {
  using Dm = SkelDims<Q, P>;
  const int64_t sp = sk.slot[2 * f], sm = sk.slot[2 * f + 1];
  const int lfp = sk.face[3 * f], lfm = sk.face[3 * f + 1], orient = sk.face[3 * f + 2];
  for (int it = tid; it < NS * Dm::ND; it += nthreads) {
    const int side = it / Dm::ND, k = it - side * Dm::ND;
    st.id[it] = op.ids[(side ? sm : sp) * Dm::ND + k];
  }
  for (int it = tid; it < NS * 24; it += nthreads) st.X[it / 24][it % 24] = op.coords[(it < 24 ? sp : sm) * 24 + it % 24];
  __syncthreads();
  // per (side, point): Jacobian, inverse, shape values and physical gradients; the plus side also dS, n and x
  for (int it = tid; it < NS * Dm::NQ; it += nthreads) {
    const int side = it / Dm::NQ, q = it - side * Dm::NQ, a = q % P, b = q / P;
    double xi[3], J[9];
    skel_xi<P>(op.tab, side ? lfm : lfp, orient, side == 1, a, b, xi);
    skel_jacobian(st.X[side], xi, J);
    // inverse of J (J[i + 3K] = dx_i/dxi_K): iJ[K + 3i] = dxi_K/dx_i
    const double det = J[0] * (J[4] * J[8] - J[7] * J[5]) - J[3] * (J[1] * J[8] - J[7] * J[2]) +
                       J[6] * (J[1] * J[5] - J[4] * J[2]);
    const double id = 1.0 / det;
    double iJ[9];
    iJ[0] = (J[4] * J[8] - J[7] * J[5]) * id;
    iJ[3] = -(J[3] * J[8] - J[6] * J[5]) * id;
    iJ[6] = (J[3] * J[7] - J[6] * J[4]) * id;
    iJ[1] = -(J[1] * J[8] - J[7] * J[2]) * id;
    iJ[4] = (J[0] * J[8] - J[6] * J[2]) * id;
    iJ[7] = -(J[0] * J[7] - J[6] * J[1]) * id;
    iJ[2] = (J[1] * J[5] - J[4] * J[2]) * id;
    iJ[5] = -(J[0] * J[5] - J[3] * J[2]) * id;
    iJ[8] = (J[0] * J[4] - J[3] * J[1]) * id;
    // iJ[r + 3 col] is the inverse of J in the same column-major layout: (J^-1)[K][i] = iJ[K + 3 i]
    for (int t = 0; t < Dm::NN; ++t) {
      const int ti = t % Q, tj = (t / Q) % Q, tk = t / (Q * Q);
      double v0, d0, v1, d1, v2, d2;
      skel_lag<Q>(xi[0], ti, v0, d0);
      skel_lag<Q>(xi[1], tj, v1, d1);
      skel_lag<Q>(xi[2], tk, v2, d2);
      const double gx[3] = {d0 * v1 * v2, v0 * d1 * v2, v0 * v1 * d2};
      st.Nv[side][q][t] = v0 * v1 * v2;
#pragma unroll
      for (int i = 0; i < 3; ++i) st.G[side][q][t][i] = gx[0] * iJ[0 + 3 * i] + gx[1] * iJ[1 + 3 * i] + gx[2] * iJ[2 + 3 * i];
    }
    if (side == 0) {
#pragma unroll
      for (int i = 0; i < 3; ++i) st.xq[q][i] = 0.0;
      for (int v = 0; v < 8; ++v) {
        const double w = (((v >> 0) & 1) ? xi[0] : 1.0 - xi[0]) * (((v >> 1) & 1) ? xi[1] : 1.0 - xi[1]) *
                         (((v >> 2) & 1) ? xi[2] : 1.0 - xi[2]);
#pragma unroll
        for (int i = 0; i < 3; ++i) st.xq[q][i] += w * st.X[0][v * 3 + i];
      }
      int d, sd, e1, e2;
      skel_face_dirs(lfp, d, sd, e1, e2);
      const double g1[3] = {J[0 + 3 * e1], J[1 + 3 * e1], J[2 + 3 * e1]};
      const double g2[3] = {J[0 + 3 * e2], J[1 + 3 * e2], J[2 + 3 * e2]};
      double n[3] = {g1[1] * g2[2] - g1[2] * g2[1], g1[2] * g2[0] - g1[0] * g2[2], g1[0] * g2[1] - g1[1] * g2[0]};
      const double nn = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
      // outward from the plus cell: along +dx/dxi_d on the face xi_d = 1, against it on xi_d = 0
      const double gd = n[0] * J[0 + 3 * d] + n[1] * J[1 + 3 * d] + n[2] * J[2 + 3 * d];
      const double sg = ((gd > 0) == (sd == 1)) ? 1.0 / nn : -1.0 / nn;
      st.dS[q] = op.tab.wq[a] * op.tab.wq[b] * nn;
#pragma unroll
      for (int i = 0; i < 3; ++i) st.nrm[q][i] = n[i] * sg;
    }
  }
  __syncthreads();
  if (tid == 0) {
    double area = 0.0;
    for (int q = 0; q < Dm::NQ; ++q) area += st.dS[q];
    st.h = sqrt(area);
  }
}

// the 12 NS slots of the nodal values U[side][c][t] at every point: S[q][24]
template <int Q, int P, int NS>
__device__ void skel_slots(const SkelStage<Q, P, NS>& st, const double* __restrict__ U, double (*S)[24], int tid,
                           int nthreads) {
  using Dm = SkelDims<Q, P>;
  for (int it = tid; it < Dm::NQ * 3 * NS; it += nthreads) {
    const int q = it / (3 * NS), r = it - 3 * NS * q, side = r / 3, c = r - 3 * side;
    const double* u = U + side * Dm::ND + c * Dm::NN;
    double v = 0.0, g0 = 0.0, g1 = 0.0, g2 = 0.0;
    for (int t = 0; t < Dm::NN; ++t) {
      const double ut = u[t];
      v += st.Nv[side][q][t] * ut;
      g0 += st.G[side][q][t][0] * ut;
      g1 += st.G[side][q][t][1] * ut;
      g2 += st.G[side][q][t][2] * ut;
    }
    double* s = S[q] + 12 * side;
    s[c] = v;
    s[3 + 3 * c + 0] = g0;
    s[3 + 3 * c + 1] = g1;
    s[3 + 3 * c + 2] = g2;
  }
}

// out[id - 1] += alpha sum_q (R[q] . slots of the test function (side, c, t)) for every free dof of both cells:
// the plus side, then the minus side (the face's nodes are dofs of both cells)
template <int Q, int P, int NS>
__device__ void skel_contract(const SkelStage<Q, P, NS>& st, const double (*R)[24], double alpha,
                              double* __restrict__ out, int tid, int nthreads) {
  using Dm = SkelDims<Q, P>;
  for (int side = 0; side < NS; ++side) {
    if (side == 1) __syncthreads();  // the plus side's writes to the shared face dofs come first
    for (int k = tid; k < Dm::ND; k += nthreads) {
      const int c = k / Dm::NN, t = k - c * Dm::NN;
      const int32_t id = st.id[side * Dm::ND + k];
      if (id <= 0) continue;
      double r = 0.0;
      for (int q = 0; q < Dm::NQ; ++q) {
        const double* Rq = R[q] + 12 * side;
        r += Rq[c] * st.Nv[side][q][t] + Rq[3 + 3 * c] * st.G[side][q][t][0] + Rq[4 + 3 * c] * st.G[side][q][t][1] +
             Rq[5 + 3 * c] * st.G[side][q][t][2];
      }
      out[id - 1] += alpha * r;  // race-free: the facets of one colour share no free dof of their cells
    }
  }
}

__device__ __forceinline__ double skel_value(int32_t id, const double* __restrict__ uf, const double* __restrict__ ud) {
  return id > 0 ? uf[id - 1] : (id < 0 && ud ? ud[-id - 1] : 0.0);
}

// residual! (KIND 0): b += r of the facets of the launch at state (xf, dvals).
// Jv (KIND 1): y += alpha (dr/dx) dx at the prepared state (xf, dvals).
// energy (KIND 2): part[blockIdx.x] = the facet's energy.
template <int Q, int P, class F, int KIND, int NS>
__global__ __launch_bounds__(256) void k_skel(OpDev op, SkelDev sk, const double* __restrict__ xf,
                                              const double* __restrict__ dx, double* __restrict__ out, double alpha) {
  using Dm = SkelDims<Q, P>;
  __shared__ SkelStage<Q, P, NS> st;
  __shared__ double U[NS * Dm::ND], dU[KIND == 1 ? NS * Dm::ND : 1];
  __shared__ double S[Dm::NQ][24], dSl[KIND == 1 ? Dm::NQ : 1][24], R[Dm::NQ][24];
  __shared__ double E[Dm::NQ];
  const int tid = threadIdx.x;
  const int64_t f = sk.order[blockIdx.x];
  skel_geometry<Q, P, NS>(op, sk, f, st, tid, 256);
  for (int it = tid; it < NS * Dm::ND; it += 256) {
    U[it] = skel_value(st.id[it], xf, op.dvals);
    if constexpr (KIND == 1) dU[it] = skel_value(st.id[it], dx, nullptr);
  }
  __syncthreads();
  skel_slots<Q, P, NS>(st, U, S, tid, 256);
  if constexpr (KIND == 1) skel_slots<Q, P, NS>(st, dU, dSl, tid, 256);
  __syncthreads();
  if (tid < Dm::NQ) {
    const int q = tid;
    if constexpr (KIND == 0) {
      double r[12 * NS];
      F::residual(S[q], st.xq[q], st.nrm[q], st.h, sk.p, r);
#pragma unroll
      for (int i = 0; i < 12 * NS; ++i) R[q][i] = st.dS[q] * r[i];
    } else if constexpr (KIND == 1) {
      double r[12 * NS];
      F::jvp(S[q], dSl[q], st.xq[q], st.nrm[q], st.h, sk.p, r);
#pragma unroll
      for (int i = 0; i < 12 * NS; ++i) R[q][i] = st.dS[q] * r[i];
    } else {
      E[q] = st.dS[q] * F::energy(S[q], st.xq[q], st.nrm[q], st.h, sk.p);
    }
  }
  __syncthreads();
  if constexpr (KIND == 2) {
    if (tid == 0) {
      double e = 0.0;
      for (int q = 0; q < Dm::NQ; ++q) e += E[q];
      out[blockIdx.x] = e;
    }
  } else {
    skel_contract<Q, P, NS>(st, R, KIND == 1 ? alpha : 1.0, out, tid, 256);
  }
}

// jacobian!: the facet matrix block of (row side si, column side sj) = (sides & 1, sides >> 1), part
// blockIdx.x % split of its entries (at most 32 per thread), added to the CSC values at the positions of its free
// (row, column) pairs, found by binary search in the column (the reference precomputes them per facet,
// skeleton_positions!, src/skeleton.jl:392-415).  One launch per block and colour: the four blocks of a facet share
// the entries of its face's dofs.  Work-groups per facet: skel_jac_split(Q).
//   K[(si,ci,ti),(sj,cj,tj)] = sum_q sum_{a in slots(ci)} sum_{b in slots(cj)} B_a(ti) dR_(si,a)/ds_(sj,b) B_b(tj)
// with the columns of dR/ds from jvp in the unit directions of the 12 slots of side sj.
constexpr int skel_jac_split(int Q) { return (9 * Q * Q * Q * Q * Q * Q + 256 * 32 - 1) / (256 * 32); }

template <int Q, int P, class F, int NS>
__global__ __launch_bounds__(256) void k_skel_jac(OpDev op, SkelDev sk, const double* __restrict__ xf,
                                                  double* __restrict__ nzval, int sides) {
  using Dm = SkelDims<Q, P>;
  constexpr int TOT = 9 * Dm::NN * Dm::NN, NSPLIT = skel_jac_split(Q), NE = (TOT + 256 * NSPLIT - 1) / (256 * NSPLIT);
  __shared__ SkelStage<Q, P, NS> st;
  __shared__ double U[NS * Dm::ND];
  __shared__ double S[Dm::NQ][24];
  __shared__ double T[Dm::NQ][12][12];  // [q][(ci, a)][(cj, b)] of the (si, sj) block, dS included
  __shared__ double H[12][3 * Dm::NN];  // [(ci, a)][(cj, tj)] at the current point
  const int tid = threadIdx.x;
  const int64_t f = sk.order[blockIdx.x / NSPLIT];
  const int si = sides & 1, sj = (sides >> 1) & 1, idx0 = (int)(blockIdx.x % NSPLIT) * 256 * NE;
  skel_geometry<Q, P, NS>(op, sk, f, st, tid, 256);
  for (int it = tid; it < NS * Dm::ND; it += 256) U[it] = skel_value(st.id[it], xf, op.dvals);
  __syncthreads();
  skel_slots<Q, P, NS>(st, U, S, tid, 256);
  __syncthreads();
  for (int it = tid; it < Dm::NQ * 12; it += 256) {
    const int q = it / 12, bcol = it - 12 * q;
    double ds[12 * NS], r[12 * NS];
#pragma unroll
    for (int i = 0; i < 12 * NS; ++i) ds[i] = 0.0;
    ds[12 * sj + bcol] = 1.0;
    F::jvp(S[q], ds, st.xq[q], st.nrm[q], st.h, sk.p, r);
#pragma unroll
    for (int arow = 0; arow < 12; ++arow) T[q][arow][bcol] = st.dS[q] * r[12 * si + arow];
  }
  double acc[NE];
#pragma unroll
  for (int e = 0; e < NE; ++e) acc[e] = 0.0;
  for (int q = 0; q < Dm::NQ; ++q) {
    __syncthreads();
    // H[(ci, a)][(cj, tj)] = sum_b T[(ci, a)][(cj, b)] B_b(tj) with B = (N, dN/dx_0, dN/dx_1, dN/dx_2) of side sj
    for (int it = tid; it < 12 * 3 * Dm::NN; it += 256) {
      const int arow = it / (3 * Dm::NN), k = it - arow * 3 * Dm::NN, cj = k / Dm::NN, tj = k - cj * Dm::NN;
      const double* Tr = T[q][arow];
      H[arow][k] = Tr[cj] * st.Nv[sj][q][tj] + Tr[3 + 3 * cj] * st.G[sj][q][tj][0] +
                   Tr[4 + 3 * cj] * st.G[sj][q][tj][1] + Tr[5 + 3 * cj] * st.G[sj][q][tj][2];
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < NE; ++e) {
      const int idx = idx0 + tid + 256 * e;
      if (idx < TOT) {
        const int i = idx % (3 * Dm::NN), j = idx / (3 * Dm::NN);  // row (ci, ti), column (cj, tj)
        const int ci = i / Dm::NN, ti = i - ci * Dm::NN;
        acc[e] += st.Nv[si][q][ti] * H[ci][j] + st.G[si][q][ti][0] * H[3 + 3 * ci][j] +
                  st.G[si][q][ti][1] * H[4 + 3 * ci][j] + st.G[si][q][ti][2] * H[5 + 3 * ci][j];
      }
    }
  }
#pragma unroll
  for (int e = 0; e < NE; ++e) {
    const int idx = idx0 + tid + 256 * e;
    if (idx < TOT) {
      const int i = idx % (3 * Dm::NN), j = idx / (3 * Dm::NN);
      const int32_t r = st.id[si * Dm::ND + i], c = st.id[sj * Dm::ND + j];
      if (r > 0 && c > 0) {
        int64_t lo = sk.colptr[c - 1], hi = sk.colptr[c] - 1;
        const int32_t want = r - 1;
        while (lo < hi) {  // first position with rowval >= want
          const int64_t mid = (lo + hi) >> 1;
          if (sk.rowval[mid] < want) lo = mid + 1;
          else hi = mid;
        }
        if (lo < sk.colptr[c] && sk.rowval[lo] == want) nzval[lo] += acc[e];  // race-free within a colour
      }
    }
  }
}

// The built-in facet form: a penalty on the jump of the displacement gradient (the continuous interior penalty /
// ghost penalty of C0 elements), energy (gamma h / 2) |grad u+ - grad u-|^2 with gamma = p[0].
struct HfkGradJump {
  __device__ static void residual(const double* s, const double* x, const double* n, double h, const double* p,
                                  double* R) {
    const double g = p[0] * h;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      R[i] = 0.0;
      R[12 + i] = 0.0;
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      const double j = s[3 + k] - s[15 + k];
      R[3 + k] = g * j;
      R[15 + k] = -g * j;
    }
  }
  __device__ static void jvp(const double* s, const double* ds, const double* x, const double* n, double h,
                             const double* p, double* dR) {
    residual(ds, x, n, h, p, dR);
  }
  __device__ static double energy(const double* s, const double* x, const double* n, double h, const double* p) {
    double e = 0.0;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      const double j = s[3 + k] - s[15 + k];
      e += j * j;
    }
    return 0.5 * p[0] * h * e;
  }
};

// The built-in boundary form: a follower pressure p[0] on the deformed face, the nominal traction
// t0 = -p cof(F) N (cof F = J F^-T, F = I + grad u, N the reference normal), r(v) = -int t0 . v dA, so
// R_c = p (cof F N)_c.  The load is not conservative on an open surface: energy() is 0.
struct HfkFollowerPressure {
  // (cof F)_{cK} N_K with cof F_{cK} = (1/2) eps_cab eps_KLM F_aL F_bM
  __device__ static void cofn(const double* s, const double* n, double* out) {
    double F[9];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int K = 0; K < 3; ++K) F[3 * c + K] = (c == K ? 1.0 : 0.0) + s[3 + 3 * c + K];
    // cof F = rows: cross products of the other two rows of F
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const int a = (c + 1) % 3, b = (c + 2) % 3;
      const double r0 = F[3 * a + 1] * F[3 * b + 2] - F[3 * a + 2] * F[3 * b + 1];
      const double r1 = F[3 * a + 2] * F[3 * b + 0] - F[3 * a + 0] * F[3 * b + 2];
      const double r2 = F[3 * a + 0] * F[3 * b + 1] - F[3 * a + 1] * F[3 * b + 0];
      out[c] = r0 * n[0] + r1 * n[1] + r2 * n[2];
    }
  }
  __device__ static void residual(const double* s, const double* x, const double* n, double h, const double* p,
                                  double* R) {
    double cn[3];
    cofn(s, n, cn);
#pragma unroll
    for (int i = 0; i < 12; ++i) R[i] = 0.0;
#pragma unroll
    for (int c = 0; c < 3; ++c) R[c] = p[0] * cn[c];
  }
  __device__ static void jvp(const double* s, const double* ds, const double* x, const double* n, double h,
                             const double* p, double* dR) {
    // d(cof F) N in direction dF: for row c, d(F_a x F_b) = dF_a x F_b + F_a x dF_b
    double F[9], dF[9];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int K = 0; K < 3; ++K) {
        F[3 * c + K] = (c == K ? 1.0 : 0.0) + s[3 + 3 * c + K];
        dF[3 * c + K] = ds[3 + 3 * c + K];
      }
#pragma unroll
    for (int i = 0; i < 12; ++i) dR[i] = 0.0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const int a = (c + 1) % 3, b = (c + 2) % 3;
      const double* Fa = F + 3 * a;
      const double* Fb = F + 3 * b;
      const double* dA = dF + 3 * a;
      const double* dB = dF + 3 * b;
      const double r0 = dA[1] * Fb[2] - dA[2] * Fb[1] + Fa[1] * dB[2] - Fa[2] * dB[1];
      const double r1 = dA[2] * Fb[0] - dA[0] * Fb[2] + Fa[2] * dB[0] - Fa[0] * dB[2];
      const double r2 = dA[0] * Fb[1] - dA[1] * Fb[0] + Fa[0] * dB[1] - Fa[1] * dB[0];
      dR[c] = p[0] * (r0 * n[0] + r1 * n[1] + r2 * n[2]);
    }
  }
  __device__ static double energy(const double* s, const double* x, const double* n, double h, const double* p) {
    return 0.0;
  }
};

}  // namespace hfk
