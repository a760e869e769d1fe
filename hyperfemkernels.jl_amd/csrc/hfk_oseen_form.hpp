// Device code of the Oseen step on tetrahedra shared by libhfk (hfk_oseen.hip) and the point forms compiled at run
// time with hiprtc (this header is embedded in the library as text): the element data, the P2-P1 shape functions,
// and the residual and Jacobian kernels templated on a point form.  Uses nothing beyond the HIP device runtime.
//
// Element: 4 vertices (affine map x = x0 + J xi), P2 velocity nodes 0..3 the vertices and 4..9 the midpoints of
// the edges (0,1) (0,2) (0,3) (1,2) (1,3) (2,3), P1 pressure on the vertices; cell dofs [c*10 + a] (velocity
// component c, node a) then [30 + a] (pressure).
#pragma once
#include "hfk_types.hpp"

namespace hfk {

constexpr int kND = 34;             // dofs per cell
constexpr int kMaxPoints = 32;

struct OseenDev {
  int64_t ncells;
  int np;
  double rho, mu, c0;
  const int32_t* ids;     // [cell][34]
  const double* coords;   // [cell][4][3]
  const double* qp;       // [q][4]: xi, eta, zeta, weight
  const double* fields;   // [cell][q][9]: w, un, f
  const double* lap;      // [cell][q][3] or null
  const double* dvals;
  const int32_t* nzpos;   // [cell][ii + 34 jj], 1-based, 0: not stored
  double* eres;           // [cell][34]: element residuals stored by a single launch over every cell, or null
};

// barycentric coordinate k (0..3) of a reference point and its (constant) reference gradient
__device__ __forceinline__ double bary(int k, double xi, double et, double ze) {
  return k == 0 ? 1.0 - xi - et - ze : (k == 1 ? xi : (k == 2 ? et : ze));
}
__device__ __forceinline__ void bary_grad(int k, double g[3]) {
  g[0] = k == 0 ? -1.0 : (k == 1 ? 1.0 : 0.0);
  g[1] = k == 0 ? -1.0 : (k == 2 ? 1.0 : 0.0);
  g[2] = k == 0 ? -1.0 : (k == 3 ? 1.0 : 0.0);
}

// shape function f of the cell at a reference point: f = 0..3 P2 vertex functions l(2l - 1), 4..9 P2 edge
// functions 4 l_a l_b (edges (0,1) (0,2) (0,3) (1,2) (1,3) (2,3)), 10..13 the P1 functions l; value and
// reference gradient
__device__ __forceinline__ void shape(int f, double xi, double et, double ze, double& n, double g[3]) {
  if (f < 4 || f >= 10) {
    const int k = f < 4 ? f : f - 10;
    const double l = bary(k, xi, et, ze);
    double gl[3];
    bary_grad(k, gl);
    const double s = f < 4 ? 4.0 * l - 1.0 : 1.0;
    n = f < 4 ? l * (2.0 * l - 1.0) : l;
#pragma unroll
    for (int d = 0; d < 3; ++d) g[d] = s * gl[d];
  } else {
    const int e = f - 4;
    const int a = e < 3 ? 0 : (e < 5 ? 1 : 2);
    const int b = e < 3 ? e + 1 : (e < 5 ? e - 1 : 3);
    const double la = bary(a, xi, et, ze), lb = bary(b, xi, et, ze);
    double ga[3], gb[3];
    bary_grad(a, ga);
    bary_grad(b, gb);
    n = 4.0 * la * lb;
#pragma unroll
    for (int d = 0; d < 3; ++d) g[d] = 4.0 * (lb * ga[d] + la * gb[d]);
  }
}

// Cell of a work-group: the cells of one colour (cells != null), or, in the single-launch mode (cells == null),
// every cell with XCD x (work-groups b = x mod 8) taking the contiguous block [x chunk, (x + 1) chunk) in mesh
// order, so that neighbouring cells, which add into the same CSC lines, meet in one L2.  -1: idle.
__device__ __forceinline__ int64_t work_item(const int32_t* cells, int64_t n, int64_t chunk) {
  if (cells) return cells[blockIdx.x];
  const int64_t c = (int64_t)(blockIdx.x & 7) * chunk + (blockIdx.x >> 3);
  return c < n ? c : -1;
}

// inverse transpose of the affine map of a cell (Jit[i + 3 j] = (J^-1)_ji, so that grad_x = Jit grad_xi), |det J|
// and the cell size h = (6 sqrt(2) |K|)^(1/3) of test/common.jl:201-205
__device__ __forceinline__ void tet_geometry(const double* X, double* Jit, double& adet, double& h) {
  double J[9];  // column d: x_{d+1} - x_0
#pragma unroll
  for (int d = 0; d < 3; ++d)
#pragma unroll
    for (int i = 0; i < 3; ++i) J[i + 3 * d] = X[3 * (d + 1) + i] - X[i];
  const double det = J[0] * (J[4] * J[8] - J[7] * J[5]) - J[3] * (J[1] * J[8] - J[7] * J[2]) +
                     J[6] * (J[1] * J[5] - J[4] * J[2]);
  const double id = 1.0 / det;
  double I[9];
  I[0] = (J[4] * J[8] - J[7] * J[5]) * id;
  I[3] = -(J[3] * J[8] - J[6] * J[5]) * id;
  I[6] = (J[3] * J[7] - J[6] * J[4]) * id;
  I[1] = -(J[1] * J[8] - J[7] * J[2]) * id;
  I[4] = (J[0] * J[8] - J[6] * J[2]) * id;
  I[7] = -(J[0] * J[7] - J[6] * J[1]) * id;
  I[2] = (J[1] * J[5] - J[4] * J[2]) * id;
  I[5] = -(J[0] * J[5] - J[3] * J[2]) * id;
  I[8] = (J[0] * J[4] - J[3] * J[1]) * id;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) Jit[i + 3 * j] = I[j + 3 * i];
  adet = fabs(det);
  h = cbrt(6.0 * sqrt(2.0) * adet / 6.0);
}

// ---------------------------------------------------------------------------------------------------------------
// Point forms.  A velocity-pressure integrand a((u,p),(v,q)) - l((v,q)) given per quadrature point by a struct
//   static void residual(const double* s, const double* d, double h, const double* prm, double* co);
//   static void tangent(const double* d, double h, const double* prm, double* k);
// (generated from the captured integrand, the role of PointForm, src/codegen.jl:166-198).  Slots, test and trial
// alike: 0..2 the velocity components, 3 + 3i + J the derivative d/dx_J of component i, 12 the pressure, 13 + J its
// derivative.  Data d: w (3), un (3), f (3), lap (3, zero when absent); h the cell size; prm: rho, mu, c0, then the
// form's own parameters.  `residual` gives the coefficient of every test slot at the state slots s; `tangent` gives
// the 50 factors of the tangent T[test slot][trial slot] = d^2 a / d v_s d u_r of a form in the factorised class
//   T[v_i, u_j] = d_ij m                 T[v_i, du_j/dx_L] = d_ij cu_L       T[dv_i/dx_J, u_j] = d_ij cv_J
//   T[dv_i/dx_J, du_j/dx_L] = d_ij K_JL + kt d_iL d_Jj + kd d_iJ d_jL
//   T[v_i, p] = zv_i   T[v_i, dp/dx_L] = d_iL zg   T[dv_i/dx_J, p] = d_iJ zd   T[dv_i/dx_J, dp/dx_L] = d_iL zw_J
//   T[q, u_j] = yv_j   T[q, du_j/dx_L] = d_jL yd   T[dq/dx_J, u_j] = d_Jj yg   T[dq/dx_J, du_j/dx_L] = d_Jj yw_L
//   T[q, p] = x0       T[q, dp/dx_L] = xg_L       T[dq/dx_J, p] = xh_J       T[dq/dx_J, dp/dx_L] = X_JL
// stored in k in that order: m, cu (3), cv (3), K (9, [J + 3L]), kt, kd, zv (3), zg, zd, zw (3), yv (3), yd, yg,
// yw (3), x0, xg (3), xh (3), X (9).  The element matrix then takes the same pair sums as the built-in kernel, with
// the groups of factors that the form declares zero left out.
constexpr int kFormCoefs = 50;
constexpr int kFormParams = 8;
struct FormParams {
  double v[kFormParams];
};
constexpr int kShV = 11;  // per (point, function): N, grad N (3), cu.g, cv.g, K g (3), zw.g, yw.g

// A form also states which factors can be nonzero: bit i of `static constexpr unsigned long long kNonzero` (a
// generated form clears the bits of factors that are the constant 0), so that the kernel drops those terms.
template <class Form>
__host__ __device__ constexpr bool any_nz(int first, int last) {
  return ((Form::kNonzero >> first) & ((2ull << (last - first)) - 1ull)) != 0ull;
}

// Dynamic shared memory of k_oseen_form_jac in doubles (np points)
__host__ __device__ constexpr int form_jac_shared(int np) { return np * 14 * kShV + np * (kFormCoefs + 1); }

// jacobian! of a point form: one work-group per cell of one colour (or every cell in the single-launch mode),
// the point factors of the form staged in LDS, then the pair sums of the factorised element matrix, the free
// entries summed into the CSC values through nzpos (as k_oseen does for the built-in form).
template <class Form, bool ATOMIC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_oseen_form_jac(
    OseenDev op, const int32_t* __restrict__ cells, int64_t chunk, double* __restrict__ nzval, FormParams prm) {
  constexpr bool M0 = any_nz<Form>(0, 0), CU = any_nz<Form>(1, 3), CV = any_nz<Form>(4, 6), KK = any_nz<Form>(7, 15);
  constexpr bool KT = any_nz<Form>(16, 16), KD = any_nz<Form>(17, 17);
  constexpr bool ZV = any_nz<Form>(18, 20), ZG = any_nz<Form>(21, 21), ZD = any_nz<Form>(22, 22), ZW = any_nz<Form>(23, 25);
  constexpr bool YV = any_nz<Form>(26, 28), YD = any_nz<Form>(29, 29), YG = any_nz<Form>(30, 30), YW = any_nz<Form>(31, 33);
  constexpr bool X0 = any_nz<Form>(34, 34), XG = any_nz<Form>(35, 37), XH = any_nz<Form>(38, 40), XX = any_nz<Form>(41, 49);
  extern __shared__ double dyn[];
  __shared__ double Ke[kND][kND + 1];
  const int tid = threadIdx.x;
  const int64_t cell = work_item(cells, op.ncells, chunk);
  if (cell < 0) return;  // uniform over the work-group, before any barrier
  const int np = op.np;
  double* Sh = dyn;                   // [q][f][kShV]
  double* Pk = dyn + np * 14 * kShV;  // [q][kFormCoefs + 1]: the form's factors, then dV
  // the CSC positions of this thread's entries and the point data of its point are requested first, so that
  // their latencies overlap; the affine map is then computed by every thread (no serial step, no barrier)
  constexpr int kPerThread = (kND * kND + 255) / 256;
  int32_t pk[kPerThread];
  double old[kPerThread];
  {
    const int32_t* pos = op.nzpos + cell * kND * kND;
#pragma unroll
    for (int k = 0; k < kPerThread; ++k) {
      const int e = tid + 256 * k;
      pk[k] = e < kND * kND ? pos[e] : 0;
    }
  }
  double d[12];
  if (tid < np) {
    const double* fq = op.fields + ((size_t)cell * np + tid) * 9;
#pragma unroll
    for (int i = 0; i < 9; ++i) d[i] = fq[i];
#pragma unroll
    for (int i = 0; i < 3; ++i) d[9 + i] = op.lap ? op.lap[((size_t)cell * np + tid) * 3 + i] : 0.0;
  }
  double Jit[9], adet, hc;
  tet_geometry(op.coords + cell * 12, Jit, adet, hc);
  if (tid < np) {
    double* k = Pk + tid * (kFormCoefs + 1);
    Form::tangent(d, hc, prm.v, k);
    k[kFormCoefs] = adet * op.qp[4 * tid + 3];
  }
  __syncthreads();
  for (int it = tid; it < np * 14; it += blockDim.x) {
    const int q = it / 14, f = it - q * 14;
    double n, g[3], gx[3];
    shape(f, op.qp[4 * q], op.qp[4 * q + 1], op.qp[4 * q + 2], n, g);
#pragma unroll
    for (int i = 0; i < 3; ++i) gx[i] = Jit[i] * g[0] + Jit[i + 3] * g[1] + Jit[i + 6] * g[2];
    double* s = Sh + (q * 14 + f) * kShV;
    s[0] = n;
    s[1] = gx[0];
    s[2] = gx[1];
    s[3] = gx[2];
    if (f < 10) {
      const double* k = Pk + q * (kFormCoefs + 1);
      if constexpr (CU) s[4] = k[1] * gx[0] + k[2] * gx[1] + k[3] * gx[2];
      if constexpr (CV) s[5] = k[4] * gx[0] + k[5] * gx[1] + k[6] * gx[2];
      if constexpr (KK) {
#pragma unroll
        for (int J = 0; J < 3; ++J) s[6 + J] = k[7 + J] * gx[0] + k[10 + J] * gx[1] + k[13 + J] * gx[2];
      }
      if constexpr (ZW) s[9] = k[23] * gx[0] + k[24] * gx[1] + k[25] * gx[2];
      if constexpr (YW) s[10] = k[31] * gx[0] + k[32] * gx[1] + k[33] * gx[2];
    }
  }
  __syncthreads();
  if (!ATOMIC) {
#pragma unroll
    for (int k = 0; k < kPerThread; ++k) old[k] = pk[k] > 0 ? nzval[pk[k] - 1] : 0.0;
  }
  if (tid < 100) {  // velocity-velocity, nodes (a, b)
    const int a = tid / 10, bn = tid - a * 10;
    double S = 0.0, P[9], Q[9];
#pragma unroll
    for (int e = 0; e < 9; ++e) P[e] = Q[e] = 0.0;
    for (int q = 0; q < np; ++q) {
      const double* k = Pk + q * (kFormCoefs + 1);
      const double* sa = Sh + (q * 14 + a) * kShV;
      const double* sb = Sh + (q * 14 + bn) * kShV;
      const double dV = k[kFormCoefs];
      double sv = 0.0;
      if constexpr (M0) sv += k[0] * sa[0] * sb[0];
      if constexpr (CU) sv += sa[0] * sb[4];
      if constexpr (CV) sv += sa[5] * sb[0];
      if constexpr (KK) sv += sa[1] * sb[6] + sa[2] * sb[7] + sa[3] * sb[8];
      S += dV * sv;
      const double wt = dV * k[16], wd = dV * k[17];
#pragma unroll
      for (int u = 0; u < 3; ++u)
#pragma unroll
        for (int v = 0; v < 3; ++v) {
          if constexpr (KT) P[u * 3 + v] += wt * sa[1 + u] * sb[1 + v];
          if constexpr (KD) Q[u * 3 + v] += wd * sa[1 + u] * sb[1 + v];
        }
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) Ke[i * 10 + a][j * 10 + bn] = (i == j ? S : 0.0) + P[j * 3 + i] + Q[i * 3 + j];
  } else if (tid < 140) {  // velocity test node a, pressure trial vertex b
    const int t = tid - 100, a = t / 4, b = t - (t / 4) * 4;
    double V[3] = {0.0, 0.0, 0.0};
    for (int q = 0; q < np; ++q) {
      const double* k = Pk + q * (kFormCoefs + 1);
      const double* sa = Sh + (q * 14 + a) * kShV;
      const double* sm = Sh + (q * 14 + 10 + b) * kShV;
      const double dV = k[kFormCoefs];
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        double v = 0.0;
        if constexpr (ZV) v += k[18 + i] * sa[0] * sm[0];
        if constexpr (ZG) v += k[21] * sa[0] * sm[1 + i];
        if constexpr (ZD) v += k[22] * sa[1 + i] * sm[0];
        if constexpr (ZW) v += sa[9] * sm[1 + i];
        V[i] += dV * v;
      }
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) Ke[i * 10 + a][30 + b] = V[i];
  } else if (tid < 180) {  // pressure test vertex a, velocity trial node b
    const int t = tid - 140, a = t / 10, bn = t - (t / 10) * 10;
    double W[3] = {0.0, 0.0, 0.0};
    for (int q = 0; q < np; ++q) {
      const double* k = Pk + q * (kFormCoefs + 1);
      const double* sm = Sh + (q * 14 + 10 + a) * kShV;
      const double* sb = Sh + (q * 14 + bn) * kShV;
      const double dV = k[kFormCoefs];
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        double w = 0.0;
        if constexpr (YV) w += k[26 + j] * sm[0] * sb[0];
        if constexpr (YD) w += k[29] * sm[0] * sb[1 + j];
        if constexpr (YG) w += k[30] * sm[1 + j] * sb[0];
        if constexpr (YW) w += sm[1 + j] * sb[10];
        W[j] += dV * w;
      }
    }
#pragma unroll
    for (int j = 0; j < 3; ++j) Ke[30 + a][j * 10 + bn] = W[j];
  } else if (tid < 196) {  // pressure-pressure
    const int t = tid - 180, a = t / 4, b = t - (t / 4) * 4;
    double Z = 0.0;
    for (int q = 0; q < np; ++q) {
      const double* k = Pk + q * (kFormCoefs + 1);
      const double* sa = Sh + (q * 14 + 10 + a) * kShV;
      const double* sb = Sh + (q * 14 + 10 + b) * kShV;
      double z = 0.0;
      if constexpr (X0) z += k[34] * sa[0] * sb[0];
#pragma unroll
      for (int L = 0; L < 3; ++L) {
        if constexpr (XG) z += k[35 + L] * sa[0] * sb[1 + L];
        if constexpr (XH) z += k[38 + L] * sa[1 + L] * sb[0];
      }
      if constexpr (XX) {
#pragma unroll
        for (int J = 0; J < 3; ++J)
#pragma unroll
          for (int L = 0; L < 3; ++L) z += k[41 + J + 3 * L] * sa[1 + J] * sb[1 + L];
      }
      Z += k[kFormCoefs] * z;
    }
    Ke[30 + a][30 + b] = Z;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kPerThread; ++k) {
    const int e = tid + 256 * k;
    if (pk[k] <= 0) continue;
    const int jj = e / kND, ii = e - jj * kND;
    if (ATOMIC) unsafeAtomicAdd(nzval + pk[k] - 1, Ke[ii][jj]);
    else nzval[pk[k] - 1] = old[k] + Ke[ii][jj];
  }
}

// residual! of a point form: as k_oseen_res (256 / np cells of one colour per work-group, one thread per (cell,
// point), then one per (cell, row)), with the coefficients of the test slots given by Form::residual.
template <class Form, bool ATOMIC>
__global__ __launch_bounds__(256) void k_oseen_form_res(OseenDev op, const int32_t* __restrict__ cells,
                                                        int64_t ncells_launch, const double* __restrict__ x,
                                                        double* __restrict__ b, FormParams prm) {
  constexpr int kMaxCells = 32;
  __shared__ double Co[256][17];
  __shared__ double Nr[kMaxPoints][14][4];
  __shared__ double JitS[kMaxCells][9];
  __shared__ double Xs[kMaxCells][kND];
  __shared__ double Hs[kMaxCells][2];
  __shared__ int32_t Is[kMaxCells][kND];
  const int tid = threadIdx.x;
  const int np = op.np;
  const int cpw = 256 / np < kMaxCells ? 256 / np : kMaxCells;
  const int64_t c0 = (int64_t)blockIdx.x * cpw;
  const int nc = (int)((ncells_launch - c0) < cpw ? (ncells_launch - c0) : cpw);
  for (int it = tid; it < np * 14; it += 256) {
    const int q = it / 14, f = it - q * 14;
    double n, g[3];
    shape(f, op.qp[4 * q], op.qp[4 * q + 1], op.qp[4 * q + 2], n, g);
    Nr[q][f][0] = n;
    Nr[q][f][1] = g[0];
    Nr[q][f][2] = g[1];
    Nr[q][f][3] = g[2];
  }
  for (int it = tid; it < nc * kND; it += 256) {
    const int lc = it / kND, dd = it - lc * kND;
    const int64_t cell = ATOMIC ? c0 + lc : cells[c0 + lc];
    const int32_t id = op.ids[cell * kND + dd];
    Is[lc][dd] = id;
    Xs[lc][dd] = id > 0 ? x[id - 1] : op.dvals[-id - 1];
  }
  if (tid < nc) {
    const int64_t cell = ATOMIC ? c0 + tid : cells[c0 + tid];
    tet_geometry(op.coords + cell * 12, JitS[tid], Hs[tid][0], Hs[tid][1]);
  }
  __syncthreads();
  if (tid < nc * np) {  // point phase: the state slots, then the form's coefficients of the test slots
    const int lc = tid / np, q = tid - lc * np;
    const int64_t cell = ATOMIC ? c0 + lc : cells[c0 + lc];
    const double* Jt = JitS[lc];
    const double* xe = Xs[lc];
    double s[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) s[i] = 0.0;
    for (int f = 0; f < 14; ++f) {
      const double n = Nr[q][f][0];
      double g[3];
#pragma unroll
      for (int J = 0; J < 3; ++J) g[J] = Jt[J] * Nr[q][f][1] + Jt[J + 3] * Nr[q][f][2] + Jt[J + 6] * Nr[q][f][3];
      if (f < 10) {
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          const double v = xe[i * 10 + f];
          s[i] += n * v;
#pragma unroll
          for (int J = 0; J < 3; ++J) s[3 + 3 * i + J] += g[J] * v;
        }
      } else {
        const double v = xe[30 + f - 10];
        s[12] += n * v;
#pragma unroll
        for (int J = 0; J < 3; ++J) s[13 + J] += g[J] * v;
      }
    }
    const double* fq = op.fields + ((size_t)cell * np + q) * 9;
    double d[12];
#pragma unroll
    for (int i = 0; i < 9; ++i) d[i] = fq[i];
#pragma unroll
    for (int i = 0; i < 3; ++i) d[9 + i] = op.lap ? op.lap[((size_t)cell * np + q) * 3 + i] : 0.0;
    double co[16];
    Form::residual(s, d, Hs[lc][1], prm.v, co);
    const double dV = Hs[lc][0] * op.qp[4 * q + 3];
#pragma unroll
    for (int i = 0; i < 16; ++i) Co[tid][i] = dV * co[i];
  }
  __syncthreads();
  for (int it = tid; it < nc * kND; it += 256) {  // row phase
    const int lc = it / kND, r = it - lc * kND;
    const int32_t id = Is[lc][r];
    if (id <= 0) continue;
    const double* Jt = JitS[lc];
    const int f = r < 30 ? r % 10 : 10 + (r - 30);
    const int i = r < 30 ? r / 10 : 0;
    double acc = 0.0;
    for (int q = 0; q < np; ++q) {
      const double* co = Co[lc * np + q];
      double g[3];
#pragma unroll
      for (int J = 0; J < 3; ++J) g[J] = Jt[J] * Nr[q][f][1] + Jt[J + 3] * Nr[q][f][2] + Jt[J + 6] * Nr[q][f][3];
      if (r < 30) acc += co[i] * Nr[q][f][0] + co[3 + 3 * i] * g[0] + co[4 + 3 * i] * g[1] + co[5 + 3 * i] * g[2];
      else acc += co[12] * Nr[q][f][0] + co[13] * g[0] + co[14] * g[1] + co[15] * g[2];
    }
    if (ATOMIC && op.eres) op.eres[(c0 + lc) * kND + r] = acc;  // summed per dof afterwards, in a fixed order
    else if (ATOMIC) unsafeAtomicAdd(b + id - 1, acc);
    else b[id - 1] += acc;
  }
}

}  // namespace hfk
