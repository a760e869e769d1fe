// One Oseen-BDF2 step of incompressible flow on tetrahedra, Taylor-Hood P2-P1 with SUPG/PSPG/LSIC stabilisation:
// coloured FP64 assembly of the residual and of the CSC Jacobian (the weak form of test/common.jl:165-228, volume
// terms; the reference assembles it through residual!/jacobian!, src/assembly.jl:909-927, and checks it against
// Gridap in test/test_cfd.jl).
//
//   a((u,p),(v,q)) = rho c u.v + 2 mu eps(u):eps(v) + rho ((grad u)' w).v - p div v - q div u
//                  + ts ((grad v)' w).Ru + ts ((grad v)' w).grad p - tp grad q.Ru - tp grad q.grad p
//                  + tl div v div u,                     Ru = rho c u + rho (grad u)' w
//   l((v,q))       = rho f.v + ts ((grad v)' w).rk - tp grad q.rk,        rk = rho f + mu lap
//   ts = ((2|un|/h)^2 + 9 (4 mu/(rho h^2))^2)^(-1/2), tp = ts/rho, tl = rho |un| h / 2, h = (6 sqrt(2) |K|)^(1/3)
// and on the outlet faces (backflow on dGb, Robin and Windkessel pressure on dGo)
//   a += -rho negpart(w.n) u.v + ZA (u.n)(v.n),   l += -Pc v.n,          negpart(s) = (s - |s|) / 2
//
// (test/common.jl:201-220, Gridap's gradient convention grad u[J,c] = d u_c / d x_J).  The residual is
// r(x) = a(x, .) - l(.) on the free test dofs with the Dirichlet values lifted; the Jacobian is a(., .) on the free
// (test, trial) pairs.  w (advecting velocity), un (velocity of the stabilisation parameters), f (BDF history) and
// lap (elementwise Laplacian of un, optional) are point data given per cell and quadrature point; the outlet faces
// come with their points in the reference coordinates of their cell, the weights dS, the outward normal and w at
// the backflow points.
//
// Element: 4 vertices (affine map x = x0 + J xi), P2 velocity nodes 0..3 the vertices and 4..9 the midpoints of
// the edges (0,1) (0,2) (0,3) (1,2) (1,3) (2,3), P1 pressure on the vertices; cell dofs [c*10 + a] (velocity
// component c, node a) then [30 + a] (pressure).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <memory>
#include <vector>

#include "hfk_op.hpp"
#include "hfk_oseen_form.hpp"
#include "hfk_solver.hpp"

namespace hfk {
namespace {

// One work-group per cell, the element matrix in factorised form.  The velocity-velocity block of test (i, a)
// and trial (j, b) is  delta_ij S_ab + P_ab[j][i] + Q_ab[i][j]  with the point sums
//   S_ab = sum dV (rho c Na Nb + mu ga.gb + rho Na (w.gb) + ts rho (w.ga)(c Nb + w.gb)),
//   P_ab = sum dV mu ga (x) gb   (the second half of 2 mu eps:eps),   Q_ab = sum dV tl ga (x) gb   (LSIC),
// so one thread per node pair (a, b) accumulates 19 numbers instead of 9 entries of the full form; the
// pressure-velocity, velocity-pressure and pressure-pressure blocks take one thread per node pair as well (3, 3
// and 1 numbers).  All points of the cell are staged in LDS first (one barrier), and the entries go to the CSC
// values through nzpos.  The residual does not form the element matrix (k_oseen_res).
constexpr int kVV = 100, kVP = 40, kPV = 40, kPP = 16, kPairs = kVV + kVP + kPV + kPP;  // 196 pair threads

template <bool ATOMIC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_oseen(OseenDev op, const int32_t* __restrict__ cells, int64_t chunk,
                                               double* __restrict__ nzval) {
  __shared__ double Sh[kMaxPoints][14][5];  // per point and shape function: N, grad N (3), w . grad N
  __shared__ double Pq[kMaxPoints][4];      // per point: dV, ts, tp, tl
  __shared__ double Ke[kND][kND + 1];       // element matrix [test][trial]
  const int tid = threadIdx.x;
  const int64_t cell = work_item(cells, op.ncells, chunk);
  if (cell < 0) return;  // uniform over the work-group, before any barrier
  const int np = op.np;
  // Everything this thread reads from memory before the pair phase is requested first, so that the latencies
  // overlap: the CSC positions of its element-matrix entries e = ii + 34 jj (entry Ke[ii][jj]), the point data of
  // its staging items and the cell's vertices.  The values the entries add to are loaded after the staging (loads
  // complete in issue order, so that their wait does not hold the staging) and arrive while the pair phase, which
  // reads only LDS, runs.  Within one colour launch no other work-group touches these entries.
  constexpr int kPerThread = (kND * kND + 255) / 256;  // the launch has 256 threads per work-group
  constexpr int kStage = (kMaxPoints * 15 + 255) / 256;  // staging items per thread
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
  double fv[kStage][3];  // w at the point (shape-function items) or un (the point-scalar item)
#pragma unroll
  for (int k = 0; k < kStage; ++k) {
    const int it = tid + 256 * k;
    if (it < np * 15) {
      const int q = it / 15, f = it - q * 15;
      const double* fq = op.fields + ((size_t)cell * np + q) * 9 + (f < 14 ? 0 : 3);
#pragma unroll
      for (int i = 0; i < 3; ++i) fv[k][i] = fq[i];
    }
  }
  // the affine map of the cell, in every thread (no barrier): J (column d: x_{d+1} - x_0), det J, the transposed
  // inverse Jit, |det J| and h
  double Jit[9], adet;
  {
    const double* X = op.coords + cell * 12;
    double J[9];
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int i = 0; i < 3; ++i) J[i + 3 * d] = X[3 * (d + 1) + i] - X[i];
    const double det = J[0] * (J[4] * J[8] - J[7] * J[5]) - J[3] * (J[1] * J[8] - J[7] * J[2]) +
                       J[6] * (J[1] * J[5] - J[4] * J[2]);
    const double id = 1.0 / det;
    double I[9];  // inverse (column-major)
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
  }
  // stage every (point, shape function) and the point scalars
#pragma unroll
  for (int k = 0; k < kStage; ++k) {
    const int it = tid + 256 * k;
    if (it >= np * 15) break;
    const int q = it / 15, f = it - q * 15;
    if (f < 14) {
      double n, g[3];
      shape(f, op.qp[4 * q], op.qp[4 * q + 1], op.qp[4 * q + 2], n, g);
      double gx[3], wg = 0.0;
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        gx[i] = Jit[i] * g[0] + Jit[i + 3] * g[1] + Jit[i + 6] * g[2];
        wg += fv[k][i] * gx[i];
      }
      Sh[q][f][0] = n;
      Sh[q][f][1] = gx[0];
      Sh[q][f][2] = gx[1];
      Sh[q][f][3] = gx[2];
      Sh[q][f][4] = wg;
    } else {
      const double un = sqrt(fv[k][0] * fv[k][0] + fv[k][1] * fv[k][1] + fv[k][2] * fv[k][2]);
      const double h = cbrt(6.0 * sqrt(2.0) * adet / 6.0);
      const double r1 = 2.0 * un / h, r2 = 3.0 * (4.0 * op.mu / (op.rho * h * h));
      const double ts = 1.0 / sqrt(r1 * r1 + r2 * r2);
      Pq[q][0] = adet * op.qp[4 * q + 3];
      Pq[q][1] = ts;
      Pq[q][2] = ts / op.rho;
      Pq[q][3] = op.rho * un * h * 0.5;
    }
  }
  __syncthreads();
  if (!ATOMIC) {
#pragma unroll
    for (int k = 0; k < kPerThread; ++k) old[k] = pk[k] > 0 ? nzval[pk[k] - 1] : 0.0;
  }
  const double rho = op.rho, mu = op.mu, c = op.c0;
  if (tid < kVV) {  // velocity-velocity, nodes (a, b)
    const int a = tid / 10, bn = tid - a * 10;
    double S = 0.0, P[9], Q[9];
#pragma unroll
    for (int e = 0; e < 9; ++e) P[e] = Q[e] = 0.0;
    for (int q = 0; q < np; ++q) {
      const double dV = Pq[q][0], ts = Pq[q][1], tl = Pq[q][3];
      const double Na = Sh[q][a][0], ga[3] = {Sh[q][a][1], Sh[q][a][2], Sh[q][a][3]}, wga = Sh[q][a][4];
      const double Nb = Sh[q][bn][0], gb[3] = {Sh[q][bn][1], Sh[q][bn][2], Sh[q][bn][3]}, wgb = Sh[q][bn][4];
      S += dV * (rho * c * Na * Nb + mu * (ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2]) + rho * Na * wgb +
                 ts * rho * wga * (c * Nb + wgb));
      const double wm = dV * mu, wl = dV * tl;
#pragma unroll
      for (int u = 0; u < 3; ++u)
#pragma unroll
        for (int v = 0; v < 3; ++v) {
          P[u * 3 + v] += wm * ga[u] * gb[v];
          Q[u * 3 + v] += wl * ga[u] * gb[v];
        }
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) Ke[i * 10 + a][j * 10 + bn] = (i == j ? S : 0.0) + P[j * 3 + i] + Q[i * 3 + j];
  } else if (tid < kVV + kVP) {  // velocity test node a, pressure trial vertex bp
    const int t = tid - kVV, a = t / 4, bp = 10 + (t - (t / 4) * 4);
    double V[3] = {0.0, 0.0, 0.0};
    for (int q = 0; q < np; ++q) {
      const double dV = Pq[q][0], ts = Pq[q][1];
      const double wga = Sh[q][a][4], Mb = Sh[q][bp][0];
#pragma unroll
      for (int i = 0; i < 3; ++i) V[i] += dV * (-Mb * Sh[q][a][1 + i] + ts * wga * Sh[q][bp][1 + i]);
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) Ke[i * 10 + a][20 + bp] = V[i];
  } else if (tid < kVV + kVP + kPV) {  // pressure test vertex ap, velocity trial node b
    const int t = tid - kVV - kVP, ap = 10 + t / 10, bn = t - (t / 10) * 10;
    double W[3] = {0.0, 0.0, 0.0};
    for (int q = 0; q < np; ++q) {
      const double dV = Pq[q][0], tp = Pq[q][2];
      const double Ma = Sh[q][ap][0], Nb = Sh[q][bn][0], wgb = Sh[q][bn][4];
#pragma unroll
      for (int j = 0; j < 3; ++j) W[j] += dV * (-Ma * Sh[q][bn][1 + j] - tp * rho * Sh[q][ap][1 + j] * (c * Nb + wgb));
    }
#pragma unroll
    for (int j = 0; j < 3; ++j) Ke[20 + ap][j * 10 + bn] = W[j];
  } else if (tid < kPairs) {  // pressure-pressure
    const int t = tid - kVV - kVP - kPV, ap = 10 + t / 4, bp = 10 + (t - (t / 4) * 4);
    double Z = 0.0;
    for (int q = 0; q < np; ++q)
      Z -= Pq[q][0] * Pq[q][2] *
           (Sh[q][ap][1] * Sh[q][bp][1] + Sh[q][ap][2] * Sh[q][bp][2] + Sh[q][ap][3] * Sh[q][bp][3]);
    Ke[20 + ap][20 + bp] = Z;
  }
  __syncthreads();
  {  // the free (test, trial) entries into the CSC values
#pragma unroll
    for (int k = 0; k < kPerThread; ++k) {
      const int e = tid + 256 * k;
      if (pk[k] <= 0) continue;
      const int jj = e / kND, ii = e - jj * kND;
      if (ATOMIC) unsafeAtomicAdd(nzval + pk[k] - 1, Ke[ii][jj]);  // hardware FP64 add in L2
      else nzval[pk[k] - 1] = old[k] + Ke[ii][jj];
    }
  }
}

// residual! without the element matrix.  Per point, from the cell's values x_e (Dirichlet values lifted):
// u, grad u, p, grad p, then the coefficients of the test functions and of their gradients,
//   velocity rows (i, a):  A_i Na + B_iJ dNa/dx_J,   A_i  = dV (rho c u_i + rho conv_i - rho f_i),
//                          B_iJ = dV (2 mu eps_Ji + (tl div u - p) delta_iJ + ts w_J (Ru_i + dp/dx_i - rk_i)),
//   pressure rows a:       C Ma + D_J dMa/dx_J,      C = -dV div u,   D_J = -dV tp (Ru_J + dp/dx_J - rk_J),
// with conv = (grad u)' w, Ru = rho c u + rho conv, rk = rho f + mu lap: the terms of a((u,p),(v,q)) - l((v,q))
// for one test function, summed over the points (about a tenth of the arithmetic of forming K x_e).  A work-group
// takes 256 / np cells of one colour, one thread per (cell, point), then one thread per (cell, row).
constexpr int kResCoef = 16;  // A (3), B (9), C, D (3)

template <bool ATOMIC>
__global__ __launch_bounds__(256) void k_oseen_res(OseenDev op, const int32_t* __restrict__ cells, int64_t ncells_launch,
                                                   const double* __restrict__ x, double* __restrict__ b) {
  constexpr int kMaxCells = 256;  // np >= 1
  __shared__ double Co[256][kResCoef + 1];       // per (cell, point) slot
  __shared__ double Nr[kMaxPoints][14][4];       // reference value and gradient of each shape function
  __shared__ double JitS[kMaxCells / 8][9];      // per cell (np >= 8 in practice; bounded below)
  __shared__ double Xs[kMaxCells / 8][kND];
  __shared__ double Hs[kMaxCells / 8][2];         // |det J|, h
  __shared__ int32_t Is[kMaxCells / 8][kND];
  const int tid = threadIdx.x;
  const int np = op.np;
  const int cpw = 256 / np < kMaxCells / 8 ? 256 / np : kMaxCells / 8;  // cells per work-group
  const int64_t c0 = (int64_t)blockIdx.x * cpw;
  const int nc = (int)((ncells_launch - c0) < cpw ? (ncells_launch - c0) : cpw);
  // reference shape table, cell geometry and values
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
    const int lc = it / kND, d = it - lc * kND;
    const int64_t cell = ATOMIC ? c0 + lc : cells[c0 + lc];
    const int32_t id = op.ids[cell * kND + d];
    Is[lc][d] = id;
    Xs[lc][d] = id > 0 ? x[id - 1] : op.dvals[-id - 1];
  }
  if (tid < nc) {
    const int64_t cell = ATOMIC ? c0 + tid : cells[c0 + tid];
    const double* X = op.coords + cell * 12;
    double J[9];
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
      for (int j = 0; j < 3; ++j) JitS[tid][i + 3 * j] = I[j + 3 * i];
    Hs[tid][0] = fabs(det);
    Hs[tid][1] = cbrt(6.0 * sqrt(2.0) * fabs(det) / 6.0);
  }
  __syncthreads();
  const double rho = op.rho, mu = op.mu, cc = op.c0;
  if (tid < nc * np) {  // point phase
    const int lc = tid / np, q = tid - lc * np;
    const int64_t cell = ATOMIC ? c0 + lc : cells[c0 + lc];
    const double* Jt = JitS[lc];
    const double* xe = Xs[lc];
    double u[3] = {0, 0, 0}, G[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, p = 0, gp[3] = {0, 0, 0};  // G[J + 3 i] = du_i/dx_J
    for (int f = 0; f < 14; ++f) {
      const double n = Nr[q][f][0];
      double g[3];
#pragma unroll
      for (int J = 0; J < 3; ++J) g[J] = Jt[J] * Nr[q][f][1] + Jt[J + 3] * Nr[q][f][2] + Jt[J + 6] * Nr[q][f][3];
      if (f < 10) {
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          const double v = xe[i * 10 + f];
          u[i] += n * v;
#pragma unroll
          for (int J = 0; J < 3; ++J) G[J + 3 * i] += g[J] * v;
        }
      } else {
        const double v = xe[30 + f - 10];
        p += n * v;
#pragma unroll
        for (int J = 0; J < 3; ++J) gp[J] += g[J] * v;
      }
    }
    const double* fq = op.fields + ((size_t)cell * np + q) * 9;
    const double* lq = op.lap ? op.lap + ((size_t)cell * np + q) * 3 : nullptr;
    const double w[3] = {fq[0], fq[1], fq[2]};
    const double unn = sqrt(fq[3] * fq[3] + fq[4] * fq[4] + fq[5] * fq[5]);
    const double h = Hs[lc][1];
    const double r1 = 2.0 * unn / h, r2 = 3.0 * (4.0 * mu / (rho * h * h));
    const double ts = 1.0 / sqrt(r1 * r1 + r2 * r2), tp = ts / rho, tl = rho * unn * h * 0.5;
    const double dV = Hs[lc][0] * op.qp[4 * q + 3];
    const double div = G[0] + G[4] + G[8];
    double* co = Co[tid];
    double ev[3];  // Ru + grad p - rk
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const double conv = G[0 + 3 * i] * w[0] + G[1 + 3 * i] * w[1] + G[2 + 3 * i] * w[2];
      const double ru = rho * cc * u[i] + rho * conv;
      const double rk = rho * fq[6 + i] + (lq ? mu * lq[i] : 0.0);
      co[i] = dV * (ru - rho * fq[6 + i]);
      ev[i] = ru + gp[i] - rk;
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int J = 0; J < 3; ++J) {
        const double eps = 0.5 * (G[J + 3 * i] + G[i + 3 * J]);
        co[3 + 3 * i + J] = dV * (2.0 * mu * eps + (i == J ? tl * div - p : 0.0) + ts * w[J] * ev[i]);
      }
    co[12] = -dV * div;
#pragma unroll
    for (int J = 0; J < 3; ++J) co[13 + J] = -dV * tp * ev[J];
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

// Outlet faces (test/common.jl:212-213, 219).  Per face: its cell, the outward unit normal, the points of the
// Robin/pressure measure dGo [q][4] (cell reference coordinates, dS) and of the backflow measure dGb [q][7]
// (reference coordinates, dS, w).
constexpr int kMaxFacePoints = 32;
struct OseenFaceDev {
  int nqo, nqb;
  double Pc, ZA;
  const int32_t* cell;    // [face]
  const double* normal;   // [face][3]
  const double* xo;       // [face][nqo][4]
  const double* xb;       // [face][nqb][7]
};

// One work-group per face of one face colour (the faces of a launch share no free dof of their cells):
// the velocity block Kf[(i,a)][(j,b)] = delta_ij Mb_ab + ZA n_i n_j Mo_ab with
//   Mb_ab = sum_dGb dS (-rho negpart(w.n)) Na Nb,   Mo_ab = sum_dGo dS Na Nb,
// and the load -Pc n_i sum_dGo dS Na, added to the CSC values through the cell's nzpos and to the residual rows.
// Velocity shape functions vanish on the faces they are not attached to, so all 10 nodes are summed.
template <bool WANT_K, bool WANT_R, bool ATOMIC>
__global__ __launch_bounds__(128) void k_oseen_face(OseenDev op, OseenFaceDev fd, const int32_t* __restrict__ faces,
                                                    int64_t nfaces, int64_t chunk, const double* __restrict__ x,
                                                    double* __restrict__ b, double* __restrict__ nzval) {
  __shared__ double Nb[kMaxFacePoints][10], No[kMaxFacePoints][10], Wb[kMaxFacePoints], Wo[kMaxFacePoints];
  __shared__ double Mb[10][10], Mo[10][10], Lo[10], Xe[30], Nrm[3];
  __shared__ int32_t Id[30];
  const int tid = threadIdx.x;
  const int64_t f = work_item(faces, nfaces, chunk);
  if (f < 0) return;  // uniform over the work-group, before any barrier
  const int64_t cell = fd.cell[f];
  if (tid < 3) Nrm[tid] = fd.normal[f * 3 + tid];
  if (tid < 30) {
    const int32_t id = op.ids[cell * kND + tid];
    Id[tid] = id;
    Xe[tid] = id > 0 ? (x ? x[id - 1] : 0.0) : op.dvals[-id - 1];
  }
  __syncthreads();
  for (int it = tid; it < (fd.nqb + fd.nqo) * 11; it += blockDim.x) {
    const int q = it / 11, a = it - q * 11;
    const bool back = q < fd.nqb;
    const double* P = back ? fd.xb + ((size_t)f * fd.nqb + q) * 7 : fd.xo + ((size_t)f * fd.nqo + (q - fd.nqb)) * 4;
    if (a < 10) {
      double n, g[3];
      shape(a, P[0], P[1], P[2], n, g);
      if (back) Nb[q][a] = n;
      else No[q - fd.nqb][a] = n;
    } else if (back) {
      const double wn = P[4] * Nrm[0] + P[5] * Nrm[1] + P[6] * Nrm[2];
      Wb[q] = -op.rho * 0.5 * (wn - fabs(wn)) * P[3];
    } else {
      Wo[q - fd.nqb] = P[3];
    }
  }
  __syncthreads();
  if (tid < 100) {
    const int a = tid / 10, bn = tid - a * 10;
    double sb = 0.0, so = 0.0;
    for (int q = 0; q < fd.nqb; ++q) sb += Wb[q] * Nb[q][a] * Nb[q][bn];
    for (int q = 0; q < fd.nqo; ++q) so += Wo[q] * No[q][a] * No[q][bn];
    Mb[a][bn] = sb;
    Mo[a][bn] = so;
  } else if (tid < 110) {
    const int a = tid - 100;
    double l = 0.0;
    for (int q = 0; q < fd.nqo; ++q) l += Wo[q] * No[q][a];
    Lo[a] = l;
  }
  __syncthreads();
  if (WANT_K) {
    const int32_t* pos = op.nzpos + cell * kND * kND;
    for (int e = tid; e < 900; e += blockDim.x) {
      const int jj = e / 30, ii = e - jj * 30;
      const int32_t p = pos[ii + kND * jj];
      if (p <= 0) continue;
      const int i = ii / 10, a = ii - i * 10, j = jj / 10, bn = jj - j * 10;
      const double v = (i == j ? Mb[a][bn] : 0.0) + fd.ZA * Nrm[i] * Nrm[j] * Mo[a][bn];
      if (ATOMIC) unsafeAtomicAdd(nzval + p - 1, v);
      else nzval[p - 1] += v;
    }
  }
  if (WANT_R && tid < 30 && Id[tid] > 0) {
    const int i = tid / 10, a = tid - i * 10;
    double r = fd.Pc * Nrm[i] * Lo[a];  // minus the load -Pc n_i int Na
    for (int jj = 0; jj < 30; ++jj) {
      const int j = jj / 10, bn = jj - j * 10;
      r += ((i == j ? Mb[a][bn] : 0.0) + fd.ZA * Nrm[i] * Nrm[j] * Mo[a][bn]) * Xe[jj];
    }
    if (ATOMIC) unsafeAtomicAdd(b + Id[tid] - 1, r);
    else b[Id[tid] - 1] += r;
  }
}

// The FE functions un = u^n and unm1 = u^{n-1} of a step: free values in the operator's numbering and Dirichlet
// values, and the coefficients of the point data made from them.
struct StepFunctions {
  const double* un;
  const double* un_dir;
  const double* unm1;
  const double* unm1_dir;
  double aw, bw, c1, c2;
};

__device__ __forceinline__ double dof_value(int32_t id, const double* fr, const double* dr) {
  return id > 0 ? fr[id - 1] : dr[-id - 1];
}

// The point data of a step from the FE functions of the two previous steps (the FE-function leaves of
// test/common.jl:183-196, which src/fefunctions.jl materialises on the CPU).  A work-group takes cpw consecutive cells
// (one thread per point of them): it stages their 30 nodal velocities of un and unm1, the P2 shape values at the
// points and, with lap != null, each cell's Laplacian of un in LDS, computes w = aw un + bw unm1, un and
// f = c1 un - c2 unm1 per point, and writes fields[cell][q][9] and lap[cell][q][3] (the points of its cells are
// consecutive, so a wave's stores fill one contiguous block).
// The Laplacian is constant on an affine cell: sum_a un_a lap(N_a) with lap(N_a) = 4 |grad l_a|^2 at vertex a and
// 8 grad l_a . grad l_b at the midpoint of edge (a, b), grad l the physical gradients of the barycentric coordinates
// (J^-T of their reference gradients; J^-T = [e2 x e3, e3 x e1, e1 x e2] / det for the edge vectors e_d = x_d - x_0).
constexpr int kFieldCells = 32;  // cells per work-group of k_oseen_fields at most

__global__ __launch_bounds__(256) void k_oseen_fields(OseenDev op, StepFunctions sf, int cpw, double* __restrict__ fields,
                                                      double* __restrict__ lap) {
  __shared__ double ue[kFieldCells][30], ve[kFieldCells][30], Lc[kFieldCells][10], lapc[kFieldCells][3];
  __shared__ double Nq[kMaxPoints][10];
  const int tid = threadIdx.x, np = op.np;
  const int64_t c0 = (int64_t)blockIdx.x * cpw;
  const int ncl = (int)(op.ncells - c0 < cpw ? op.ncells - c0 : cpw);
  for (int it = tid; it < ncl * 30; it += 256) {
    const int cell = it / 30, j = it - cell * 30;
    const int32_t id = op.ids[(c0 + cell) * kND + j];
    ue[cell][j] = dof_value(id, sf.un, sf.un_dir);
    ve[cell][j] = dof_value(id, sf.unm1, sf.unm1_dir);
  }
  for (int it = tid; it < np * 10; it += 256) {
    const int q = it / 10, a = it - q * 10;
    double n, g[3];
    shape(a, op.qp[q * 4], op.qp[q * 4 + 1], op.qp[q * 4 + 2], n, g);
    Nq[q][a] = n;
  }
  if (lap && tid < ncl) {
    const double* X = op.coords + (c0 + tid) * 12;
    double e[3][3];  // e[d] = x_{d+1} - x_0
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int i = 0; i < 3; ++i) e[d][i] = X[3 * (d + 1) + i] - X[i];
    double G[4][3];  // physical gradients of the barycentric coordinates
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double* p = e[(d + 1) % 3];
      const double* q = e[(d + 2) % 3];
      G[d + 1][0] = p[1] * q[2] - p[2] * q[1];
      G[d + 1][1] = p[2] * q[0] - p[0] * q[2];
      G[d + 1][2] = p[0] * q[1] - p[1] * q[0];
    }
    const double det = e[0][0] * G[1][0] + e[0][1] * G[1][1] + e[0][2] * G[1][2];
#pragma unroll
    for (int k = 1; k < 4; ++k)
#pragma unroll
      for (int i = 0; i < 3; ++i) G[k][i] /= det;
#pragma unroll
    for (int i = 0; i < 3; ++i) G[0][i] = -(G[1][i] + G[2][i] + G[3][i]);
#pragma unroll
    for (int a = 0; a < 4; ++a) Lc[tid][a] = 4.0 * (G[a][0] * G[a][0] + G[a][1] * G[a][1] + G[a][2] * G[a][2]);
#pragma unroll
    for (int ed = 0; ed < 6; ++ed) {
      const int a = ed < 3 ? 0 : (ed < 5 ? 1 : 2);
      const int b = ed < 3 ? ed + 1 : (ed < 5 ? ed - 1 : 3);
      Lc[tid][4 + ed] = 8.0 * (G[a][0] * G[b][0] + G[a][1] * G[b][1] + G[a][2] * G[b][2]);
    }
  }
  __syncthreads();
  if (lap && tid < ncl * 3) {
    const int cell = tid / 3, c = tid - cell * 3;
    double s = 0.0;
    for (int a = 0; a < 10; ++a) s += Lc[cell][a] * ue[cell][c * 10 + a];
    lapc[cell][c] = s;
  }
  __syncthreads();
  if (tid < ncl * np) {
    const int cell = tid / np, q = tid - cell * np;
    double u[3] = {0.0, 0.0, 0.0}, v[3] = {0.0, 0.0, 0.0};
    for (int a = 0; a < 10; ++a) {
      const double n = Nq[q][a];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        u[c] += n * ue[cell][c * 10 + a];
        v[c] += n * ve[cell][c * 10 + a];
      }
    }
    double* fo = fields + (c0 * np + tid) * 9;  // the cells' points are consecutive: one contiguous block
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      fo[c] = sf.aw * u[c] + sf.bw * v[c];
      fo[3 + c] = u[c];
      fo[6 + c] = sf.c1 * u[c] - sf.c2 * v[c];
    }
    if (lap) {
#pragma unroll
      for (int c = 0; c < 3; ++c) lap[(c0 * np + tid) * 3 + c] = lapc[cell][c];
    }
  }
}

// w = aw un + bw unm1 at the dGb points of the outlet faces, into xb[face][q][4..6]: one thread per point
__global__ __launch_bounds__(256) void k_oseen_face_w(OseenDev op, StepFunctions sf, int64_t nfaces, int nqb,
                                                      const int32_t* __restrict__ face_cell, double* __restrict__ xb) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= nfaces * nqb) return;
  const int64_t cell = face_cell[t / nqb];
  double* xq = xb + t * 7;
  const int32_t* ids = op.ids + cell * kND;
  double w[3] = {0.0, 0.0, 0.0};
  for (int a = 0; a < 10; ++a) {
    double n, g[3];
    shape(a, xq[0], xq[1], xq[2], n, g);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const int32_t id = ids[c * 10 + a];
      w[c] += n * (sf.aw * dof_value(id, sf.un, sf.un_dir) + sf.bw * dof_value(id, sf.unm1, sf.unm1_dir));
    }
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) xq[4 + c] = w[c];
}

// b[i] = the element residuals of free dof i summed in ascending (cell, row) order: rows of a single residual launch
// over every cell (each stored once, no atomics), gathered so that the sum is deterministic without colour launches
__global__ __launch_bounds__(256) void k_oseen_gather(int64_t nfree, const int64_t* __restrict__ ptr,
                                                      const int32_t* __restrict__ idx,
                                                      const double* __restrict__ eres, double* __restrict__ b) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nfree; i += (int64_t)gridDim.x * blockDim.x) {
    double v = 0.0;
    const int64_t e = ptr[i + 1];
    int64_t k = ptr[i];
    for (; k + 4 <= e; k += 4) {  // four rows in flight at a time, added in order
      const int32_t j0 = idx[k], j1 = idx[k + 1], j2 = idx[k + 2], j3 = idx[k + 3];
      const double e0 = eres[j0], e1 = eres[j1], e2 = eres[j2], e3 = eres[j3];
      v += e0;
      v += e1;
      v += e2;
      v += e3;
    }
    for (; k < e; ++k) v += eres[idx[k]];
    b[i] = v;
  }
}

}  // namespace
}  // namespace hfk

using namespace hfk;

struct hfk_oseen {
  hfk_ctx* ctx = nullptr;
  int64_t ncells = 0, nfree = 0, ndir = 0;
  int np = 0;
  double rho = 1, mu = 0.01, c0 = 0;
  std::vector<int32_t> cell_dofs, color;  // cell_dofs in the library's local order
  std::vector<int32_t> perm;              // perm[i] = the host's local dof of the library's local dof i
  std::vector<uint8_t> block_mask;        // ncells x 4 (velocity/pressure test x trial blocks) or empty: all
  std::vector<int64_t> color_start;  // cells of colour c: cells_by_color[color_start[c] .. color_start[c+1])
  bool pattern_built = false, has_fields = false;
  DevPattern pat;  // the CSC pattern and the cells' positions, on the device only
  DevBuf ids, coords, qp, fields, lap, dvals, cells_d, x_stage, y_stage, nz_stage;
  bool has_lap = false;
  // colour by colour with plain adds (bitwise reproducible), or one launch with FP64 atomic adds
  bool deterministic = true;
  // the deterministic residual: element residuals of one launch ([cell][34]) and, per free dof, the positions of
  // its rows in them (ascending), summed by k_oseen_gather
  DevBuf eres, gat_ptr, gat_idx;
  bool gather_built = false;
  // outlet faces, ordered by their colour (faces coloured among themselves by shared free dofs)
  int64_t nfaces = 0;
  int nqo = 0, nqb = 0;
  double Pc = 0, ZA = 0;
  std::vector<int64_t> face_start;  // faces of launch group g: faces_d[face_start[g] .. face_start[g+1])
  DevBuf face_cell, face_normal, face_xo, face_xb, faces_d;
  DevBuf step_stage;  // host FE-function values of hfk_oseen_step_fields
  // a point form compiled at run time (hfk_oseen_set_form) in place of the built-in integrand, with its own
  // parameters after rho, mu, c0
  std::unique_ptr<OseenFormKernels> form;
  std::vector<double> form_params;
  hipStream_t stream(void* s) const { return s ? static_cast<hipStream_t>(s) : ctx->stream; }
  FormParams form_args() const {
    FormParams p{};
    p.v[0] = rho;
    p.v[1] = mu;
    p.v[2] = c0;
    for (size_t i = 0; i < form_params.size(); ++i) p.v[3 + i] = form_params[i];
    return p;
  }
  void drop_form() {
    if (!form) return;
    if (hipSetDevice(ctx->device) == hipSuccess) (void)hipDeviceSynchronize();  // launches may still be queued
    (void)hipModuleUnload(form->mod);
    form.reset();
  }
  ~hfk_oseen() { drop_form(); }
  OseenDev dev() const {
    OseenDev d;
    d.ncells = ncells;
    d.np = np;
    d.rho = rho;
    d.mu = mu;
    d.c0 = c0;
    d.ids = ids.as<int32_t>();
    d.coords = coords.as<double>();
    d.qp = qp.as<double>();
    d.fields = fields.as<double>();
    d.lap = has_lap ? lap.as<double>() : nullptr;
    d.dvals = dvals.as<double>();
    d.nzpos = pat.nzpos.as<int32_t>();
    d.eres = nullptr;
    return d;
  }
  OseenFaceDev face_dev() const {
    OseenFaceDev d;
    d.nqo = nqo;
    d.nqb = nqb;
    d.Pc = Pc;
    d.ZA = ZA;
    d.cell = face_cell.as<int32_t>();
    d.normal = face_normal.as<double>();
    d.xo = face_xo.as<double>();
    d.xb = face_xb.as<double>();
    return d;
  }
};

namespace {

int oseen_bind(hfk_oseen* op) {
  if (!op || !op->ctx) return set_err(nullptr, HFK_EARG, "HyperFEMKernels: null Oseen operator");
  HIPCHK(op->ctx, hipSetDevice(op->ctx->device));
  return HFK_OK;
}

int oseen_pattern(hfk_oseen* op) {
  if (op->pattern_built) return HFK_OK;
  BlockInfo bi;
  std::vector<int8_t> field_of(kND);
  for (int i = 0; i < kND; ++i) field_of[i] = i < 30 ? 0 : 1;  // velocity, then pressure (the library's local order)
  if (!op->block_mask.empty()) {
    bi.field_of = field_of.data();
    bi.nf = 2;
    bi.mask = op->block_mask.data();
  }
  const int rc = device_pattern(op->ctx, op->ctx->stream, op->cell_dofs.data(), kND, op->ncells, op->nfree, bi, nullptr,
                                op->pat);
  if (rc) return rc;
  op->pattern_built = true;
  return HFK_OK;
}

// colour by colour (or in one launch with atomic adds); x may be null (zero free values) when only K is wanted
// the volume kernels of a compiled point form over `n` cells (cells: the list of one colour, or null for every
// cell in the single-launch mode)
int form_launch(hfk_oseen* op, bool atomic, const int32_t* cells, int64_t n, const double* x, double* b,
                double* nzval, hipStream_t s, double* eres = nullptr) {
  OseenDev d = op->dev();
  d.eres = eres;
  FormParams prm = op->form_args();
  const int a = atomic ? 1 : 0;
  if (nzval) {
    const int64_t chunk = atomic ? (op->ncells + 7) / 8 : 0;
    const unsigned grid = (unsigned)(atomic ? 8 * chunk : n);
    void* args[] = {&d, &cells, (void*)&chunk, &nzval, &prm};
    HIPCHK(op->ctx, hipModuleLaunchKernel(op->form->jac[a], grid, 1, 1, 256, 1, 1,
                                          (unsigned)(form_jac_shared(op->np) * sizeof(double)), s, args, nullptr));
  }
  if (b) {
    const int cpw = std::min(256 / op->np, 32);
    const unsigned grid = (unsigned)((n + cpw - 1) / cpw);
    void* args[] = {&d, &cells, &n, &x, &b, &prm};
    HIPCHK(op->ctx, hipModuleLaunchKernel(op->form->res[a], grid, 1, 1, 256, 1, 1, 0, s, args, nullptr));
  }
  return HFK_OK;
}

// the per-dof lists of k_oseen_gather, built once from the cell dof table
int oseen_build_gather(hfk_oseen* op) {
  const int64_t nf = op->nfree, nc = op->ncells;
  std::vector<int64_t> ptr((size_t)nf + 1, 0);
  for (int64_t k = 0; k < nc * kND; ++k)
    if (op->cell_dofs[k] > 0) ptr[op->cell_dofs[k]]++;
  for (int64_t i = 0; i < nf; ++i) ptr[i + 1] += ptr[i];
  std::vector<int64_t> fill(ptr.begin(), ptr.end() - 1);
  std::vector<int32_t> idx((size_t)ptr[nf]);
  for (int64_t k = 0; k < nc * kND; ++k)
    if (op->cell_dofs[k] > 0) idx[fill[op->cell_dofs[k] - 1]++] = (int32_t)k;
  HIPCHK(op->ctx, op->eres.alloc((size_t)std::max<int64_t>(nc * kND, 1) * sizeof(double)));
  HIPCHK(op->ctx, op->gat_ptr.alloc(ptr.size() * sizeof(int64_t)));
  HIPCHK(op->ctx, hipMemcpy(op->gat_ptr.p, ptr.data(), ptr.size() * sizeof(int64_t), hipMemcpyHostToDevice));
  HIPCHK(op->ctx, op->gat_idx.alloc(std::max<size_t>(idx.size(), 1) * sizeof(int32_t)));
  if (!idx.empty())
    HIPCHK(op->ctx, hipMemcpy(op->gat_idx.p, idx.data(), idx.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  op->gather_built = true;
  return HFK_OK;
}

// the deterministic residual without colour launches: one launch storing every cell's element residual, then one
// thread per free dof summing its rows in ascending (cell, row) order (b overwritten); used when the cell table's
// positions fit 32 bits
int oseen_residual_gather(hfk_oseen* op, const double* x, double* b, hipStream_t s) {
  if (!op->gather_built) {
    const int rc = oseen_build_gather(op);
    if (rc) return rc;
  }
  double* er = op->eres.as<double>();
  if (op->form) {
    const int rc = form_launch(op, true, nullptr, op->ncells, x, b, nullptr, s, er);
    if (rc) return rc;
  } else {
    OseenDev d = op->dev();
    d.eres = er;
    const int cpw = std::min(256 / op->np, 32);
    hipLaunchKernelGGL((k_oseen_res<true>), dim3((unsigned)((op->ncells + cpw - 1) / cpw)), dim3(256), 0, s, d, nullptr,
                       op->ncells, x, b);
  }
  const int64_t nf = op->nfree;
  const unsigned g = (unsigned)std::max<int64_t>(1, std::min<int64_t>((nf + 255) / 256, 8192));
  hipLaunchKernelGGL(k_oseen_gather, dim3(g), dim3(256), 0, s, nf, op->gat_ptr.as<int64_t>(),
                     op->gat_idx.as<int32_t>(), er, b);
  HIPCHK(op->ctx, hipGetLastError());
  return HFK_OK;
}

template <bool A>
int oseen_launch(hfk_oseen* op, const double* x, double* b, double* nzval, hipStream_t s) {
  const OseenDev d = op->dev();
  const dim3 t(256), tf(128);
  const int cpw = std::min(256 / op->np, 32);  // cells per work-group of k_oseen_res
  auto faces_launch = [&](const int32_t* faces, int64_t chunk, unsigned grid) {
    const OseenFaceDev fd = op->face_dev();
    const dim3 g(grid);
    const int64_t nf = op->nfaces;
    if (b && nzval) hipLaunchKernelGGL((k_oseen_face<true, true, A>), g, tf, 0, s, d, fd, faces, nf, chunk, x, b, nzval);
    else if (b) hipLaunchKernelGGL((k_oseen_face<false, true, A>), g, tf, 0, s, d, fd, faces, nf, chunk, x, b, nzval);
    else hipLaunchKernelGGL((k_oseen_face<true, false, A>), g, tf, 0, s, d, fd, faces, nf, chunk, x, b, nzval);
  };
  if (A) {  // one launch each, cells in mesh order
    const int64_t chunk = (op->ncells + 7) / 8;
    if (op->form) {
      const int rc = form_launch(op, true, nullptr, op->ncells, x, b, nzval, s);
      if (rc) return rc;
    } else {
      if (nzval) hipLaunchKernelGGL((k_oseen<true>), dim3((unsigned)(8 * chunk)), t, 0, s, d, nullptr, chunk, nzval);
      if (b)
        hipLaunchKernelGGL((k_oseen_res<true>), dim3((unsigned)((op->ncells + cpw - 1) / cpw)), t, 0, s, d, nullptr,
                           op->ncells, x, b);
    }
    HIPCHK(op->ctx, hipGetLastError());
    if (op->nfaces > 0) {
      const int64_t fchunk = (op->nfaces + 7) / 8;
      faces_launch(nullptr, fchunk, (unsigned)(8 * fchunk));
      HIPCHK(op->ctx, hipGetLastError());
    }
    return HFK_OK;
  }
  const bool gather = b && op->ncells * kND < (int64_t)INT32_MAX;  // the residual by element rows and a gather
  if (gather) {
    const int rc = oseen_residual_gather(op, x, b, s);
    if (rc) return rc;
  }
  double* bc = gather ? nullptr : b;  // what the colour launches still add into
  for (size_t c = 0; c + 1 < op->color_start.size(); ++c) {
    const int64_t n = op->color_start[c + 1] - op->color_start[c];
    if (n <= 0 || (!nzval && !bc)) continue;
    const int32_t* cells = op->cells_d.as<int32_t>() + op->color_start[c];
    if (op->form) {
      const int rc = form_launch(op, false, cells, n, x, bc, nzval, s);
      if (rc) return rc;
      continue;
    }
    if (nzval) hipLaunchKernelGGL((k_oseen<false>), dim3((unsigned)n), t, 0, s, d, cells, 0, nzval);  // a cell each
    if (bc) hipLaunchKernelGGL((k_oseen_res<false>), dim3((unsigned)((n + cpw - 1) / cpw)), t, 0, s, d, cells, n, x, bc);
    HIPCHK(op->ctx, hipGetLastError());
  }
  if (op->nfaces > 0) {
    for (size_t c = 0; c + 1 < op->face_start.size(); ++c) {
      const int64_t n = op->face_start[c + 1] - op->face_start[c];
      if (n <= 0) continue;
      faces_launch(op->faces_d.as<int32_t>() + op->face_start[c], 0, (unsigned)n);  // one work-group per face
      HIPCHK(op->ctx, hipGetLastError());
    }
  }
  return HFK_OK;
}

int oseen_run(hfk_oseen* op, const double* x, double* b, double* nzval, hipStream_t s) {
  if (!op->has_fields) return set_err(op->ctx, HFK_ESTATE, "HyperFEMKernels: Oseen fields (w, un, f) not set");
  if (b) HIPCHK(op->ctx, launch_fill(b, op->nfree, 0.0, s));
  if (nzval) HIPCHK(op->ctx, launch_fill(nzval, op->pat.nnz, 0.0, s));
  return op->deterministic ? oseen_launch<false>(op, x, b, nzval, s) : oseen_launch<true>(op, x, b, nzval, s);
}

// the point data of a step on the device (k_oseen_fields); enables the Laplacian term when with_lap, removes it
// otherwise
int oseen_step_fields(hfk_oseen* op, const StepFunctions& sf, bool with_lap, hipStream_t s) {
  const int64_t ncp = op->ncells * op->np;
  HIPCHK(op->ctx, op->fields.alloc((size_t)ncp * 9 * sizeof(double)));
  if (with_lap) HIPCHK(op->ctx, op->lap.alloc((size_t)ncp * 3 * sizeof(double)));
  const int cpw = std::min(256 / op->np, kFieldCells);
  hipLaunchKernelGGL(k_oseen_fields, dim3((unsigned)((op->ncells + cpw - 1) / cpw)), dim3(256), 0, s, op->dev(), sf, cpw,
                     op->fields.as<double>(), with_lap ? op->lap.as<double>() : nullptr);
  HIPCHK(op->ctx, hipGetLastError());
  const int64_t nfp = op->nfaces * op->nqb;
  if (nfp > 0) {
    hipLaunchKernelGGL(k_oseen_face_w, dim3((unsigned)((nfp + 255) / 256)), dim3(256), 0, s, op->dev(), sf, op->nfaces,
                       op->nqb, op->face_cell.as<int32_t>(), op->face_xb.as<double>());
    HIPCHK(op->ctx, hipGetLastError());
  }
  op->has_fields = true;
  op->has_lap = with_lap;
  return HFK_OK;
}

}  // namespace

extern "C" {

int hfk_oseen_create(hfk_ctx* ctx, int64_t ncells, const int32_t* host_dofs, const int32_t* nodeperm,
                     const int32_t* pnodeperm, const int32_t* ldof, const double* coords, int np,
                     const double* qpoints, const double* qweights, int64_t nfree, int64_t ndir, double rho, double mu,
                     double c0, hfk_oseen** out) {
  if (!ctx || !out || !host_dofs || !coords || !qpoints || !qweights)
    return set_err(ctx, HFK_EARG, "HyperFEMKernels: null argument");
  if (ncells <= 0 || nfree <= 0 || ndir < 0 || np < 1 || np > kMaxPoints)
    return set_err(ctx, HFK_EARG, "HyperFEMKernels: invalid Oseen sizes (cells, dofs, 1..32 points)");
  if (!(rho > 0) || !(mu > 0) || !(c0 >= 0)) return set_err(ctx, HFK_EARG, "HyperFEMKernels: invalid rho, mu or c0");
  // the host's local order (FieldTable ldof and the node permutations, src/tables.jl:171-237, 329-342): the library's
  // local dof c*10 + f (P2 node f of its own order) is the host's ldof[c*10 + nodeperm[f]], its 30 + v (P1 node v)
  // the host's ldof[30 + pnodeperm[v]]; NULL arrays mean the library's own order
  auto is_perm = [](const int32_t* a, int n) {
    std::vector<char> seen((size_t)n, 0);
    for (int i = 0; i < n; ++i) {
      if (a[i] < 0 || a[i] >= n || seen[a[i]]) return false;
      seen[a[i]] = 1;
    }
    return true;
  };
  if ((nodeperm && !is_perm(nodeperm, 10)) || (pnodeperm && !is_perm(pnodeperm, 4)) || (ldof && !is_perm(ldof, kND)))
    return set_err(ctx, HFK_EARG, "HyperFEMKernels: nodeperm (10), pnodeperm (4) and ldof (34) must be permutations");
  std::vector<int32_t> perm(kND);
  for (int c = 0; c < 3; ++c)
    for (int f = 0; f < 10; ++f) {
      const int h = c * 10 + (nodeperm ? nodeperm[f] : f);
      perm[c * 10 + f] = ldof ? ldof[h] : h;
    }
  for (int v = 0; v < 4; ++v) {
    const int h = 30 + (pnodeperm ? pnodeperm[v] : v);
    perm[30 + v] = ldof ? ldof[h] : h;
  }
  std::vector<int32_t> cell_dofs((size_t)ncells * kND);
  for (int64_t cell = 0; cell < ncells; ++cell)
    for (int i = 0; i < kND; ++i) cell_dofs[(size_t)cell * kND + i] = host_dofs[(size_t)cell * kND + perm[i]];
  for (int64_t i = 0; i < ncells * kND; ++i) {
    const int32_t id = cell_dofs[i];
    if (id == 0 || id > nfree || -(int64_t)id > ndir)
      return set_err(ctx, HFK_EDIM, "HyperFEMKernels: cell dof id out of range");
  }
  HIPCHK(ctx, hipSetDevice(ctx->device));
  std::unique_ptr<hfk_oseen> op(new hfk_oseen());
  op->ctx = ctx;
  op->ncells = ncells;
  op->nfree = nfree;
  op->ndir = ndir;
  op->np = np;
  op->rho = rho;
  op->mu = mu;
  op->c0 = c0;
  op->cell_dofs = std::move(cell_dofs);
  op->perm = std::move(perm);
  if (!greedy_colors(op->cell_dofs.data(), kND, ncells, nfree, op->color))
    return set_err(ctx, HFK_ESTATE, "HyperFEMKernels: more than 128 colours needed");
  const int nc = *std::max_element(op->color.begin(), op->color.end());
  op->color_start.assign((size_t)nc + 1, 0);
  for (int32_t c : op->color) op->color_start[c]++;
  for (int c = 0; c < nc; ++c) op->color_start[c + 1] += op->color_start[c];
  std::vector<int64_t> fill(op->color_start.begin(), op->color_start.end() - 1);
  std::vector<int32_t> cells((size_t)ncells);
  for (int64_t cell = 0; cell < ncells; ++cell) cells[fill[op->color[cell] - 1]++] = (int32_t)cell;
  std::vector<double> q((size_t)np * 4);
  for (int i = 0; i < np; ++i) {
    for (int d = 0; d < 3; ++d) q[4 * i + d] = qpoints[3 * i + d];
    q[4 * i + 3] = qweights[i];
  }
  HIPCHK(ctx, op->ids.alloc((size_t)ncells * kND * sizeof(int32_t)));
  HIPCHK(ctx, hipMemcpy(op->ids.p, op->cell_dofs.data(), (size_t)ncells * kND * sizeof(int32_t), hipMemcpyHostToDevice));
  HIPCHK(ctx, op->coords.alloc((size_t)ncells * 12 * sizeof(double)));
  HIPCHK(ctx, hipMemcpy(op->coords.p, coords, (size_t)ncells * 12 * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(ctx, op->qp.alloc(q.size() * sizeof(double)));
  HIPCHK(ctx, hipMemcpy(op->qp.p, q.data(), q.size() * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(ctx, op->cells_d.alloc(cells.size() * sizeof(int32_t)));
  HIPCHK(ctx, hipMemcpy(op->cells_d.p, cells.data(), cells.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  HIPCHK(ctx, op->dvals.alloc((size_t)std::max<int64_t>(ndir, 1) * sizeof(double)));
  HIPCHK(ctx, hipMemset(op->dvals.p, 0, (size_t)std::max<int64_t>(ndir, 1) * sizeof(double)));
  *out = op.release();
  return HFK_OK;
}

int hfk_oseen_destroy(hfk_oseen* op) {
  delete op;
  return HFK_OK;
}

int hfk_oseen_set_fields(hfk_oseen* op, const double* fields) {
  int rc = oseen_bind(op);
  if (rc) return rc;
  if (!fields) return set_err(op->ctx, HFK_EARG, "HyperFEMKernels: null fields");
  const size_t n = (size_t)op->ncells * op->np * 9;
  HIPCHK(op->ctx, op->fields.alloc(n * sizeof(double)));
  HIPCHK(op->ctx, hipMemcpy(op->fields.p, fields, n * sizeof(double), hipMemcpyHostToDevice));
  op->has_fields = true;
  return HFK_OK;
}

int hfk_oseen_step_fields_dev(hfk_oseen* op, const double* un, const double* un_dir, const double* unm1,
                              const double* unm1_dir, double aw, double bw, double c1, double c2, int laplacian,
                              void* stream) {
  int rc = oseen_bind(op);
  if (rc) return rc;
  if (!un || !unm1 || (op->ndir > 0 && (!un_dir || !unm1_dir)))
    return set_err(op->ctx, HFK_EARG, "HyperFEMKernels: null FE-function values");
  const StepFunctions sf{un, un_dir, unm1, unm1_dir, aw, bw, c1, c2};
  return oseen_step_fields(op, sf, laplacian != 0, op->stream(stream));
}

int hfk_oseen_step_fields(hfk_oseen* op, const double* un, const double* un_dir, const double* unm1,
                          const double* unm1_dir, double aw, double bw, double c1, double c2, int laplacian) {
  int rc = oseen_bind(op);
  if (rc) return rc;
  if (!un || !unm1 || (op->ndir > 0 && (!un_dir || !unm1_dir)))
    return set_err(op->ctx, HFK_EARG, "HyperFEMKernels: null FE-function values");
  const size_t nf = (size_t)op->nfree, nd = (size_t)op->ndir;
  HIPCHK(op->ctx, op->step_stage.alloc((2 * (nf + nd)) * sizeof(double)));
  double* d = op->step_stage.as<double>();
  hipStream_t s = op->ctx->stream;
  HIPCHK(op->ctx, hipMemcpyAsync(d, un, nf * sizeof(double), hipMemcpyHostToDevice, s));
  HIPCHK(op->ctx, hipMemcpyAsync(d + nf + nd, unm1, nf * sizeof(double), hipMemcpyHostToDevice, s));
  if (nd) {
    HIPCHK(op->ctx, hipMemcpyAsync(d + nf, un_dir, nd * sizeof(double), hipMemcpyHostToDevice, s));
    HIPCHK(op->ctx, hipMemcpyAsync(d + 2 * nf + nd, unm1_dir, nd * sizeof(double), hipMemcpyHostToDevice, s));
  }
  const StepFunctions sf{d, d + nf, d + nf + nd, d + 2 * nf + nd, aw, bw, c1, c2};
  if ((rc = oseen_step_fields(op, sf, laplacian != 0, s))) return rc;
  HIPCHK(op->ctx, hipStreamSynchronize(s));
  return HFK_OK;
}

int hfk_oseen_set_laplacian(hfk_oseen* op, const double* lap_points) {
  int rc = oseen_bind(op);
  if (rc) return rc;
  if (!lap_points) {
    op->has_lap = false;
    return HFK_OK;
  }
  const size_t n = (size_t)op->ncells * op->np * 3;
  HIPCHK(op->ctx, op->lap.alloc(n * sizeof(double)));
  HIPCHK(op->ctx, hipMemcpy(op->lap.p, lap_points, n * sizeof(double), hipMemcpyHostToDevice));
  op->has_lap = true;
  return HFK_OK;
}

int hfk_oseen_set_outlet(hfk_oseen* op, int64_t nfaces, const int32_t* face_cell, const double* normals, int nqo,
                         const double* xo, int nqb, const double* xb, double Pc, double ZA) {
  int rc = oseen_bind(op);
  if (rc) return rc;
  if (nfaces == 0) {
    op->nfaces = 0;
    return HFK_OK;
  }
  if (nfaces < 0 || !face_cell || !normals || !xo || !xb || nqo < 1 || nqb < 1 || nqo > kMaxFacePoints ||
      nqb > kMaxFacePoints)
    return set_err(op->ctx, HFK_EARG, "HyperFEMKernels: invalid outlet faces (sizes, 1..32 points per measure)");
  for (int64_t f = 0; f < nfaces; ++f)
    if (face_cell[f] < 0 || face_cell[f] >= op->ncells)
      return set_err(op->ctx, HFK_EDIM, "HyperFEMKernels: outlet face cell out of range");
  // faces coloured among themselves, greedily in face order: a face takes the lowest colour that no earlier face
  // sharing a free dof of its cell holds (two faces of one cell share every dof), so the faces of one launch add
  // into disjoint CSC entries and residual rows; the launches run in colour order after every cell colour
  std::vector<int> fcol((size_t)nfaces, 0);
  size_t ng = 0;
  {
    std::vector<int64_t> dstart((size_t)op->nfree + 1, 0);  // free dof -> the faces touching it
    auto cell_ids = [&](int64_t f) { return op->cell_dofs.data() + (size_t)face_cell[f] * kND; };
    for (int64_t f = 0; f < nfaces; ++f)
      for (int k = 0; k < kND; ++k)
        if (cell_ids(f)[k] > 0) dstart[cell_ids(f)[k]]++;
    for (int64_t v = 0; v < op->nfree; ++v) dstart[v + 1] += dstart[v];
    std::vector<int64_t> dfill(dstart.begin(), dstart.end() - 1);
    std::vector<int32_t> dface((size_t)dstart[op->nfree]);
    for (int64_t f = 0; f < nfaces; ++f)
      for (int k = 0; k < kND; ++k)
        if (cell_ids(f)[k] > 0) dface[dfill[cell_ids(f)[k] - 1]++] = (int32_t)f;
    std::vector<int64_t> stamp((size_t)nfaces + 1, -1);  // stamp[c] == f: colour c taken by a neighbour of f
    for (int64_t f = 0; f < nfaces; ++f) {
      for (int k = 0; k < kND; ++k) {
        const int32_t id = cell_ids(f)[k];
        if (id <= 0) continue;
        for (int64_t e = dstart[id - 1]; e < dstart[id]; ++e)
          if (dface[e] < f) stamp[fcol[dface[e]]] = f;
      }
      int c = 0;
      while (stamp[c] == f) ++c;
      fcol[f] = c;
      ng = std::max(ng, (size_t)c + 1);
    }
  }
  auto group = [&](int64_t f) { return (size_t)fcol[f]; };
  std::vector<int64_t> start(ng + 1, 0);
  for (int64_t f = 0; f < nfaces; ++f) start[group(f) + 1]++;
  for (size_t g = 0; g < ng; ++g) start[g + 1] += start[g];
  std::vector<int64_t> fill(start.begin(), start.end() - 1);
  std::vector<int32_t> order((size_t)nfaces);
  for (int64_t f = 0; f < nfaces; ++f) order[fill[group(f)]++] = (int32_t)f;
  HIPCHK(op->ctx, op->face_cell.alloc((size_t)nfaces * sizeof(int32_t)));
  HIPCHK(op->ctx, hipMemcpy(op->face_cell.p, face_cell, (size_t)nfaces * sizeof(int32_t), hipMemcpyHostToDevice));
  HIPCHK(op->ctx, op->face_normal.alloc((size_t)nfaces * 3 * sizeof(double)));
  HIPCHK(op->ctx, hipMemcpy(op->face_normal.p, normals, (size_t)nfaces * 3 * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(op->ctx, op->face_xo.alloc((size_t)nfaces * nqo * 4 * sizeof(double)));
  HIPCHK(op->ctx, hipMemcpy(op->face_xo.p, xo, (size_t)nfaces * nqo * 4 * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(op->ctx, op->face_xb.alloc((size_t)nfaces * nqb * 7 * sizeof(double)));
  HIPCHK(op->ctx, hipMemcpy(op->face_xb.p, xb, (size_t)nfaces * nqb * 7 * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(op->ctx, op->faces_d.alloc((size_t)nfaces * sizeof(int32_t)));
  HIPCHK(op->ctx, hipMemcpy(op->faces_d.p, order.data(), (size_t)nfaces * sizeof(int32_t), hipMemcpyHostToDevice));
  op->face_start = start;
  op->nfaces = nfaces;
  op->nqo = nqo;
  op->nqb = nqb;
  op->Pc = Pc;
  op->ZA = ZA;
  return HFK_OK;
}

int hfk_oseen_set_deterministic(hfk_oseen* op, int deterministic) {
  int rc = oseen_bind(op);
  if (rc) return rc;
  op->deterministic = deterministic != 0;
  return HFK_OK;
}

int hfk_oseen_set_parameters(hfk_oseen* op, double rho, double mu, double c0) {
  int rc = oseen_bind(op);
  if (rc) return rc;
  if (!(rho > 0) || !(mu > 0) || !(c0 >= 0)) return set_err(op->ctx, HFK_EARG, "HyperFEMKernels: invalid rho, mu or c0");
  op->rho = rho;
  op->mu = mu;
  op->c0 = c0;
  return HFK_OK;
}

int hfk_oseen_set_form(hfk_oseen* op, const char* source, const char* name, const double* params, int nparams) {
  int rc = oseen_bind(op);
  if (rc) return rc;
  if (nparams < 0 || nparams > kFormParams - 3 || (nparams > 0 && !params))
    return set_err(op->ctx, HFK_EARG, "HyperFEMKernels: a point form takes 0..5 parameters after rho, mu, c0");
  if (!source) {  // back to the built-in integrand
    op->drop_form();
    op->form_params.clear();
    return HFK_OK;
  }
  std::unique_ptr<OseenFormKernels> k(new OseenFormKernels());
  if ((rc = rtc_oseen_form(op->ctx, source, name, k.get()))) return rc;
  op->drop_form();
  op->form = std::move(k);
  op->form_params.assign(params, params + nparams);
  return HFK_OK;
}

int hfk_oseen_set_blocks(hfk_oseen* op, const int32_t* mask, const int32_t* cell_mask) {
  int rc = oseen_bind(op);
  if (rc) return rc;
  if (!mask && !cell_mask) {
    op->block_mask.clear();
  } else {
    std::vector<uint8_t> m((size_t)op->ncells * 4);
    for (int64_t cell = 0; cell < op->ncells; ++cell)
      for (int k = 0; k < 4; ++k) m[(size_t)cell * 4 + k] = (cell_mask ? cell_mask[cell * 4 + k] : mask[k]) != 0;
    bool all = true;
    for (uint8_t v : m) all = all && v;
    if (all) m.clear();  // every block of every cell: the full pattern (cell_block_mask returns nothing)
    op->block_mask = std::move(m);
  }
  op->pattern_built = false;  // the pattern, its positions and the device copy are rebuilt at the next use
  op->pat.colptr.reset();
  op->pat.rowval.reset();
  op->pat.nzpos.reset();
  op->pat.nnz = 0;
  return HFK_OK;
}

int hfk_oseen_set_outlet_pressure(hfk_oseen* op, double Pc) {
  int rc = oseen_bind(op);
  if (rc) return rc;
  op->Pc = Pc;
  return HFK_OK;
}

int hfk_oseen_set_dirichlet(hfk_oseen* op, const double* dvals) {
  int rc = oseen_bind(op);
  if (rc) return rc;
  if (!dvals && op->ndir > 0) return set_err(op->ctx, HFK_EARG, "HyperFEMKernels: null Dirichlet values");
  if (op->ndir > 0) HIPCHK(op->ctx, hipMemcpy(op->dvals.p, dvals, (size_t)op->ndir * sizeof(double), hipMemcpyHostToDevice));
  return HFK_OK;
}

int hfk_oseen_info(hfk_oseen* op, int64_t* info) {
  if (!op || !info) return set_err(op ? op->ctx : nullptr, HFK_EARG, "HyperFEMKernels: null argument");
  info[0] = op->nfree;
  info[1] = op->ndir;
  info[2] = op->ncells;
  info[3] = op->np;
  info[4] = (int64_t)op->color_start.size() - 1;
  info[5] = op->pattern_built ? op->pat.nnz : -1;
  return HFK_OK;
}

int hfk_oseen_colors(hfk_oseen* op, int32_t* colors) {
  if (!op || !colors) return set_err(op ? op->ctx : nullptr, HFK_EARG, "HyperFEMKernels: null argument");
  std::copy(op->color.begin(), op->color.end(), colors);
  return HFK_OK;
}

int hfk_oseen_pattern_nnz(hfk_oseen* op, int64_t* nnz) {
  int rc = oseen_bind(op);
  if (rc) return rc;
  if (!nnz) return set_err(op->ctx, HFK_EARG, "HyperFEMKernels: null argument");
  if ((rc = oseen_pattern(op))) return rc;
  *nnz = op->pat.nnz;
  return HFK_OK;
}

int hfk_oseen_pattern(hfk_oseen* op, int64_t* colptr, int64_t* rowval, int32_t* nzpos) {
  int rc = oseen_bind(op);
  if (rc) return rc;
  if ((rc = oseen_pattern(op))) return rc;
  // nzpos in the host's local order: [cell][perm[ii] + 34 perm[jj]]
  return device_pattern_to_host(op->ctx, op->ctx->stream, op->pat, op->nfree, op->ncells, kND, colptr, rowval, nzpos,
                                op->perm.data());
}

int hfk_oseen_residual_dev(hfk_oseen* op, const double* x_dev, double* b_dev, void* stream) {
  int rc = oseen_bind(op);
  if (rc) return rc;
  if (!x_dev || !b_dev) return set_err(op->ctx, HFK_EARG, "HyperFEMKernels: null vector");
  return oseen_run(op, x_dev, b_dev, nullptr, op->stream(stream));
}

int hfk_oseen_jacobian_dev(hfk_oseen* op, const double* x_dev, double* nzval_dev, void* stream) {
  int rc = oseen_bind(op);
  if (rc) return rc;
  if (!nzval_dev) return set_err(op->ctx, HFK_EARG, "HyperFEMKernels: null vector");
  if ((rc = oseen_pattern(op))) return rc;
  (void)x_dev;  // the Oseen step is linear: the Jacobian does not depend on the state
  return oseen_run(op, nullptr, nullptr, nzval_dev, op->stream(stream));
}

int hfk_oseen_residual_jacobian_dev(hfk_oseen* op, const double* x_dev, double* b_dev, double* nzval_dev,
                                    void* stream) {
  int rc = oseen_bind(op);
  if (rc) return rc;
  if (!x_dev || !b_dev || !nzval_dev) return set_err(op->ctx, HFK_EARG, "HyperFEMKernels: null vector");
  if ((rc = oseen_pattern(op))) return rc;
  return oseen_run(op, x_dev, b_dev, nzval_dev, op->stream(stream));
}

int hfk_oseen_residual(hfk_oseen* op, const double* x, double* b) {
  int rc = oseen_bind(op);
  if (rc) return rc;
  if (!x || !b) return set_err(op->ctx, HFK_EARG, "HyperFEMKernels: null vector");
  hipStream_t s = op->ctx->stream;
  HIPCHK(op->ctx, op->x_stage.alloc((size_t)op->nfree * sizeof(double)));
  HIPCHK(op->ctx, op->y_stage.alloc((size_t)op->nfree * sizeof(double)));
  HIPCHK(op->ctx, hipMemcpyAsync(op->x_stage.p, x, (size_t)op->nfree * sizeof(double), hipMemcpyHostToDevice, s));
  if ((rc = oseen_run(op, op->x_stage.as<double>(), op->y_stage.as<double>(), nullptr, s))) return rc;
  HIPCHK(op->ctx, hipMemcpyAsync(b, op->y_stage.p, (size_t)op->nfree * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(op->ctx, hipStreamSynchronize(s));
  return HFK_OK;
}

int hfk_oseen_jacobian(hfk_oseen* op, const double* x, double* nzval) {
  int rc = oseen_bind(op);
  if (rc) return rc;
  if (!nzval) return set_err(op->ctx, HFK_EARG, "HyperFEMKernels: null vector");
  (void)x;
  if ((rc = oseen_pattern(op))) return rc;
  hipStream_t s = op->ctx->stream;
  const size_t nnz = (size_t)op->pat.nnz;
  HIPCHK(op->ctx, op->nz_stage.alloc(std::max<size_t>(nnz, 1) * sizeof(double)));
  if ((rc = oseen_run(op, nullptr, nullptr, op->nz_stage.as<double>(), s))) return rc;
  HIPCHK(op->ctx, hipMemcpyAsync(nzval, op->nz_stage.p, nnz * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(op->ctx, hipStreamSynchronize(s));
  return HFK_OK;
}

}  // extern "C"
