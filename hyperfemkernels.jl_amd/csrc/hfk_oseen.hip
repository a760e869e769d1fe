// One Oseen-BDF2 step of incompressible flow on tetrahedra, Taylor-Hood P2-P1 with SUPG/PSPG/LSIC stabilisation:
// coloured FP64 assembly of the residual and of the CSC Jacobian (the weak form of test/common.jl:165-228, volume
// terms; the reference assembles it through residual!/jacobian!, src/assembly.jl:909-927, and checks it against
// Gridap in test/test_cfd.jl).
//
//   a((u,p),(v,q)) = rho c u.v + 2 mu eps(u):eps(v) + rho ((grad u)' w).v - p div v - q div u
//                  + ts ((grad v)' w).Ru + ts ((grad v)' w).grad p - tp grad q.Ru - tp grad q.grad p
//                  + tl div v div u,                     Ru = rho c u + rho (grad u)' w
//   l((v,q))       = rho f.v + ts ((grad v)' w).(rho f) - tp grad q.(rho f)
//   ts = ((2|un|/h)^2 + 9 (4 mu/(rho h^2))^2)^(-1/2), tp = ts/rho, tl = rho |un| h / 2, h = (6 sqrt(2) |K|)^(1/3)
//
// (test/common.jl:201-205, Gridap's gradient convention grad u[J,c] = d u_c / d x_J).  The residual is
// r(x) = a(x, .) - l(.) on the free test dofs with the Dirichlet values lifted; the Jacobian is a(., .) on the free
// (test, trial) pairs.  w (advecting velocity), un (velocity of the stabilisation parameters) and f (BDF history)
// are point data given per cell and quadrature point.
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
#include "hfk_solver.hpp"

namespace hfk {
namespace {

constexpr int kND = 34;             // dofs per cell
constexpr int kCellsPerGroup = 3;   // 3 x 2 x 34 = 204 of 256 threads: two threads per element-matrix row
constexpr int kMaxPoints = 64;

struct OseenDev {
  int64_t ncells;
  int np;
  double rho, mu, c0;
  const int32_t* ids;     // [cell][34]
  const double* coords;   // [cell][4][3]
  const double* qp;       // [q][4]: xi, eta, zeta, weight
  const double* fields;   // [cell][q][9]: w, un, f
  const double* dvals;
  const int32_t* nzpos;   // [cell][ii + 34 jj], 1-based, 0: not stored
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

// The 17 entries [J0, J0 + 17) of the element-matrix row of test function (vel: velocity component ic of P2
// node a; else pressure vertex a - 10) at one point, added to acc with the weight dV.  S: the cell's 14 shape
// functions at the point (value, physical gradient, w . gradient).
template <int J0>
__device__ __forceinline__ void accumulate(const double (*S)[5], bool vel, int ic, int a, double dV, double ts,
                                           double tp, double tl, double c, double rho, double mu, double* acc) {
  const double Na = S[a][0], ga[3] = {S[a][1], S[a][2], S[a][3]}, wga = S[a][4];
  const double gai = ic == 0 ? ga[0] : (ic == 1 ? ga[1] : ga[2]);
#pragma unroll
  for (int t = 0; t < 17; ++t) {
    const int jj = J0 + t;
    double v;
    if (jj < 30) {  // trial velocity component jc at P2 node nb
      const int jc = jj / 10, nb = jj % 10;
      const double Nb = S[nb][0], wgb = S[nb][4], gb_j = S[nb][1 + jc];
      if (vel) {
        v = mu * ga[jc] * S[nb][1 + ic] + tl * gai * gb_j;
        if (jc == ic) {
          const double gab = ga[0] * S[nb][1] + ga[1] * S[nb][2] + ga[2] * S[nb][3];
          v += rho * c * Na * Nb + mu * gab + rho * Na * wgb + ts * rho * wga * (c * Nb + wgb);
        }
      } else {
        v = -Na * gb_j - tp * rho * ga[jc] * (c * Nb + wgb);
      }
    } else {  // trial pressure at vertex nb
      const int nb = 10 + jj - 30;
      if (vel) v = -S[nb][0] * gai + ts * wga * S[nb][1 + ic];
      else v = -tp * (ga[0] * S[nb][1] + ga[1] * S[nb][2] + ga[2] * S[nb][3]);
    }
    acc[t] += dV * v;
  }
}

// One work-group: kCellsPerGroup cells of one colour, two threads per element-matrix row (test function), each
// for 17 of its 34 entries.  Per quadrature point the cell's 14 shape functions (value, physical gradient,
// w . gradient) and the point scalars are staged in LDS; each thread accumulates its 17 entries (and the first of
// the two the row's load).  Then the entries are scattered into the CSC values through nzpos (free trial dofs),
// and/or the two halves of K x_e - l are added in LDS and written to the residual (free test dof).  The cells of
// one colour share no free dof, so the plain read-modify-writes do not race and the sums follow the colour order.
template <bool WANT_K, bool WANT_R>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_oseen(OseenDev op, const int32_t* __restrict__ cells, int64_t ncell_color,
                                               const double* __restrict__ x, double* __restrict__ b,
                                               double* __restrict__ nzval) {
  __shared__ double S[kCellsPerGroup][14][5];  // per cell: N, grad N (3), w . grad N
  __shared__ double P[kCellsPerGroup][12];     // per cell and point: dV, ts, tp, tl, c, w(3), f(3), spare
  __shared__ double Jit[kCellsPerGroup][9];    // inverse transposed Jacobian (physical gradient = Jit * ref)
  __shared__ double Cm[kCellsPerGroup][2];     // |det J|, h
  __shared__ double Xe[kCellsPerGroup][kND];
  __shared__ double Rp[kCellsPerGroup][2][kND];  // the two halves of the residual rows
  const int tid = threadIdx.x;
  const int lc = tid / (2 * kND), rem = tid - lc * 2 * kND;
  const int half = rem / kND, row = rem - half * kND;
  const int64_t k = (int64_t)blockIdx.x * kCellsPerGroup + lc;
  const bool active = lc < kCellsPerGroup && k < ncell_color;
  const int64_t cell = active ? cells[k] : 0;
  if (active && rem == 0) {
    const double* X = op.coords + cell * 12;
    double J[9];  // column d: x_{d+1} - x_0
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int i = 0; i < 3; ++i) J[i + 3 * d] = X[3 * (d + 1) + i] - X[i];
    const double det = J[0] * (J[4] * J[8] - J[7] * J[5]) - J[3] * (J[1] * J[8] - J[7] * J[2]) +
                       J[6] * (J[1] * J[5] - J[4] * J[2]);
    const double id = 1.0 / det;
    // inverse (column-major), then its transpose
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
      for (int j = 0; j < 3; ++j) Jit[lc][i + 3 * j] = I[j + 3 * i];
    const double vol = fabs(det) / 6.0;
    Cm[lc][0] = fabs(det);
    Cm[lc][1] = cbrt(6.0 * sqrt(2.0) * vol);
  }
  int32_t myid = 0;
  if (active) {
    myid = op.ids[cell * kND + row];
    if (half == 0) Xe[lc][row] = myid > 0 ? (x ? x[myid - 1] : 0.0) : op.dvals[-myid - 1];
  }
  // the row's test function: velocity component ic of node a, or pressure node a
  const bool vel = row < 30;
  const int ic = vel ? row / 10 : 0, a = vel ? row % 10 : 10 + (row - 30);
  double acc[17];
#pragma unroll
  for (int j = 0; j < 17; ++j) acc[j] = 0.0;
  double load = 0.0;
  for (int q = 0; q < op.np; ++q) {
    __syncthreads();
    if (active && rem < 14) {  // shape function `rem` of this cell at point q
      double n, g[3];
      shape(rem, op.qp[4 * q], op.qp[4 * q + 1], op.qp[4 * q + 2], n, g);
      const double* Ji = Jit[lc];
      const double* fq = op.fields + ((size_t)cell * op.np + q) * 9;
      double gx[3], wg = 0.0;
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        gx[i] = Ji[i] * g[0] + Ji[i + 3] * g[1] + Ji[i + 6] * g[2];
        wg += fq[i] * gx[i];
      }
      S[lc][rem][0] = n;
      S[lc][rem][1] = gx[0];
      S[lc][rem][2] = gx[1];
      S[lc][rem][3] = gx[2];
      S[lc][rem][4] = wg;
    }
    if (active && rem == 14) {  // point scalars: dV, stabilisation parameters, w, f
      const double* fq = op.fields + ((size_t)cell * op.np + q) * 9;
      const double un = sqrt(fq[3] * fq[3] + fq[4] * fq[4] + fq[5] * fq[5]);
      const double h = Cm[lc][1];
      const double r1 = 2.0 * un / h, r2 = 3.0 * (4.0 * op.mu / (op.rho * h * h));
      const double ts = 1.0 / sqrt(r1 * r1 + r2 * r2);
      P[lc][0] = Cm[lc][0] * op.qp[4 * q + 3];
      P[lc][1] = ts;
      P[lc][2] = ts / op.rho;
      P[lc][3] = op.rho * un * h * 0.5;
      P[lc][4] = op.c0;
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        P[lc][5 + i] = fq[i];
        P[lc][8 + i] = fq[6 + i];
      }
    }
    __syncthreads();
    if (!active) continue;
    const double dV = P[lc][0], ts = P[lc][1], tp = P[lc][2], tl = P[lc][3], c = P[lc][4];
    if (half == 0) {
      accumulate<0>(S[lc], vel, ic, a, dV, ts, tp, tl, c, op.rho, op.mu, acc);
      const double Na = S[lc][a][0], wga = S[lc][a][4];
      const double f[3] = {P[lc][8], P[lc][9], P[lc][10]};
      if (vel) {
        const double fi = ic == 0 ? f[0] : (ic == 1 ? f[1] : f[2]);
        load += dV * (op.rho * fi * Na + ts * wga * op.rho * fi);
      } else {
        load += dV * (-tp * op.rho * (S[lc][a][1] * f[0] + S[lc][a][2] * f[1] + S[lc][a][3] * f[2]));
      }
    } else {
      accumulate<17>(S[lc], vel, ic, a, dV, ts, tp, tl, c, op.rho, op.mu, acc);
    }
  }
  const bool assemble = active && myid > 0;  // Dirichlet test rows are not assembled
  const int j0 = half * 17;
  if (WANT_K && assemble) {
    const int32_t* pos = op.nzpos + cell * kND * kND + row;
#pragma unroll
    for (int t = 0; t < 17; ++t) {
      const int32_t p = pos[kND * (j0 + t)];
      if (p > 0) nzval[p - 1] += acc[t];
    }
  }
  if (WANT_R) {
    if (active) {
      double r = half == 0 ? -load : 0.0;
#pragma unroll
      for (int t = 0; t < 17; ++t) r += acc[t] * Xe[lc][j0 + t];
      Rp[lc][half][row] = r;
    }
    __syncthreads();
    if (assemble && half == 0) b[myid - 1] += Rp[lc][0][row] + Rp[lc][1][row];
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
  std::vector<int32_t> cell_dofs, color;
  std::vector<int64_t> color_start;  // cells of colour c: cells_by_color[color_start[c] .. color_start[c+1])
  bool pattern_built = false, has_fields = false;
  std::vector<int64_t> colptr, rowval;
  std::vector<int32_t> nzpos;
  DevBuf ids, coords, qp, fields, dvals, cells_d, nzpos_d, x_stage, y_stage, nz_stage;
  hipStream_t stream(void* s) const { return s ? static_cast<hipStream_t>(s) : ctx->stream; }
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
    d.dvals = dvals.as<double>();
    d.nzpos = nzpos_d.as<int32_t>();
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
  csc_pattern(op->cell_dofs.data(), kND, op->ncells, op->nfree, op->colptr, op->rowval, op->nzpos);
  HIPCHK(op->ctx, op->nzpos_d.alloc(op->nzpos.size() * sizeof(int32_t)));
  HIPCHK(op->ctx, hipMemcpy(op->nzpos_d.p, op->nzpos.data(), op->nzpos.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  op->pattern_built = true;
  return HFK_OK;
}

// colour by colour; x may be null (zero free values) when only K is wanted
int oseen_run(hfk_oseen* op, const double* x, double* b, double* nzval, hipStream_t s) {
  if (!op->has_fields) return set_err(op->ctx, HFK_ESTATE, "HyperFEMKernels: Oseen fields (w, un, f) not set");
  const OseenDev d = op->dev();
  if (b) HIPCHK(op->ctx, launch_fill(b, op->nfree, 0.0, s));
  if (nzval) HIPCHK(op->ctx, launch_fill(nzval, (int64_t)op->rowval.size(), 0.0, s));
  for (size_t c = 0; c + 1 < op->color_start.size(); ++c) {
    const int64_t n = op->color_start[c + 1] - op->color_start[c];
    if (n <= 0) continue;
    const int32_t* cells = op->cells_d.as<int32_t>() + op->color_start[c];
    const dim3 g((unsigned)((n + kCellsPerGroup - 1) / kCellsPerGroup)), t(256);
    if (b && nzval) hipLaunchKernelGGL((k_oseen<true, true>), g, t, 0, s, d, cells, n, x, b, nzval);
    else if (b) hipLaunchKernelGGL((k_oseen<false, true>), g, t, 0, s, d, cells, n, x, b, nzval);
    else hipLaunchKernelGGL((k_oseen<true, false>), g, t, 0, s, d, cells, n, x, b, nzval);
    HIPCHK(op->ctx, hipGetLastError());
  }
  return HFK_OK;
}

}  // namespace

extern "C" {

int hfk_oseen_create(hfk_ctx* ctx, int64_t ncells, const int32_t* cell_dofs, const double* coords, int np,
                     const double* qpoints, const double* qweights, int64_t nfree, int64_t ndir, double rho, double mu,
                     double c0, hfk_oseen** out) {
  if (!ctx || !out || !cell_dofs || !coords || !qpoints || !qweights)
    return set_err(ctx, HFK_EARG, "HyperFEMKernels: null argument");
  if (ncells <= 0 || nfree <= 0 || ndir < 0 || np < 1 || np > kMaxPoints)
    return set_err(ctx, HFK_EARG, "HyperFEMKernels: invalid Oseen sizes (cells, dofs, 1..64 points)");
  if (!(rho > 0) || !(mu > 0) || !(c0 >= 0)) return set_err(ctx, HFK_EARG, "HyperFEMKernels: invalid rho, mu or c0");
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
  op->cell_dofs.assign(cell_dofs, cell_dofs + ncells * kND);
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
  HIPCHK(ctx, hipMemcpy(op->ids.p, cell_dofs, (size_t)ncells * kND * sizeof(int32_t), hipMemcpyHostToDevice));
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
  info[5] = op->pattern_built ? (int64_t)op->rowval.size() : -1;
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
  *nnz = (int64_t)op->rowval.size();
  return HFK_OK;
}

int hfk_oseen_pattern(hfk_oseen* op, int64_t* colptr, int64_t* rowval, int32_t* nzpos) {
  int rc = oseen_bind(op);
  if (rc) return rc;
  if ((rc = oseen_pattern(op))) return rc;
  if (colptr) std::copy(op->colptr.begin(), op->colptr.end(), colptr);
  if (rowval) std::copy(op->rowval.begin(), op->rowval.end(), rowval);
  if (nzpos) std::copy(op->nzpos.begin(), op->nzpos.end(), nzpos);
  return HFK_OK;
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
  const size_t nnz = op->rowval.size();
  HIPCHK(op->ctx, op->nz_stage.alloc(std::max<size_t>(nnz, 1) * sizeof(double)));
  if ((rc = oseen_run(op, nullptr, nullptr, op->nz_stage.as<double>(), s))) return rc;
  HIPCHK(op->ctx, hipMemcpyAsync(nzval, op->nz_stage.p, nnz * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(op->ctx, hipStreamSynchronize(s));
  return HFK_OK;
}

}  // extern "C"
