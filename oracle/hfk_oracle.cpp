// TEST INFRASTRUCTURE ONLY — the oracle (see hfk_oracle.hpp for conventions and who may load it).
//
// Parity pins for every function here live in tests/test_oracle_*.py: closed forms (Gauss-Legendre,
// Lagrange, linear-elastic tangents at F = I), finite differences (tangent = derivative of P, Jacobian =
// derivative of the residual), brute force on tiny meshes (sum factorised = dense element matrices,
// pattern = Python set of pairs, colouring invariant), and the reference's own known answers.
#include "hfk_oracle.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

thread_local std::string g_err;
int fail(const std::string& s) {
  g_err = s;
  return -1;
}

// 3x3 helpers, column-major a[i + 3 j]
double det3(const double* a) {
  return a[0] * (a[4] * a[8] - a[7] * a[5]) - a[3] * (a[1] * a[8] - a[7] * a[2]) + a[6] * (a[1] * a[5] - a[4] * a[2]);
}
void inv3(const double* a, double* r) {
  double d = det3(a);
  r[0] = (a[4] * a[8] - a[7] * a[5]) / d;
  r[3] = -(a[3] * a[8] - a[6] * a[5]) / d;
  r[6] = (a[3] * a[7] - a[6] * a[4]) / d;
  r[1] = -(a[1] * a[8] - a[7] * a[2]) / d;
  r[4] = (a[0] * a[8] - a[6] * a[2]) / d;
  r[7] = -(a[0] * a[7] - a[6] * a[1]) / d;
  r[2] = (a[1] * a[5] - a[4] * a[2]) / d;
  r[5] = -(a[0] * a[5] - a[3] * a[2]) / d;
  r[8] = (a[0] * a[4] - a[3] * a[1]) / d;
}

}  // namespace

struct orc_op {
  int Q, P, nn, nc = 3, nd, np, ncells;
  int64_t nfree, ndir;
  std::vector<int32_t> dofs, nodeperm, ldof, tensor_of;  // tensor_of: table node -> tensor node
  std::vector<double> coords, body, mufield, dvals;
  int law;
  double lam, mu;
  std::vector<double> xq1, wq1, N1, dN1;  // N1[a*Q + i]
  // dense ND tables in table-node order: Nt[q*nn + n], dNt[(q*nn + n)*3 + K]
  std::vector<double> Nt, dNt, wq;
  std::vector<std::vector<int32_t>> colors;
};

extern "C" {

const char* orc_last_error(void) { return g_err.c_str(); }

// Gauss-Legendre points and weights on [0,1], ascending (Newton on the Legendre recurrence).
void orc_gauss01(int P, double* x, double* w) {
  for (int i = 0; i < P; ++i) {
    double z = std::cos(M_PI * (i + 0.75) / (P + 0.5)), dp = 0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1, p1 = z;
      for (int j = 2; j <= P; ++j) {
        double p2 = ((2 * j - 1) * z * p1 - (j - 1) * p0) / j;
        p0 = p1;
        p1 = p2;
      }
      if (P == 1) { p1 = z; p0 = 1; }
      dp = P * (z * p1 - p0) / (z * z - 1);
      double dz = p1 / dp;
      z -= dz;
      if (std::fabs(dz) < 1e-16) break;
    }
    x[i] = 0.5 * (1 - z);
    w[i] = 1.0 / ((1 - z * z) * dp * dp);  // = w_[-1,1] / 2
  }
}

// tensorproduct.jl:24-47: Lagrange basis on Q equispaced nodes of [0,1] (midpoint if Q = 1)
void orc_lagrange1d(int Q, double xi, double* N, double* dN) {
  std::vector<double> xs(Q);
  for (int i = 0; i < Q; ++i) xs[i] = Q == 1 ? 0.5 : double(i) / (Q - 1);
  for (int i = 0; i < Q; ++i) {
    double v = 1;
    for (int j = 0; j < Q; ++j)
      if (j != i) v *= (xi - xs[j]) / (xs[i] - xs[j]);
    N[i] = v;
    double d = 0;
    for (int m = 0; m < Q; ++m) {
      if (m == i) continue;
      double t = 1.0 / (xs[i] - xs[m]);
      for (int j = 0; j < Q; ++j)
        if (j != i && j != m) t *= (xi - xs[j]) / (xs[i] - xs[j]);
      d += t;
    }
    dN[i] = d;
  }
}

// Pointwise laws in slot form: g[J + 3c] = du_c/dx_J, F_cJ = delta + g[J+3c]; R[J + 3c] = P_cJ;
// T[sv + 9 su] = dP_cJ/dF_kL (sv = J + 3c, su = L + 3k).
void orc_law_P(int law, double lam, double mu, const double* g, double* R) {
  double F[9];  // F[c + 3J] column-major F_cJ
  for (int c = 0; c < 3; ++c)
    for (int J = 0; J < 3; ++J) F[c + 3 * J] = (c == J ? 1.0 : 0.0) + g[J + 3 * c];
  double Pm[9];
  if (law == 0) {  // SVK: P = F S, S = lam tr(E) I + 2 mu E, E = (F'F - I)/2  (test/common.jl:2-8)
    double E[9], S[9];
    for (int I = 0; I < 3; ++I)
      for (int J = 0; J < 3; ++J) {
        double s = 0;
        for (int k = 0; k < 3; ++k) s += F[k + 3 * I] * F[k + 3 * J];
        E[I + 3 * J] = 0.5 * (s - (I == J ? 1.0 : 0.0));
      }
    double tr = E[0] + E[4] + E[8];
    for (int I = 0; I < 3; ++I)
      for (int J = 0; J < 3; ++J) S[I + 3 * J] = (I == J ? lam * tr : 0.0) + 2 * mu * E[I + 3 * J];
    for (int c = 0; c < 3; ++c)
      for (int J = 0; J < 3; ++J) {
        double s = 0;
        for (int I = 0; I < 3; ++I) s += F[c + 3 * I] * S[I + 3 * J];
        Pm[c + 3 * J] = s;
      }
  } else if (law == 1) {  // neo-Hookean: P = mu (F - F^-T) + lam lnJ F^-T
    double A[9];
    inv3(F, A);  // A = F^-1, A[J + 3c] = (F^-1)_Jc
    double lnJ = std::log(det3(F));
    for (int c = 0; c < 3; ++c)
      for (int J = 0; J < 3; ++J) Pm[c + 3 * J] = mu * F[c + 3 * J] + (lam * lnJ - mu) * A[J + 3 * c];
  } else {  // linear elasticity on G = F - I: sigma = lam tr(G) I + mu (G + G')
    double trG = g[0] + g[4] + g[8];
    for (int c = 0; c < 3; ++c)
      for (int J = 0; J < 3; ++J)
        Pm[c + 3 * J] = (c == J ? lam * trG : 0.0) + mu * (g[J + 3 * c] + g[c + 3 * J]);
  }
  for (int c = 0; c < 3; ++c)
    for (int J = 0; J < 3; ++J) R[J + 3 * c] = Pm[c + 3 * J];
}

void orc_law_T(int law, double lam, double mu, const double* g, double* T) {
  double F[9];
  for (int c = 0; c < 3; ++c)
    for (int J = 0; J < 3; ++J) F[c + 3 * J] = (c == J ? 1.0 : 0.0) + g[J + 3 * c];
  auto d = [](int a, int b) { return a == b ? 1.0 : 0.0; };
  if (law == 0) {
    // dP = dF S + F dS, dS = lam tr(dE) I + 2 mu dE, dE = (dF'F + F'dF)/2; T = dP_cJ/dF_kL
    double E[9], S[9], C[9];
    for (int I = 0; I < 3; ++I)
      for (int J = 0; J < 3; ++J) {
        double s = 0;
        for (int k = 0; k < 3; ++k) s += F[k + 3 * I] * F[k + 3 * J];
        C[I + 3 * J] = s;
        E[I + 3 * J] = 0.5 * (s - d(I, J));
      }
    double tr = E[0] + E[4] + E[8];
    for (int I = 0; I < 3; ++I)
      for (int J = 0; J < 3; ++J) S[I + 3 * J] = d(I, J) * lam * tr + 2 * mu * E[I + 3 * J];
    (void)C;
    for (int c = 0; c < 3; ++c)
      for (int J = 0; J < 3; ++J)
        for (int k = 0; k < 3; ++k)
          for (int L = 0; L < 3; ++L) {
            // d(F S)_cJ / dF_kL = delta_ck S_LJ + sum_I F_cI dS_IJ/dF_kL
            // dE_IJ/dF_kL = (delta_IL F_kJ + F_kI delta_JL)/2 ; tr(dE)/dF_kL = F_kL
            double v = d(c, k) * S[L + 3 * J];
            for (int I = 0; I < 3; ++I) {
              double dE = 0.5 * (d(I, L) * F[k + 3 * J] + F[k + 3 * I] * d(J, L));
              double dS = d(I, J) * lam * F[k + 3 * L] + 2 * mu * dE;
              v += F[c + 3 * I] * dS;
            }
            T[(J + 3 * c) + 9 * (L + 3 * k)] = v;
          }
  } else if (law == 1) {
    // T = mu d_ck d_JL + (mu - lam lnJ) A_Jk A_Lc + lam A_Jc A_Lk, A = F^-1
    double A[9];
    inv3(F, A);
    double lnJ = std::log(det3(F));
    for (int c = 0; c < 3; ++c)
      for (int J = 0; J < 3; ++J)
        for (int k = 0; k < 3; ++k)
          for (int L = 0; L < 3; ++L)
            T[(J + 3 * c) + 9 * (L + 3 * k)] = mu * d(c, k) * d(J, L) + (mu - lam * lnJ) * A[J + 3 * k] * A[L + 3 * c] +
                                               lam * A[J + 3 * c] * A[L + 3 * k];
  } else {
    for (int c = 0; c < 3; ++c)
      for (int J = 0; J < 3; ++J)
        for (int k = 0; k < 3; ++k)
          for (int L = 0; L < 3; ++L)
            T[(J + 3 * c) + 9 * (L + 3 * k)] = lam * d(c, J) * d(k, L) + mu * (d(c, k) * d(J, L) + d(J, k) * d(c, L));
  }
}

// Stored energy densities whose first derivative is orc_law_P (the pins in tests/ check P = dW/dF and
// T = dP/dF by finite differences): SVK W = lam/2 tr(E)^2 + mu E:E; neo-Hookean
// W = mu/2 (I1 - 3) - mu lnJ + lam/2 lnJ^2; linear W = lam/2 tr(eps)^2 + mu eps:eps, eps = sym(G).
double orc_law_W(int law, double lam, double mu, const double* g) {
  double F[9];
  for (int c = 0; c < 3; ++c)
    for (int J = 0; J < 3; ++J) F[c + 3 * J] = (c == J ? 1.0 : 0.0) + g[J + 3 * c];
  if (law == 0) {
    double E[9], tr = 0, ee = 0;
    for (int I = 0; I < 3; ++I)
      for (int J = 0; J < 3; ++J) {
        double s = 0;
        for (int k = 0; k < 3; ++k) s += F[k + 3 * I] * F[k + 3 * J];
        E[I + 3 * J] = 0.5 * (s - (I == J ? 1.0 : 0.0));
      }
    for (int i = 0; i < 3; ++i) tr += E[i + 3 * i];
    for (int i = 0; i < 9; ++i) ee += E[i] * E[i];
    return 0.5 * lam * tr * tr + mu * ee;
  } else if (law == 1) {
    double I1 = 0;
    for (int i = 0; i < 9; ++i) I1 += F[i] * F[i];
    double lnJ = std::log(det3(F));
    return 0.5 * mu * (I1 - 3) - mu * lnJ + 0.5 * lam * lnJ * lnJ;
  } else {
    double tr = g[0] + g[4] + g[8], ee = 0;
    for (int c = 0; c < 3; ++c)
      for (int J = 0; J < 3; ++J) {
        double e = 0.5 * (g[J + 3 * c] + g[c + 3 * J]);
        ee += e * e;
      }
    return 0.5 * lam * tr * tr + mu * ee;
  }
}

// coloring.jl:4-28: greedy first fit on the free test dofs, 128-bit masks
static int color_cells(orc_op* op) {
  int nd = op->nd;
  std::vector<unsigned __int128> masks(op->nfree, 0);
  std::vector<int> color(op->ncells);
  int ncol = 0;
  for (int cell = 0; cell < op->ncells; ++cell) {
    unsigned __int128 used = 0;
    for (int j = 0; j < nd; ++j) {
      int32_t r = op->dofs[(size_t)cell * nd + j];
      if (r > 0) used |= masks[r - 1];
    }
    int c = 0;
    while (c < 128 && ((used >> c) & 1)) ++c;  // trailing_ones
    if (c >= 128) return fail("oracle: more than 128 colours needed");
    color[cell] = c;
    unsigned __int128 bit = (unsigned __int128)1 << c;
    for (int j = 0; j < nd; ++j) {
      int32_t r = op->dofs[(size_t)cell * nd + j];
      if (r > 0) masks[r - 1] |= bit;
    }
    ncol = std::max(ncol, c + 1);
  }
  op->colors.assign(ncol, {});
  for (int cell = 0; cell < op->ncells; ++cell) op->colors[color[cell]].push_back(cell);
  return 0;
}

orc_op* orc_op_create(int Q, int P, int ncells, const int32_t* cell_dofs, const int32_t* nodeperm, const int32_t* ldof,
                      const double* coords, int64_t nfree, int64_t ndir, int law, double lam, double mu,
                      const double* body) {
  orc_op* op = new orc_op();
  op->Q = Q;
  op->P = P;
  op->nn = Q * Q * Q;
  op->nd = 3 * op->nn;
  op->np = P * P * P;
  op->ncells = ncells;
  op->nfree = nfree;
  op->ndir = ndir;
  op->law = law;
  op->lam = lam;
  op->mu = mu;
  op->dofs.assign(cell_dofs, cell_dofs + (size_t)ncells * op->nd);
  op->nodeperm.assign(nodeperm, nodeperm + op->nn);
  op->ldof.assign(ldof, ldof + op->nd);
  op->tensor_of.assign(op->nn, 0);
  for (int t = 0; t < op->nn; ++t) op->tensor_of[nodeperm[t]] = t;
  op->coords.assign(coords, coords + (size_t)ncells * 24);
  op->body.assign(3, 0.0);
  if (body) op->body.assign(body, body + 3);
  op->dvals.assign(ndir, 0.0);
  op->xq1.resize(P);
  op->wq1.resize(P);
  orc_gauss01(P, op->xq1.data(), op->wq1.data());
  op->N1.resize(P * Q);
  op->dN1.resize(P * Q);
  for (int a = 0; a < P; ++a) orc_lagrange1d(Q, op->xq1[a], &op->N1[a * Q], &op->dN1[a * Q]);
  int nn = op->nn, np = op->np;
  op->Nt.resize((size_t)np * nn);
  op->dNt.resize((size_t)np * nn * 3);
  op->wq.resize(np);
  for (int q = 0; q < np; ++q) {
    int a = q % P, b = (q / P) % P, c = q / (P * P);
    op->wq[q] = op->wq1[a] * op->wq1[b] * op->wq1[c];
    for (int n = 0; n < nn; ++n) {
      int t = op->tensor_of[n];
      int i = t % Q, j = (t / Q) % Q, k = t / (Q * Q);
      double Na = op->N1[a * Q + i], Nb = op->N1[b * Q + j], Nc = op->N1[c * Q + k];
      op->Nt[(size_t)q * nn + n] = Na * Nb * Nc;
      op->dNt[((size_t)q * nn + n) * 3 + 0] = op->dN1[a * Q + i] * Nb * Nc;
      op->dNt[((size_t)q * nn + n) * 3 + 1] = Na * op->dN1[b * Q + j] * Nc;
      op->dNt[((size_t)q * nn + n) * 3 + 2] = Na * Nb * op->dN1[c * Q + k];
    }
  }
  if (color_cells(op) != 0) {
    delete op;
    return nullptr;
  }
  return op;
}

void orc_op_destroy(orc_op* op) { delete op; }
void orc_op_set_mu_field(orc_op* op, const double* mu) {
  if (mu)
    op->mufield.assign(mu, mu + (size_t)op->ncells * op->np);
  else
    op->mufield.clear();
}
void orc_op_set_dirichlet(orc_op* op, const double* dvals) { op->dvals.assign(dvals, dvals + op->ndir); }
int orc_op_ncolors(orc_op* op) { return (int)op->colors.size(); }
int orc_op_colors(orc_op* op, int32_t* color_of_cell) {
  for (size_t c = 0; c < op->colors.size(); ++c)
    for (int32_t cell : op->colors[c]) color_of_cell[cell] = (int32_t)c + 1;
  return 0;
}

}  // extern "C"

namespace {

// assembly.jl:1073-1111: Jt[K + 3 i] = sum_v dM_v/dxi_K x_v[i]; geometry per point (dV, inv(Jt))
void point_geometry(const orc_op* op, int cell, int q, double* iJt, double* dV) {
  int P = op->P;
  int a = q % P, b = (q / P) % P, c = q / (P * P);
  double xi[3] = {op->xq1[a], op->xq1[b], op->xq1[c]};
  double Jt[9] = {0};
  for (int v = 0; v < 8; ++v) {
    int bits[3] = {v & 1, (v >> 1) & 1, (v >> 2) & 1};
    double m[3], dm[3];
    for (int d = 0; d < 3; ++d) {
      m[d] = bits[d] ? xi[d] : 1 - xi[d];
      dm[d] = bits[d] ? 1.0 : -1.0;
    }
    double dM[3] = {dm[0] * m[1] * m[2], m[0] * dm[1] * m[2], m[0] * m[1] * dm[2]};
    const double* x = &op->coords[((size_t)cell * 8 + v) * 3];
    for (int K = 0; K < 3; ++K)
      for (int i = 0; i < 3; ++i) Jt[K + 3 * i] += dM[K] * x[i];
  }
  inv3(Jt, iJt);
  *dV = op->wq[q] * std::fabs(det3(Jt));
}

// assembly.jl:1060-1070: local values, free from x and Dirichlet from dvals
void gather_state(const orc_op* op, int cell, const double* x, double* xl) {
  for (int j = 0; j < op->nd; ++j) {
    int32_t id = op->dofs[(size_t)cell * op->nd + j];
    xl[j] = id > 0 ? x[id - 1] : op->dvals[-id - 1];
  }
}

// kernels.jl:210-232 with a nodal vector field: g[J + 3c] = sum_K iJt[J,K] sum_n x_{n,c} dN_n/dxi_K
void point_gradient(const orc_op* op, const double* xl, int q, const double* iJt, double* g) {
  int nn = op->nn;
  for (int c = 0; c < 3; ++c) {
    double G[3] = {0, 0, 0};
    for (int n = 0; n < nn; ++n) {
      double xv = xl[op->ldof[c * nn + n]];
      for (int K = 0; K < 3; ++K) G[K] += xv * op->dNt[((size_t)q * nn + n) * 3 + K];
    }
    for (int J = 0; J < 3; ++J) {
      double s = 0;
      for (int K = 0; K < 3; ++K) s += iJt[J + 3 * K] * G[K];
      g[J + 3 * c] = s;
    }
  }
}

double mu_at(const orc_op* op, int cell, int q) {
  return op->mufield.empty() ? op->mu : op->mufield[(size_t)cell * op->np + q];
}

// tensorproduct.jl:125-159 (3D): value and reference gradient at the tensor points from tensor nodal values
void tp_interpolate(const orc_op* op, const double* u, double* out /*np*4*/) {
  int Q = op->Q, P = op->P;
  const double *N1 = op->N1.data(), *dN1 = op->dN1.data();
  std::vector<double> tA(P * Q * Q), tB(P * Q * Q), v2(P * P * Q), g1(P * P * Q), g2(P * P * Q);
  for (int k = 0; k < Q; ++k)
    for (int j = 0; j < Q; ++j)
      for (int a = 0; a < P; ++a) {
        double s = 0, ds = 0;
        for (int i = 0; i < Q; ++i) {
          double v = u[i + Q * j + Q * Q * k];
          s += N1[a * Q + i] * v;
          ds += dN1[a * Q + i] * v;
        }
        tA[a + P * (j + Q * k)] = s;
        tB[a + P * (j + Q * k)] = ds;
      }
  for (int k = 0; k < Q; ++k)
    for (int b = 0; b < P; ++b)
      for (int a = 0; a < P; ++a) {
        double v = 0, x1 = 0, x2 = 0;
        for (int j = 0; j < Q; ++j) {
          double ta = tA[a + P * (j + Q * k)], tb = tB[a + P * (j + Q * k)];
          v += N1[b * Q + j] * ta;
          x1 += N1[b * Q + j] * tb;
          x2 += dN1[b * Q + j] * ta;
        }
        int r = a + P * b;
        v2[r + P * P * k] = v;
        g1[r + P * P * k] = x1;
        g2[r + P * P * k] = x2;
      }
  for (int c = 0; c < P; ++c)
    for (int r = 0; r < P * P; ++r) {
      double v = 0, x1 = 0, x2 = 0, x3 = 0;
      for (int k = 0; k < Q; ++k) {
        v += N1[c * Q + k] * v2[r + P * P * k];
        x1 += N1[c * Q + k] * g1[r + P * P * k];
        x2 += N1[c * Q + k] * g2[r + P * P * k];
        x3 += dN1[c * Q + k] * v2[r + P * P * k];
      }
      int pl = r + P * P * c;
      out[pl * 4 + 0] = v;
      out[pl * 4 + 1] = x1;
      out[pl * 4 + 2] = x2;
      out[pl * 4 + 3] = x3;
    }
}

// tensorproduct.jl:189-224 (3D): r[n] = sum_pl N_n coef0 + sum_d dN_n/dxi_d coef_d
void tp_contract(const orc_op* op, const double* coef /*np*4*/, double* r) {
  int Q = op->Q, P = op->P;
  const double *N1 = op->N1.data(), *dN1 = op->dN1.data();
  std::vector<double> t0(P * P * Q), t1(P * P * Q), t2(P * P * Q), s0(P * Q * Q), s1(P * Q * Q);
  for (int k = 0; k < Q; ++k)
    for (int r2 = 0; r2 < P * P; ++r2) {
      double a0 = 0, a1 = 0, a2 = 0;
      for (int c = 0; c < P; ++c) {
        int pl = r2 + P * P * c;
        a0 += N1[c * Q + k] * coef[pl * 4 + 0] + dN1[c * Q + k] * coef[pl * 4 + 3];
        a1 += N1[c * Q + k] * coef[pl * 4 + 1];
        a2 += N1[c * Q + k] * coef[pl * 4 + 2];
      }
      t0[r2 + P * P * k] = a0;
      t1[r2 + P * P * k] = a1;
      t2[r2 + P * P * k] = a2;
    }
  for (int k = 0; k < Q; ++k)
    for (int j = 0; j < Q; ++j)
      for (int a = 0; a < P; ++a) {
        double b0 = 0, b1 = 0;
        for (int b = 0; b < P; ++b) {
          int r2 = a + P * b;
          b0 += N1[b * Q + j] * t0[r2 + P * P * k] + dN1[b * Q + j] * t2[r2 + P * P * k];
          b1 += N1[b * Q + j] * t1[r2 + P * P * k];
        }
        s0[a + P * (j + Q * k)] = b0;
        s1[a + P * (j + Q * k)] = b1;
      }
  for (int k = 0; k < Q; ++k)
    for (int j = 0; j < Q; ++j)
      for (int i = 0; i < Q; ++i) {
        double s = 0;
        for (int a = 0; a < P; ++a) s += N1[a * Q + i] * s0[a + P * (j + Q * k)] + dN1[a * Q + i] * s1[a + P * (j + Q * k)];
        r[i + Q * j + Q * Q * k] = s;
      }
}

}  // namespace

extern "C" {

int orc_prepare(orc_op* op, const double* x, double* T, double* iJt, double* dV) {
  int np = op->np, nd = op->nd;
#pragma omp parallel for schedule(static)
  for (int cell = 0; cell < op->ncells; ++cell) {
    std::vector<double> xl(nd);
    gather_state(op, cell, x, xl.data());
    for (int q = 0; q < np; ++q) {
      double ij[9], dv, g[9], Tq[81];
      point_geometry(op, cell, q, ij, &dv);
      point_gradient(op, xl.data(), q, ij, g);
      orc_law_T(op->law, op->lam, mu_at(op, cell, q), g, Tq);
      size_t base = (size_t)cell * np + q;
      for (int i = 0; i < 81; ++i) T[base * 81 + i] = Tq[i];
      for (int i = 0; i < 9; ++i) iJt[base * 9 + i] = ij[i];
      dV[base] = dv;
    }
  }
  return 0;
}

int orc_jv(orc_op* op, const double* T, const double* iJt, const double* dV, const double* dx, double* y, int nthreads) {
  int np = op->np, nd = op->nd, nn = op->nn;
  for (int64_t i = 0; i < op->nfree; ++i) y[i] = 0.0;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
  for (const auto& col : op->colors) {
    int ncol = (int)col.size();
#pragma omp parallel
    {
      std::vector<double> xl(nd), re(nd), uc(nn), rc(nn), pts(np * 4), coef(np * 4), su(9 * np), sv(9 * np);
#pragma omp for schedule(static)
      for (int ic = 0; ic < ncol; ++ic) {
        int cell = col[ic];
        const int32_t* ids = &op->dofs[(size_t)cell * nd];
        for (int j = 0; j < nd; ++j) xl[j] = ids[j] > 0 ? dx[ids[j] - 1] : 0.0;
        for (int c = 0; c < 3; ++c) {
          for (int t = 0; t < nn; ++t) uc[t] = xl[op->ldof[c * nn + op->nodeperm[t]]];
          tp_interpolate(op, uc.data(), pts.data());
          for (int q = 0; q < np; ++q) {
            const double* ij = &iJt[((size_t)cell * np + q) * 9];
            for (int J = 0; J < 3; ++J) {
              double s = 0;
              for (int K = 0; K < 3; ++K) s += ij[J + 3 * K] * pts[q * 4 + 1 + K];
              su[(J + 3 * c) + 9 * q] = s;
            }
          }
        }
        for (int q = 0; q < np; ++q) {
          const double* Tq = &T[((size_t)cell * np + q) * 81];
          double dv = dV[(size_t)cell * np + q];
          for (int i = 0; i < 9; ++i) {
            double s = 0;
            for (int j = 0; j < 9; ++j) s += Tq[i + 9 * j] * su[j + 9 * q];
            sv[i + 9 * q] = s * dv;
          }
        }
        std::fill(re.begin(), re.end(), 0.0);
        for (int c = 0; c < 3; ++c) {
          for (int q = 0; q < np; ++q) {
            const double* ij = &iJt[((size_t)cell * np + q) * 9];
            coef[q * 4] = 0.0;
            for (int K = 0; K < 3; ++K) {
              double s = 0;
              for (int J = 0; J < 3; ++J) s += sv[(J + 3 * c) + 9 * q] * ij[J + 3 * K];
              coef[q * 4 + 1 + K] = s;
            }
          }
          tp_contract(op, coef.data(), rc.data());
          for (int t = 0; t < nn; ++t) re[op->ldof[c * nn + op->nodeperm[t]]] += rc[t];
        }
        for (int j = 0; j < nd; ++j)
          if (ids[j] > 0) y[ids[j] - 1] += re[j];
      }
    }
  }
  return 0;
}

// A dead traction on faces of the cells (a boundary group, assembly.jl:539-578): f[dof] += sum_q dS t_c N_n(xi_q)
// over the P x P Gauss points of each face (local face 2d + s: xi_d = s), with dS = w_a w_b |dx/dxi_e1 x dx/dxi_e2|
// from the cell's trilinear map and N_n the tensor-product Lagrange function of table node n evaluated at the face
// point; free ids into f[id - 1], Dirichlet ids into f[nfree - id - 1].
int orc_face_load(orc_op* op, int64_t nfaces, const int32_t* face_cell, const int32_t* face_local, const double* traction,
                  double* f) {
  const int Q = op->Q, P = op->P, nn = op->nn;
  std::fill(f, f + op->nfree + op->ndir, 0.0);
  std::vector<double> N[3], dN[3];
  for (int k = 0; k < 3; ++k) {
    N[k].resize(Q);
    dN[k].resize(Q);
  }
  for (int64_t fc = 0; fc < nfaces; ++fc) {
    const int cell = face_cell[fc], d = face_local[fc] / 2, side = face_local[fc] % 2;
    const int e1 = d == 0 ? 1 : 0, e2 = d == 2 ? 1 : 2;
    for (int b = 0; b < P; ++b)
      for (int a = 0; a < P; ++a) {
        double xi[3];
        xi[d] = side;
        xi[e1] = op->xq1[a];
        xi[e2] = op->xq1[b];
        double Jt[9] = {0};  // Jt[K + 3i] = d x_i / d xi_K
        for (int v = 0; v < 8; ++v) {
          double m[3], dm[3];
          for (int k = 0; k < 3; ++k) {
            const int bit = (v >> k) & 1;
            m[k] = bit ? xi[k] : 1 - xi[k];
            dm[k] = bit ? 1.0 : -1.0;
          }
          const double dM[3] = {dm[0] * m[1] * m[2], m[0] * dm[1] * m[2], m[0] * m[1] * dm[2]};
          const double* x = &op->coords[((size_t)cell * 8 + v) * 3];
          for (int K = 0; K < 3; ++K)
            for (int i = 0; i < 3; ++i) Jt[K + 3 * i] += dM[K] * x[i];
        }
        double t1[3], t2[3];
        for (int i = 0; i < 3; ++i) {
          t1[i] = Jt[e1 + 3 * i];
          t2[i] = Jt[e2 + 3 * i];
        }
        const double nrm[3] = {t1[1] * t2[2] - t1[2] * t2[1], t1[2] * t2[0] - t1[0] * t2[2], t1[0] * t2[1] - t1[1] * t2[0]};
        const double dS = op->wq1[a] * op->wq1[b] * std::sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
        for (int k = 0; k < 3; ++k) orc_lagrange1d(Q, xi[k], N[k].data(), dN[k].data());
        for (int n = 0; n < nn; ++n) {
          const int t = op->tensor_of[n];
          const double Nn = N[0][t % Q] * N[1][(t / Q) % Q] * N[2][t / (Q * Q)];
          if (Nn == 0.0) continue;
          for (int c = 0; c < 3; ++c) {
            const int32_t id = op->dofs[(size_t)cell * op->nd + op->ldof[c * nn + n]];
            const double v = dS * Nn * traction[3 * fc + c];
            if (id > 0) f[id - 1] += v;
            else f[op->nfree - id - 1] += v;
          }
        }
      }
  }
  return 0;
}

// A Robin (spring) term k int u.v dGamma on faces of the cells, as a dense matrix over all dofs (free ids first, then
// the Dirichlet ids: index id - 1, or nfree - id - 1): M[i + ntot j] += k dS N_n1 N_n2 for the same component, over
// the P x P Gauss points of each face, with the tensor Lagrange functions of every table node evaluated at the point.
int orc_face_mass(orc_op* op, int64_t nfaces, const int32_t* face_cell, const int32_t* face_local, const double* k,
                  double* M) {
  const int Q = op->Q, P = op->P, nn = op->nn;
  const int64_t ntot = op->nfree + op->ndir;
  std::fill(M, M + ntot * ntot, 0.0);
  std::vector<double> N[3], dN[3], Nn(nn);
  for (int kk = 0; kk < 3; ++kk) {
    N[kk].resize(Q);
    dN[kk].resize(Q);
  }
  auto index = [&](int32_t id) { return id > 0 ? (int64_t)id - 1 : op->nfree - id - 1; };
  for (int64_t fc = 0; fc < nfaces; ++fc) {
    const int cell = face_cell[fc], d = face_local[fc] / 2, side = face_local[fc] % 2;
    const int e1 = d == 0 ? 1 : 0, e2 = d == 2 ? 1 : 2;
    for (int b = 0; b < P; ++b)
      for (int a = 0; a < P; ++a) {
        double xi[3];
        xi[d] = side;
        xi[e1] = op->xq1[a];
        xi[e2] = op->xq1[b];
        double Jt[9] = {0};
        for (int v = 0; v < 8; ++v) {
          double m[3], dm[3];
          for (int kk = 0; kk < 3; ++kk) {
            const int bit = (v >> kk) & 1;
            m[kk] = bit ? xi[kk] : 1 - xi[kk];
            dm[kk] = bit ? 1.0 : -1.0;
          }
          const double dM[3] = {dm[0] * m[1] * m[2], m[0] * dm[1] * m[2], m[0] * m[1] * dm[2]};
          const double* x = &op->coords[((size_t)cell * 8 + v) * 3];
          for (int K = 0; K < 3; ++K)
            for (int i = 0; i < 3; ++i) Jt[K + 3 * i] += dM[K] * x[i];
        }
        double t1[3], t2[3];
        for (int i = 0; i < 3; ++i) {
          t1[i] = Jt[e1 + 3 * i];
          t2[i] = Jt[e2 + 3 * i];
        }
        const double nrm[3] = {t1[1] * t2[2] - t1[2] * t2[1], t1[2] * t2[0] - t1[0] * t2[2], t1[0] * t2[1] - t1[1] * t2[0]};
        const double w = k[fc] * op->wq1[a] * op->wq1[b] *
                         std::sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
        for (int kk = 0; kk < 3; ++kk) orc_lagrange1d(Q, xi[kk], N[kk].data(), dN[kk].data());
        for (int n = 0; n < nn; ++n) {
          const int t = op->tensor_of[n];
          Nn[n] = N[0][t % Q] * N[1][(t / Q) % Q] * N[2][t / (Q * Q)];
        }
        for (int n1 = 0; n1 < nn; ++n1) {
          if (Nn[n1] == 0.0) continue;
          for (int n2 = 0; n2 < nn; ++n2) {
            if (Nn[n2] == 0.0) continue;
            for (int c = 0; c < 3; ++c) {
              const int64_t i = index(op->dofs[(size_t)cell * op->nd + op->ldof[c * nn + n1]]);
              const int64_t j = index(op->dofs[(size_t)cell * op->nd + op->ldof[c * nn + n2]]);
              M[i + ntot * j] += w * Nn[n1] * Nn[n2];
            }
          }
        }
      }
  }
  return 0;
}

// the displacement gradient at every quadrature point, G[(cell np + q) 9 + J + 3c] = d u_c / d x_J (Gridap's
// grad u[J, c]): the pointwise input of a functional's integrand (functionals.jl:9-21)
int orc_point_gradients(orc_op* op, const double* x, double* G) {
  int np = op->np, nd = op->nd;
  std::vector<double> xl(nd);
  for (int cell = 0; cell < op->ncells; ++cell) {
    gather_state(op, cell, x, xl.data());
    for (int q = 0; q < np; ++q) {
      double ij[9], dv;
      point_geometry(op, cell, q, ij, &dv);
      point_gradient(op, xl.data(), q, ij, &G[((size_t)cell * np + q) * 9]);
    }
  }
  return 0;
}

double orc_energy(orc_op* op, const double* x) {
  int np = op->np, nd = op->nd;
  double total = 0;
  for (int cell = 0; cell < op->ncells; ++cell) {
    std::vector<double> xl(nd);
    gather_state(op, cell, x, xl.data());
    for (int q = 0; q < np; ++q) {
      double ij[9], dv, g[9];
      point_geometry(op, cell, q, ij, &dv);
      point_gradient(op, xl.data(), q, ij, g);
      total += orc_law_W(op->law, op->lam, mu_at(op, cell, q), g) * dv;
      for (int c = 0; c < 3; ++c) {  // body force potential: - b . u
        double u = 0;
        for (int n = 0; n < op->nn; ++n) u += xl[op->ldof[c * op->nn + n]] * op->Nt[(size_t)q * op->nn + n];
        total -= op->body[c] * u * dv;
      }
    }
  }
  return total;
}

int orc_residual(orc_op* op, const double* x, double* b) {
  int np = op->np, nd = op->nd, nn = op->nn;
  for (int64_t i = 0; i < op->nfree; ++i) b[i] = 0.0;
  for (const auto& col : op->colors) {
    int ncol = (int)col.size();
#pragma omp parallel
    {
      std::vector<double> xl(nd), re(nd);
#pragma omp for schedule(static)
      for (int ic = 0; ic < ncol; ++ic) {
        int cell = col[ic];
        gather_state(op, cell, x, xl.data());
        std::fill(re.begin(), re.end(), 0.0);
        for (int q = 0; q < np; ++q) {
          double ij[9], dv, g[9], R[9];
          point_geometry(op, cell, q, ij, &dv);
          point_gradient(op, xl.data(), q, ij, g);
          double m = mu_at(op, cell, q);
          orc_law_P(op->law, op->lam, m, g, R);
          for (int c = 0; c < 3; ++c) {
            double r0 = -dv * op->body[c], rt[3];
            for (int K = 0; K < 3; ++K) {
              double s = 0;
              for (int J = 0; J < 3; ++J) s += ij[J + 3 * K] * R[J + 3 * c];
              rt[K] = dv * s;
            }
            for (int n = 0; n < nn; ++n) {
              const double* dN = &op->dNt[((size_t)q * nn + n) * 3];
              re[op->ldof[c * nn + n]] += op->Nt[(size_t)q * nn + n] * r0 + dN[0] * rt[0] + dN[1] * rt[1] + dN[2] * rt[2];
            }
          }
        }
        const int32_t* ids = &op->dofs[(size_t)cell * nd];
        for (int j = 0; j < nd; ++j)
          if (ids[j] > 0) b[ids[j] - 1] += re[j];
      }
    }
  }
  return 0;
}

int orc_element_jacobians(orc_op* op, const double* x, double* Ke) {
  int np = op->np, nd = op->nd, nn = op->nn;
#pragma omp parallel for schedule(static)
  for (int cell = 0; cell < op->ncells; ++cell) {
    std::vector<double> xl(nd), W(nn * 3);
    double* K = &Ke[(size_t)cell * nd * nd];
    std::fill(K, K + (size_t)nd * nd, 0.0);
    gather_state(op, cell, x, xl.data());
    for (int q = 0; q < np; ++q) {
      double ij[9], dv, g[9], Tq[81];
      point_geometry(op, cell, q, ij, &dv);
      point_gradient(op, xl.data(), q, ij, g);
      orc_law_T(op->law, op->lam, mu_at(op, cell, q), g, Tq);
      for (int n = 0; n < nn; ++n)
        for (int J = 0; J < 3; ++J) {
          double s = 0;
          for (int Kd = 0; Kd < 3; ++Kd) s += ij[J + 3 * Kd] * op->dNt[((size_t)q * nn + n) * 3 + Kd];
          W[n * 3 + J] = s;
        }
      for (int c = 0; c < 3; ++c)
        for (int n = 0; n < nn; ++n) {
          int ii = op->ldof[c * nn + n];
          for (int k = 0; k < 3; ++k)
            for (int m = 0; m < nn; ++m) {
              int jj = op->ldof[k * nn + m];
              double s = 0;
              for (int J = 0; J < 3; ++J)
                for (int L = 0; L < 3; ++L) s += W[n * 3 + J] * Tq[(J + 3 * c) + 9 * (L + 3 * k)] * W[m * 3 + L];
              K[ii + (size_t)nd * jj] += dv * s;
            }
        }
    }
  }
  return 0;
}

// assembly.jl:773-818 (full keys, single field, test == trial numbering)
int64_t orc_pattern_nnz(orc_op* op) {
  int nd = op->nd;
  std::vector<uint64_t> keys;
  keys.reserve((size_t)op->ncells * nd * nd);
  for (int cell = 0; cell < op->ncells; ++cell) {
    const int32_t* ids = &op->dofs[(size_t)cell * nd];
    for (int jj = 0; jj < nd; ++jj) {
      if (ids[jj] <= 0) continue;
      for (int ii = 0; ii < nd; ++ii)
        if (ids[ii] > 0) keys.push_back(((uint64_t)ids[jj] << 32) | (uint64_t)ids[ii]);
    }
  }
  std::sort(keys.begin(), keys.end());
  return (int64_t)(std::unique(keys.begin(), keys.end()) - keys.begin());
}

int orc_pattern(orc_op* op, int64_t* colptr, int64_t* rowval, int32_t* nzpos) {
  int nd = op->nd;
  std::vector<uint64_t> keys;
  keys.reserve((size_t)op->ncells * nd * nd);
  for (int cell = 0; cell < op->ncells; ++cell) {
    const int32_t* ids = &op->dofs[(size_t)cell * nd];
    for (int jj = 0; jj < nd; ++jj) {
      if (ids[jj] <= 0) continue;
      for (int ii = 0; ii < nd; ++ii)
        if (ids[ii] > 0) keys.push_back(((uint64_t)ids[jj] << 32) | (uint64_t)ids[ii]);
    }
  }
  std::sort(keys.begin(), keys.end());
  keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
  int64_t ncols = op->nfree;
  std::vector<int64_t> cnt(ncols, 0);
  for (size_t k = 0; k < keys.size(); ++k) {
    cnt[(keys[k] >> 32) - 1]++;
    rowval[k] = (int64_t)(keys[k] & 0xffffffffu);
  }
  colptr[0] = 1;
  for (int64_t c = 0; c < ncols; ++c) colptr[c + 1] = colptr[c] + cnt[c];
  // cell_positions! with FULL = true: searchsortedfirst in the column's rows
#pragma omp parallel for schedule(static)
  for (int cell = 0; cell < op->ncells; ++cell) {
    const int32_t* ids = &op->dofs[(size_t)cell * nd];
    int32_t* pos = &nzpos[(size_t)cell * nd * nd];
    for (int jj = 0; jj < nd; ++jj) {
      int32_t c = ids[jj];
      for (int ii = 0; ii < nd; ++ii) {
        int32_t r = ids[ii];
        if (c <= 0 || r <= 0) {
          pos[ii + nd * jj] = 0;
          continue;
        }
        int64_t lo = colptr[c - 1] - 1, hi = colptr[c] - 1;  // 0-based [lo, hi)
        const int64_t* it = std::lower_bound(rowval + lo, rowval + hi, (int64_t)r);
        pos[ii + nd * jj] = (int32_t)(it - rowval + 1);
      }
    }
  }
  return 0;
}

int orc_jacobian(orc_op* op, const double* x, const int32_t* nzpos, double* nzval, int64_t nnz) {
  int nd = op->nd;
  std::vector<double> Ke((size_t)op->ncells * nd * nd);
  orc_element_jacobians(op, x, Ke.data());
  for (int64_t i = 0; i < nnz; ++i) nzval[i] = 0.0;
  for (const auto& col : op->colors)
    for (int32_t cell : col) {
      const int32_t* pos = &nzpos[(size_t)cell * nd * nd];
      const double* K = &Ke[(size_t)cell * nd * nd];
      for (int jj = 0; jj < nd; ++jj)
        for (int ii = 0; ii < nd; ++ii) {
          int32_t p = pos[ii + nd * jj];
          if (p > 0) nzval[p - 1] += K[ii + (size_t)nd * jj];
        }
    }
  return 0;
}

// preconditioners.jl:151-171 (field_blocks): one block per node with a free component, first-seen order
int orc_node_blocks(orc_op* op, int32_t* block_of, int32_t* block_dofs, int64_t cap, int64_t* nblocks) {
  int nn = op->nn, nd = op->nd;
  std::map<std::tuple<int32_t, int32_t, int32_t>, int32_t> keys;
  int64_t nb = 0;
  for (int cell = 0; cell < op->ncells; ++cell) {
    const int32_t* ids = &op->dofs[(size_t)cell * nd];
    for (int n = 0; n < nn; ++n) {
      int32_t k0 = ids[op->ldof[0 * nn + n]], k1 = ids[op->ldof[1 * nn + n]], k2 = ids[op->ldof[2 * nn + n]];
      if (k0 <= 0 && k1 <= 0 && k2 <= 0) {
        block_of[(size_t)cell * nn + n] = -1;
        continue;
      }
      auto key = std::make_tuple(k0, k1, k2);
      auto it = keys.find(key);
      int32_t blk;
      if (it == keys.end()) {
        if (nb >= cap) return fail("oracle: block capacity exceeded");
        blk = (int32_t)nb++;
        keys[key] = blk;
        block_dofs[3 * blk + 0] = k0 > 0 ? k0 : 0;
        block_dofs[3 * blk + 1] = k1 > 0 ? k1 : 0;
        block_dofs[3 * blk + 2] = k2 > 0 ? k2 : 0;
      } else {
        blk = it->second;
      }
      block_of[(size_t)cell * nn + n] = blk;
    }
  }
  *nblocks = nb;
  return 0;
}

// preconditioners.jl:234-263 (_blocks_from_tangents!) then 292-307 (_invert_blocks!)
int orc_blocks_from_tangents(orc_op* op, const double* T, const double* iJt, const double* dV, const int32_t* block_of,
                             const int32_t* block_dofs, int64_t nblocks, double* Binv) {
  int np = op->np, nn = op->nn;
  std::vector<double> B((size_t)nblocks * 9, 0.0), W(nn * 3);
  for (const auto& col : op->colors)
    for (int32_t cell : col)
      for (int q = 0; q < np; ++q) {
        const double* ij = &iJt[((size_t)cell * np + q) * 9];
        const double* Tq = &T[((size_t)cell * np + q) * 81];
        double dv = dV[(size_t)cell * np + q];
        for (int n = 0; n < nn; ++n)
          for (int J = 0; J < 3; ++J) {
            double s = 0;
            for (int K = 0; K < 3; ++K) s += ij[J + 3 * K] * op->dNt[((size_t)q * nn + n) * 3 + K];
            W[n * 3 + J] = s;
          }
        for (int n = 0; n < nn; ++n) {
          int32_t blk = block_of[(size_t)cell * nn + n];
          if (blk < 0) continue;
          for (int cj = 0; cj < 3; ++cj)
            for (int ci = 0; ci < 3; ++ci) {
              double s = 0;
              for (int b = 0; b < 3; ++b) {
                double wu = W[n * 3 + b];
                int ju = b + 3 * cj;
                for (int a = 0; a < 3; ++a) s += W[n * 3 + a] * Tq[(a + 3 * ci) + 9 * ju] * wu;
              }
              B[(size_t)blk * 9 + ci + 3 * cj] += dv * s;
            }
        }
      }
  for (int64_t blk = 0; blk < nblocks; ++blk) {
    double M[9];
    for (int j = 0; j < 3; ++j)
      for (int i = 0; i < 3; ++i)
        M[i + 3 * j] = (block_dofs[3 * blk + i] == 0 || block_dofs[3 * blk + j] == 0) ? (i == j ? 1.0 : 0.0)
                                                                                      : B[blk * 9 + i + 3 * j];
    bool finite = true;
    for (double v : M) finite &= std::isfinite(v);
    if (finite)
      inv3(M, &Binv[blk * 9]);
    else
      std::memcpy(&Binv[blk * 9], M, sizeof(M));
  }
  return 0;
}

// preconditioners.jl:327-342
void orc_apply_blocks(const double* Binv, const int32_t* block_dofs, int64_t nblocks, const double* r, double* y) {
  for (int64_t blk = 0; blk < nblocks; ++blk)
    for (int i = 0; i < 3; ++i) {
      int32_t id = block_dofs[3 * blk + i];
      if (id == 0) continue;
      double s = 0;
      for (int j = 0; j < 3; ++j) {
        int32_t jd = block_dofs[3 * blk + j];
        if (jd != 0) s += Binv[blk * 9 + i + 3 * j] * r[jd - 1];
      }
      y[id - 1] = s;
    }
}

}  // extern "C"
