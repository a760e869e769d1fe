// TEST INFRASTRUCTURE ONLY — the oracle.
//
// CPU restatement, in FP64, of the HyperFEMKernels.jl (v0.2.6) algorithms on the matrix-free and assembly
// hot path. It is the checker of the HIP library: only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it, and never as the thing measured or shipped.
//
// Each function cites the reference file:line it restates. Conventions shared with the C-ABI (include/hfk.h):
//   * cell_dofs[cell*nd + j]   Gridap's signed ids: >0 free (1-based), <0 Dirichlet (-1-based)
//   * ldof[c*nn + n]           local dof of table node n, component c (FieldTable.ldof, tables.jl:171-237)
//   * nodeperm[t]              tensor node t = i + Q j + Q^2 k -> table node (TPTable.nodeperm, tensorproduct.jl:13-91)
//   * coords[(cell*8 + v)*3+i] Q1 geometry vertices in tensor order v = i + 2 j + 4 k
//   * points in tensor order pl = a + P b + P^2 c (pointperm = identity)
//   * slots s = J + 3 c  <->  grad(u)[J,c] = du_c/dx_J (codegen.jl:62-82); F_cJ = delta_cJ + grad(u)[J,c]
//   * tangents T[sv + 9 su] = dP_{c J}/dF_{k L} with sv = J + 3c, su = L + 3k (matrixfree.jl:15)
#pragma once
#include <cstdint>

extern "C" {

// --- tables (tensorproduct.jl:24-47; Gauss-Legendre on [0,1]) ---
void orc_gauss01(int P, double* x, double* w);
void orc_lagrange1d(int Q, double xi, double* N, double* dN);

// --- pointwise laws: 0 = St Venant-Kirchhoff (test/common.jl:2-8), 1 = compressible neo-Hookean
//     W = mu/2 (I1-3) - mu lnJ + lam/2 lnJ^2, 2 = linear elasticity (test_pmultigrid.jl:8) ---
void orc_law_P(int law, double lam, double mu, const double* g /*9 slots*/, double* R /*9 slots*/);
void orc_law_T(int law, double lam, double mu, const double* g, double* T /*81*/);
double orc_law_W(int law, double lam, double mu, const double* g);  // stored energy density W(F)

// --- operator handle ---
typedef struct orc_op orc_op;
orc_op* orc_op_create(int Q, int P, int ncells, const int32_t* cell_dofs, const int32_t* nodeperm,
                      const int32_t* ldof, const double* coords, int64_t nfree, int64_t ndir, int law,
                      double lam, double mu, const double* body /*3 or null*/);
void orc_op_destroy(orc_op*);
void orc_op_set_mu_field(orc_op*, const double* mu /*ncells*np, tensor point order; null = constant*/);
void orc_op_set_dirichlet(orc_op*, const double* dvals);
int orc_op_ncolors(orc_op*);
int orc_op_colors(orc_op*, int32_t* color_of_cell /*ncells, 1-based*/);
const char* orc_last_error(void);

// prepare (matrixfree.jl:140-159,199-228): T (81*np*ncells, reference layout T[i, q, cell]), iJt (9*np*ncells), dV (np*ncells)
int orc_prepare(orc_op*, const double* x, double* T, double* iJt, double* dV);
// matrix-free product with stored tangents, sum factorised, colour by colour (matrixfree.jl:246-260,341-424,496-520)
int orc_jv(orc_op*, const double* T, const double* iJt, const double* dV, const double* dx, double* y, int nthreads);
// total stored energy sum_cells sum_q W(F) dV at state x (the potential whose gradient is the residual)
double orc_energy(orc_op*, const double* x);
// residual (assembly.jl:909-916,1219-1241; kernels.jl:319-356)
int orc_residual(orc_op*, const double* x, double* b);
// dense element Jacobians Ke[cell][ii + nd*jj] (test ii, trial jj) with dense ND tables (kernels.jl:568-599)
int orc_element_jacobians(orc_op*, const double* x, double* Ke);
// sparsity (assembly.jl:773-899): size query, then colptr (ncols+1), rowval (nnz), nzpos (nd*nd*ncells), all 1-based
int64_t orc_pattern_nnz(orc_op*);
int orc_pattern(orc_op*, int64_t* colptr, int64_t* rowval, int32_t* nzpos);
// assembled values in colour order through nzpos (assembly.jl:918-927,1402-1408)
int orc_jacobian(orc_op*, const double* x, const int32_t* nzpos, double* nzval, int64_t nnz);

// --- block Jacobi (preconditioners.jl:151-342) on prepared tangents ---
// blocks per global node: node_dofs (nnodes*3, 1-based free ids, 0 = fixed) and per (cell, table node) the block id (0-based, -1 none)
int orc_node_blocks(orc_op*, int32_t* block_of /*ncells*nn*/, int32_t* block_dofs /*cap*3*/, int64_t cap, int64_t* nblocks);
int orc_blocks_from_tangents(orc_op*, const double* T, const double* iJt, const double* dV, const int32_t* block_of,
                             const int32_t* block_dofs, int64_t nblocks, double* Binv /*nblocks*9*/);
void orc_apply_blocks(const double* Binv, const int32_t* block_dofs, int64_t nblocks, const double* r, double* y);
}
