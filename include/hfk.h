/* libhfk — MI355X (gfx950) kernels for the matrix-free and assembly hot path of HyperFEMKernels.jl.
 *
 * C ABI, bound by `ccall` from the Julia host (hyperfemkernels.jl_amd/julia/) and by ctypes from the Python
 * mirror (hyperfemkernels.jl_amd/hfk/).  The reference is pure Julia and has no FFI: each entry point below
 * replaces a Julia method of the reference, cited as file:line of HyperFEMKernels.jl v0.2.6.
 *
 * Conventions
 *   - Every function returns HFK_OK (0) or a negative status; hfk_last_error() gives the message.  Misuse
 *     maps to the reference's exceptions (ArgumentError -> HFK_EARG, DimensionMismatch -> HFK_EDIM,
 *     ErrorException -> HFK_ESTATE).  Numerical failure is never an error status (solver calls report it
 *     through their status outputs, as the reference's `converged`/`indefinite` flags).
 *   - Host pointers are borrowed for the duration of the call.  Device memory is owned by the handles and
 *     freed by *_destroy / hfk_close.  `*_dev` variants take device pointers and a hipStream_t (NULL: the
 *     context's stream) and return without synchronising.
 *   - Ids follow Gridap's cell dof ids: cell_dofs[cell*nd + j] > 0 is a free dof (1-based), < 0 a Dirichlet
 *     value (-1-based).  Table indices (ldof, nodeperm) are 0-based.  CSC arrays (colptr, rowval) and the
 *     positions nzpos are 1-based as in Julia's SparseMatrixCSC, nzpos = 0 meaning "not stored".
 *   - A context is bound to one device and one stream; neither a context nor an operator is reentrant
 *     (src/assembly.jl:363-367).
 */
#ifndef HFK_H
#define HFK_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HFK_OK 0
#define HFK_EARG (-1)   /* ArgumentError */
#define HFK_EDIM (-2)   /* DimensionMismatch */
#define HFK_ESTATE (-3) /* ErrorException: call out of order, unsupported configuration */
#define HFK_EHIP (-4)   /* HIP runtime failure */
#define HFK_EALLOC (-5) /* device or host allocation failure */

/* Pointwise laws of the built-in forms  ∫ ∇(v)' ⊙ P(F(∇u')) dΩ  with F = I + ∇u' (README.md:53-67,
 * test/common.jl:2-8): tangent T[(J,c),(L,k)] = ∂P_cJ/∂F_kL (src/factorforms.jl:317-322). */
#define HFK_LAW_SVK 0         /* St Venant-Kirchhoff, test/common.jl:2-8 */
#define HFK_LAW_NEOHOOKE 1    /* W = mu/2 (I1-3) - mu lnJ + lam/2 (lnJ)^2 */
#define HFK_LAW_LINEAR 2      /* sigma = lam tr(eps) I + 2 mu eps on G = F - I */

/* Storage of the prepared point tangents (src/matrixfree.jl:14-23 stores 81 FP64 per point). */
#define HFK_STORE_FULL64 0    /* 81 FP64 per point (the reference's layout, in reference coordinates) */
#define HFK_STORE_SYM64 1     /* 45 FP64 per point: the symmetric part of a tangent that is symmetric */
#define HFK_STORE_FULL32 2    /* 81 FP32 per point (mixed precision) */
#define HFK_STORE_SYM32 3     /* 45 FP32 per point */
#define HFK_STORE_FAC64 4     /* 17 FP64 per point: the factors of an isotropic tangent mu d_ck d_JL + m A_Jk A_Lc +
                                 lam A_Jc A_Lk (built-in neo-Hookean, A = F^-1; built-in linear, A = I), from which
                                 the product rebuilds D su in registers (no counterpart in the reference, which stores
                                 the 81 entries) */
#define HFK_STORE_FAC32 5     /* the same 17 factors in FP32 (mixed precision) */
#define HFK_STORE_AFAC64 6    /* the factorised tangent on affine cells (parallelepipeds; refused otherwise): 11 FP64
                                 per point (iJt' F^-1, dV m, dV mu) and 8 per cell (the metric iJt' iJt and det J, the
                                 per-cell geometry of the reference's Jv byte model) */
#define HFK_STORE_AFAC32 7    /* the same 11 per point in FP32 */

typedef struct hfk_ctx hfk_ctx;
typedef struct hfk_op hfk_op;

/* ---- context ----------------------------------------------------------------------------------- */
int hfk_version(void);
/* Bind a context to HIP device `device` with its own stream (default priority). */
int hfk_open(int device, hfk_ctx** ctx);
int hfk_close(hfk_ctx* ctx);
/* Message of the last failing call on `ctx` (or of this thread when ctx is NULL). */
const char* hfk_last_error(const hfk_ctx* ctx);
int hfk_synchronize(hfk_ctx* ctx);

/* ---- operator: KernelOperator(res, jac, U, V; backend = :mi355x)  (src/assembly.jl:630-666) ----- */
/* One volume group of a vector Q_{Q-1} Lagrangian field on hexahedra with P^3 tensor Gauss points.
 *   cell_dofs  nd*ncells, nd = 3*Q^3, Gridap order (get_cell_dof_ids, src/tables.jl:329-342)
 *   ldof       3*Q^3: ldof[c*nn + n] = local dof of table node n, component c (src/tables.jl:171-237)
 *   nodeperm   Q^3:   tensor node i + Q j + Q^2 k -> table node (src/tensorproduct.jl:13-91)
 *   coords     ncells*8*3 vertices of the trilinear cell map, v = i + 2j + 4k (src/assembly.jl:1073-1111)
 * Colours the cells with the reference's greedy first fit (src/coloring.jl:4-28). */
int hfk_op_create(hfk_ctx* ctx, int Q, int P, int64_t ncells, const int32_t* cell_dofs, const int32_t* ldof,
                  const int32_t* nodeperm, const double* coords, int64_t nfree, int64_t ndir, int law,
                  double lam, double mu, hfk_op** op);
int hfk_op_destroy(hfk_op* op);
/* Declare `fine` a nested refinement of `coarse`'s mesh (no counterpart in the reference, whose p-multigrid keeps
 * one mesh): cell f of `fine` lies in cell parent[f] of `coarse` as its child child[f] = cx + 2 cy + 4 cz (cx, cy,
 * cz in {0, 1}: the half of the coarse reference cell along each direction).  A p-multigrid whose consecutive levels
 * are such a pair transfers between them geometrically (trilinear prolongation, its transpose, nodal injection of
 * the state): an h-hierarchy below the p-levels.  coarse = NULL removes the declaration. */
int hfk_op_set_refinement(hfk_op* fine, hfk_op* coarse, const int32_t* parent, const int32_t* child);
/* Dirichlet values, re-read by every assembly (refresh_dirichlet_values!, src/assembly.jl:393-410). */
int hfk_op_set_dirichlet(hfk_op* op, const double* dvals);
/* The numbers of the form, rebound without rebuilding anything (the fast path of update(op, res, jac) for a capture
 * of the same structure, src/assembly.jl:693-744 and src/tree.jl:270-283: the colouring, the pattern, the tables and
 * compiled laws, which take lam and mu as arguments, are kept).  The stored tangents are invalidated (prepare! again). */
int hfk_op_set_parameters(hfk_op* op, double lam, double mu);
/* A boundary traction, the Neumann load  -Λ ∫ t·v dΓ  of a boundary group (src/assembly.jl:539-578; the load term of
 * test/solvers/test_functionals.jl:40-45), on nfaces faces of the cells: face_cell (0-based cell), face_local (0..5:
 * the reference-cube face xi_d = s is 2d + s, vertex order v = i + 2j + 4k), traction nfaces x 3 (a dead load per
 * unit area of the physical face), integrated with the operator's P x P Gauss rule on the face.  residual! subtracts
 * Λ f on the free rows and the energy subtracts Λ (f·x + f_D·u_D); the Jacobian does not change.  The load vector f
 * is computed on the device when the faces are set; nfaces = 0 removes it. */
int hfk_op_set_traction(hfk_op* op, int64_t nfaces, const int32_t* face_cell, const int32_t* face_local,
                        const double* traction);
/* A Robin (spring) term  k ∫ u·v dΓ  on nfaces faces of the cells (a boundary group whose bilinear form enters the
 * Jacobian, src/assembly.jl:539-578; the Robin outlet term ZA (u·n)(v·n) of test/common.jl:212-213 is its normal
 * part): faces as in hfk_op_set_traction, stiffness k per face.  The face mass matrices are computed on the device
 * when the faces are set; residual! adds k M u (Dirichlet values lifted), Jv, the assembled Jacobian and the node
 * blocks add k M, and the energy adds k/2 u·M u.  The stored tangents are invalidated (prepare! again); nfaces = 0
 * removes the term. */
int hfk_op_set_robin(hfk_op* op, int64_t nfaces, const int32_t* face_cell, const int32_t* face_local,
                     const double* stiffness);
/* Skeleton integrals (SkeletonTriangulation terms, src/skeleton.jl; the facet kernels of cells!, :291-330, the cross
 * blocks of skeleton_keys!, :355-390): interior facets, each with a plus cell and a minus cell (0-based), the local
 * face of each (0..5: face xi_d = side is 2 d + side) and the orientation o (0..7) that maps the plus face's
 * coordinates (s, t) along its tangential directions (in increasing order) to the minus face's:
 * (s', t') = o & 1 ? (t, s) : (s, t), then s' -> 1 - s' if o & 2, t' -> 1 - t' if o & 4.  The faces must coincide
 * (corners, and the dof ids of every face node) and each (cell, face) may appear once.  source/name: a facet
 * point form (a struct with static residual(s, x, n, h, p, R), jvp(s, ds, x, n, h, p, dR) and energy(s, x, n, h, p)
 * of the 24 facet slots s, the point x, the normal n, h and p; see hfk_skel.hpp); source NULL:
 * the built-in penalty on the jump of the displacement gradient, energy (gamma h / 2) |grad u+ - grad u-|^2 with
 * gamma = params[0] and h = sqrt(facet area).  params: up to 8 numbers.  residual!, mul!, jacobian!, the energy and
 * the sparsity pattern (which gains the cross blocks of every facet) then include the facets; BlockJacobi and the
 * p-multigrid refuse such an operator, as the reference's do.  nfacets = 0 removes the skeleton. */
int hfk_op_set_skeleton(hfk_op* op, int64_t nfacets, const int32_t* plus_cell, const int32_t* plus_face,
                        const int32_t* minus_cell, const int32_t* minus_face, const int32_t* orient,
                        const char* source, const char* name, const double* params, int nparams);
/* the colour (1-based) of every facet of the skeleton (color_facets, src/skeleton.jl:39-69) */
int hfk_op_skeleton_colors(hfk_op* op, int32_t* color);
/* A boundary group with a captured integrand (src/assembly.jl:539-578: the boundary classes of a
 * BoundaryTriangulation): faces of the cells (cell 0-based, local face 0..5, each pair once) and a point form of the
 * 12 slots of the cell at the face point (u, then du_c/dx_K at 3 + 3c + K), the point x, the outward unit normal n
 * of the reference configuration, h = sqrt(face area) and params (static residual, jvp and energy as for the
 * skeleton).  source NULL: the built-in follower pressure p = params[0] on the deformed face, nominal traction
 * -p cof(F) N (not conservative: no energy).  residual!, mul!, jacobian! and the energy include it; BlockJacobi and
 * the p-multigrid refuse such an operator.  nfaces = 0 removes it. */
int hfk_op_set_boundary_form(hfk_op* op, int64_t nfaces, const int32_t* face_cell, const int32_t* face_local,
                             const char* source, const char* name, const double* params, int nparams);
/* Compile the face kernels of a form with `sides` sides (2: skeleton, 1: boundary; source NULL: the built-in one)
 * for (Q, P) without a device: HFK_OK, or HFK_EARG with the compiler's log in `log` (emit!'s checks at capture time,
 * src/codegen.jl:166-198). */
int hfk_skel_form_check(int Q, int P, int sides, const char* source, const char* name, char* log, int64_t log_size);
/* The load factor Λ of the traction (1 by default; a number of the form, rebound without recomputing f). */
int hfk_op_set_load_factor(hfk_op* op, double load_factor);
/* The traction load vector f (unscaled): nfree values on the free ids, then ndir on the Dirichlet ids. */
int hfk_op_traction_load(hfk_op* op, double* f);
/* Pointwise shear modulus mu(x) at the quadrature points, np*ncells in tensor point order, [cell][q]
 * (dense auxiliary point data; NULL restores the constant mu). */
int hfk_op_set_mu_field(hfk_op* op, const double* mu_points);
/* A function of the physical coordinates materialised on the device at the quadrature points as the pointwise
 * shear modulus (the function-of-x leaves of src/coordfunctions.jl:150-213, materialize_coord!: M[q, cell] =
 * f(X[q, cell]) at the physical points X of the cell's quadrature points).  `source` (HIP) defines
 *   __device__ double name(const double* x, const double* p);     x[3] the physical point, p the parameters
 * compiled with hiprtc for gfx950 (cached per process by the source text, like the laws); params: nparams doubles
 * copied to the device (nparams may be 0).  The points are the trilinear images of the tensor Gauss points.  Replaces
 * a field of hfk_op_set_mu_field; compile errors return HFK_EARG with the compiler log. */
int hfk_op_set_mu_function(hfk_op* op, const char* source, const char* name, const double* params, int64_t nparams);
/* The operator's shear modulus at the points, np*ncells, [cell][q] (the constant mu without a field). */
int hfk_op_mu_field(hfk_op* op, double* mu_points);
/* info[0..7] = nfree, ndir, ncells, nd, np, ncolors, tangent bytes of the current storage, nnz (or -1). */
int hfk_op_info(hfk_op* op, int64_t* info);
/* Colour of each cell, 1-based (the reference's `color[cell]`, src/coloring.jl:4-28). */
int hfk_op_colors(hfk_op* op, int32_t* color_of_cell);
/* Matrix-free products colour by colour (1, the default: the reference's summation order, bitwise
 * reproducible, src/matrixfree.jl:246-260) or in one launch with FP64 atomic adds (0: faster, the order of the
 * additions into one entry varies between runs). */
int hfk_op_set_deterministic(hfk_op* op, int deterministic);

/* ---- pointwise laws compiled at run time (src/codegen.jl:166-198 PointForm, compiled per structural
 * signature, src/tree.jl:244-283) ------------------------------------------------------------------ */
/* Compile a constitutive law given as HIP source with hiprtc for gfx950: `source` defines a struct `name` with
 *   __device__ void setup(const double* F, double lam, double mu);   F[c + 3J] = d x_c / d X_J
 *   __device__ double stress(int c, int J) const;                     first Piola-Kirchhoff P_cJ
 *   __device__ double tangent(int c, int J, int k, int L) const;     d P_cJ / d F_kL
 * (helpers hfk_det3(a) and hfk_inv3(a, r, det) on column-major 3x3 arrays are in scope).  Code objects are
 * cached per process by a hash of the full source: a second law of the same text is not compiled again.
 * Compile errors return HFK_EARG with the compiler log in hfk_last_error. */
typedef struct hfk_law hfk_law;
int hfk_law_create(hfk_ctx* ctx, const char* source, const char* name, hfk_law** law);
int hfk_law_destroy(hfk_law* law);
/* Compile only (no device): HFK_OK, or HFK_EARG with the compiler log (also copied into log[0..log_size)). */
int hfk_law_check(const char* source, const char* name, char* log, int64_t log_size);
/* key: the source hash; from_cache: 1 when the code object came from the cache */
int hfk_law_info(hfk_law* law, uint64_t* key, int* from_cache);
/* Evaluate the operator's residual!, prepare! and jacobian! with `law` (NULL: the built-in law of
 * hfk_op_create).  The law must outlive its use by the operator; the stored tangents are invalidated.  The
 * packed symmetric storages assume a symmetric tangent, as for the built-in laws. */
int hfk_op_set_law(hfk_op* op, hfk_law* law);

/* ---- residual: Algebra.residual!(b, op, x)  (src/assembly.jl:909-916, 1219-1241) ----------------- */
int hfk_residual(hfk_op* op, const double* x, double* b);
int hfk_residual_dev(hfk_op* op, const double* x_dev, double* b_dev, void* stream);

/* ---- scalar functionals: KernelFunctional(f, U; reduce)  (src/functionals.jl:9-154) --------------------------- */
/* The strain energy sum_cells sum_q W(F(x)) dV of the operator's law at the free values x (Dirichlet values lifted):
 * KernelFunctional(u -> ∫(Ψ ∘ F(∇u))dΩ, U) with reduce = :sum, the energy of a Newton line search.  Built-in laws, or
 * a compiled law that defines `__device__ double energy() const` (W at the F of setup; HFK_EARG otherwise).  Summed
 * per batch and then over the batches in a fixed order (bitwise reproducible).  _dev: device pointers, asynchronous
 * on `stream`, the energy written to energy_dev[0]. */
int hfk_energy(hfk_op* op, const double* x, double* energy);
int hfk_energy_dev(hfk_op* op, const double* x_dev, double* energy_dev, void* stream);
/* hfk_energy_dev without any allocation (no counterpart in the reference): for callers that cannot allocate device
 * memory (code under HIP graph capture, tight line-search loops, harnesses that forbid hipMalloc).  Same kernels and
 * summation order as hfk_energy_dev, so the same value bit for bit; the partial sums go to a static 1 MiB scratch
 * array per device (set up by hfk_open) instead of a buffer of the operator.  Limit: batches + 512 (traction load)
 * + Robin faces <= 131,072 partial sums (Q2: about 1.17M cells), else HFK_EDIM.  A compiled law must have been
 * used once by the operator (residual, prepare or energy) so that its kernels are loaded, else HFK_ESTATE.  Every
 * call on one device shares that scratch: calls that overlap on one device (two streams, two threads) race. */
int hfk_energy_noalloc_dev(hfk_op* op, const double* x_dev, double* energy_dev, void* stream);
/* A pointwise function of the displacement gradient and the gradient of an increment, compiled with hiprtc:
 * `source` defines  __device__ double name(const double* Gu, const double* Gdu, const double* p);
 * G[c + 3J] = d u_c / d x_J (Gridap's ∇(u)'), p the parameters.  Cached per process by the source text. */
typedef struct hfk_pointfn hfk_pointfn;
int hfk_pointfn_create(hfk_ctx* ctx, const char* source, const char* name, hfk_pointfn** fn);
int hfk_pointfn_destroy(hfk_pointfn* fn);
/* KernelFunctional(u -> ∫(g ∘ (∇(u)', ∇(duh)'))dΩ, U; reduce = :min / :max) (src/functionals.jl:12-21; the
 * injectivity step of a line search): the minimum (reduce 0) or maximum (1) of g over every quadrature point, no
 * weights, with u the free values x and the operator's Dirichlet values and du the free values dx with the Dirichlet
 * values dx_dir (ndir; NULL: zero).  params: nparams doubles (p).  NaN at any point gives NaN. */
int hfk_extremum(hfk_op* op, hfk_pointfn* fn, const double* x, const double* dx, const double* dx_dir,
                 const double* params, int64_t nparams, int reduce, double* value);

/* ---- matrix-free Jacobian (src/matrixfree.jl) ---------------------------------------------------- */
/* prepare!(mf, x)  (src/matrixfree.jl:140-159, 199-228): point tangents at state x, stored per `storage`. */
int hfk_mf_prepare(hfk_op* op, const double* x, int storage);
int hfk_mf_prepare_dev(hfk_op* op, const double* x_dev, int storage, void* stream);
/* mul!(y, mf, dx)  (src/matrixfree.jl:246-260, 341-424, 496-520): y = J(x) dx on free dofs. */
int hfk_mf_apply(hfk_op* op, const double* dx, double* y);
int hfk_mf_apply_dev(hfk_op* op, const double* dx_dev, double* y_dev, void* stream);
/* The prepared tangent in reference coordinates, D = dV iJt' T iJt per point, [cell][q][i + 9 j] (81 per
 * point, expanded from the stored form).  Inspection / tests. */
int hfk_mf_tangents(hfk_op* op, double* D);

/* ---- assembled Jacobian (src/assembly.jl:750-899, 909-927, 1316-1408) ---------------------------- */
/* sparsity_pattern / nonzero_positions: nnz first, then colptr (nfree+1), rowval (nnz), nzpos
 * (nd*nd*ncells, [cell][ii + nd*jj]).  Columns are trial dofs, rows test dofs. */
int hfk_pattern_nnz(hfk_op* op, int64_t* nnz);
int hfk_pattern(hfk_op* op, int64_t* colptr, int64_t* rowval, int32_t* nzpos);
/* jacobian!(A, op, x): nonzeros of the CSC matrix, summed colour by colour through nzpos. */
int hfk_jacobian(hfk_op* op, const double* x, double* nzval);
int hfk_jacobian_dev(hfk_op* op, const double* x_dev, double* nzval_dev, void* stream);

/* ---- BlockJacobi (src/preconditioners.jl:100-342) ---------------------------------------------------- */
/* preconditioner_setup/update!: nodal 3x3 blocks of the prepared tangent (block_diagonal! route :tangents),
 * inverted (fixed components -> identity).  ok = 0 when an inverted block is not finite (non-strict mode). */
int hfk_block_jacobi_update(hfk_op* op, int* ok);
/* ldiv!(y, blocks, r) (free rows of every block written) */
int hfk_block_jacobi_apply(hfk_op* op, const double* r, double* y);
int hfk_block_jacobi_apply_dev(hfk_op* op, const double* r_dev, double* y_dev, void* stream);
/* the block numbering (first-seen order over cells and table nodes): nblocks, then 3*nblocks free ids (0 = fixed
 * component) and the 9*nblocks inverted blocks [i + 3j] of the last update (either output may be NULL) */
int hfk_block_jacobi_blocks(hfk_op* op, int64_t* nblocks, int32_t* dofs, double* Binv);

/* ---- Chebyshev smoother (src/smoothers.jl) ----------------------------------------------------------- */
typedef struct hfk_cheb hfk_cheb;
/* Chebyshev(degree, range, inflate, eig_iters) on the block-preconditioned tangent of op */
int hfk_cheb_create(hfk_op* op, int degree, double range, double inflate, int eig_iters, hfk_cheb** out);
int hfk_cheb_destroy(hfk_cheb* ch);
/* smoother_update!: blocks and interval at op's current tangents.  status[0] = ok, status[1] = indefinite;
 * interval = (lambda_min, lambda_max, a, b) */
int hfk_cheb_update(hfk_cheb* ch, int* status, double* interval);
/* smooth!(x, st, rhs; zero_guess, residual) (src/smoothers.jl:116-146); residual may be NULL */
int hfk_cheb_smooth(hfk_cheb* ch, double* x, const double* rhs, int zero_guess, double* residual);
int hfk_cheb_smooth_dev(hfk_cheb* ch, double* x_dev, const double* rhs_dev, int zero_guess, double* residual_dev,
                        void* stream);
/* eigenvalue_bounds(mf, blocks; k, b = chebyshev_start) (src/preconditioners.jl:374-389) */
int hfk_eigenvalue_bounds(hfk_op* op, int k, double* lambda_min, double* lambda_max);

/* ---- transfers between nested Q_k spaces on one mesh (src/transfers.jl) ------------------------------- */
typedef struct hfk_transfer hfk_transfer;
/* Prolongation(Uc, Uf): both operators on the same cells; errors when the Dirichlet sets differ */
int hfk_transfer_create(hfk_op* coarse, hfk_op* fine, hfk_transfer** out);
int hfk_transfer_destroy(hfk_transfer* T);
int hfk_prolongate(hfk_transfer* T, const double* xc, double* yf);  /* prolongate! */
int hfk_restrict(hfk_transfer* T, const double* yf, double* yc);    /* restrict! */
int hfk_inject(hfk_transfer* T, const double* xf, const double* df, double* xc, double* dc);  /* inject! */

/* ---- p-multigrid (src/pmultigrid.jl) ------------------------------------------------------------------ */
typedef struct hfk_pmg hfk_pmg;
/* PMultigrid(spaces; smoother = Chebyshev(...), coarse): levels[0] is the Newton's operator (prepared with
 * hfk_mf_prepare*), levels[1..n-2] matrix-free, levels[n-1] assembled.  coarse_kind 0: block-Jacobi PCG on the
 * assembled coarse matrix to coarse_rtol (at most coarse_maxiter iterations; a variable cycle, so the outer
 * Krylov method should be the flexible CG); coarse_kind 1: a direct solve (the reference's LUSolver): Cholesky of
 * the assembled coarse matrix at every coarse set-up, at most 4000 coarse dofs (the bottom of an h-hierarchy, see
 * hfk_op_set_refinement), a fixed cycle (CG and GMRES apply).  The coarse matrix must be symmetric.  Consecutive
 * levels on one mesh are p-levels; a level declared a refinement of the next one is an h-level.  smoother_storage (an
 * HFK_STORE_* value, or -1 for the working operator's own tangents) selects the tangents every smoother and
 * matrix-free coarse level applies: HFK_STORE_SYM32 gives FP32 level applies and smoothers while the outer
 * Krylov product keeps the working operator's FP64 tangents (mixed precision). */
int hfk_pmg_create(hfk_op** levels, int nlevels, int degree, double range, double inflate, int eig_iters,
                   int smoother_storage, int coarse_kind, double coarse_rtol, int coarse_maxiter, hfk_pmg** out);
int hfk_pmg_destroy(hfk_pmg* g);
/* update_levels! + smoother_update!: inject the state of levels[0] (its last prepare), re-prepare / re-assemble
 * the levels, blocks and intervals.  status[0] = ok (preconditioner_ok), status[1] = indefinite */
int hfk_pmg_update(hfk_pmg* g, int* status);
/* PMultigrid(...; coarse_every) (src/pmultigrid.jl:229-262, 200-216): the assembled coarse level is re-assembled and
 * set up on the first update and then every coarse_every updates (its state is injected at every update; a lagged
 * coarse matrix only changes the preconditioner, the outer tolerance still fixes the answer).  Default 1. */
int hfk_pmg_set_coarse_every(hfk_pmg* g, int coarse_every);
/* number of coarse set-ups so far (the reference's nbuild) */
int hfk_pmg_coarse_builds(hfk_pmg* g, int64_t* builds);
/* The products of the coarse block-Jacobi PCG on FP32 copies of the coarse matrix values (bits = 32; vectors,
 * dot products and accumulation stay FP64) or on the FP64 values (64, the default).  No counterpart in the reference
 * (its coarse solver is LU or AMG); set before the first update.  The network cycle's coarse loop keeps FP64. */
int hfk_pmg_set_coarse_precision(hfk_pmg* g, int bits);
/* ldiv!: one V-cycle z = M^-1 r (NaN-filled when not ok) */
int hfk_pmg_vcycle(hfk_pmg* g, const double* r, double* z);
int hfk_pmg_vcycle_dev(hfk_pmg* g, const double* r_dev, double* z_dev, void* stream);
/* info[0..5] = nlevels, ok, cycles, inner coarse iterations of the last cycle, their total, coarse level ok;
 * then per smoothed level (lambda_min, lambda_max, a, b, smoother ok, blocks ok): 6 + 6 (nlevels - 1) values */
int hfk_pmg_info(hfk_pmg* g, double* info);
/* y = A_c x with the assembled coarse matrix of the last update (host arrays of the coarse level's free dofs):
 * through the node-block ELL product the solver uses (scalar_csc = 0) or the scalar CSC product (1). */
int hfk_pmg_coarse_matvec(hfk_pmg* g, const double* x, double* y, int scalar_csc);

/* ---- learned coarse correction (no counterpart in the reference) ------------------------------------- */
/* A 4-level 3D convolutional U-Net (channels 16-32-64-64, bias-free 3x3x3 convolutions with ReLU, fixed
 * full-weighting restriction and trilinear prolongation between levels, skip concatenations), BF16 weights and
 * activations on MFMA with FP32 accumulation and FP32 master weights.  It maps a residual of the assembled Q1
 * coarse level to an approximation of its coarse-level error: z = (s/h) N(r/s), s = RMS of r, h = cell size.
 *   n_lattice  vertices per direction of the Q1 lattice (2^3 m + 1, e.g. 17, 33, 65)
 *   coarsen    0: the U-Net runs on that lattice; 1: on every other vertex ((n_lattice + 1) / 2 per direction), fed
 *              by full-weighting restriction of the input and prolongated trilinearly back
 *   node_dofs  n_lattice^3 * 3 signed coarse dof ids per vertex and component (x fastest; > 0 free) */
typedef struct hfk_net hfk_net;
int hfk_net_create(hfk_ctx* ctx, int n_lattice, int coarsen, const int32_t* node_dofs, int64_t ncoarse, double h,
                   uint64_t seed, hfk_net** out);
int hfk_net_destroy(hfk_net* net);
int64_t hfk_net_num_params(const hfk_net* net);
int hfk_net_get_weights(hfk_net* net, float* w);
int hfk_net_set_weights(hfk_net* net, const float* w);
int hfk_net_apply(hfk_net* net, const double* r, double* z);
int hfk_net_apply_dev(hfk_net* net, const double* r_dev, double* z_dev, void* stream);
/* loss (sum over samples of ||z_k - e_k||^2 / ||e_k||^2 in the network's units) and its gradient with respect to
 * the FP32 weights, for samples [k][ncoarse] on the host; grad may be NULL */
int hfk_net_gradient(hfk_net* net, const double* r, const double* e, int nsamples, float* grad, double* loss);
/* Adam training on device-resident pairs; history[steps] receives the mean loss of each step (may be NULL) */
int hfk_net_train_dev(hfk_net* net, const double* r_dev, const double* e_dev, int nsamples, int steps, int batch,
                      double lr, uint64_t seed, double* history, void* stream);
/* The network as the coarse level of a p-multigrid: mode 0 off (block-Jacobi PCG), 1 the network's output is the
 * initial guess of the coarse PCG (same stopping rule), 2 the network's output is the coarse correction, 3 the
 * coarse solve is a flexible CG (same stopping rule) preconditioned by one network cycle, 4 one network cycle is
 * the coarse correction.  The network cycle is x = S b; x += N(b - A x); x += S(b - A x) with S the Chebyshev
 * smoother of the assembled coarse matrix (degree smoother_degree, 0 for the p-multigrid's; the p-multigrid's range
 * and inflation); modes 3 and 4 fall back to mode 0 when that smoother has no usable interval.  scaled = 1: the
 * network acts on the diagonally scaled coarse system, z = D^-1/2 N(D^-1/2 r) with D = diag(A) (also in training
 * and in hfk_pmg_coarse_network_error); 0: z = N(r).  The setting stays with the p-multigrid. */
int hfk_pmg_set_coarse_network(hfk_pmg* g, hfk_net* net, int mode, int smoother_degree, int scaled);
/* The plain p-multigrid as the fallback of a solve with a coarse network (hfk_krylov_solve with precond 2): when the
 * residual norm has not fallen by `factor` (0 < factor < 1) over the last `window` iterations, or the preconditioned
 * direction is not finite, the network is set aside for the rest of that solve and the plain cycle (the coarse PCG
 * or direct solve) preconditions from there on; CG/FCG restart their direction there.  window 0 turns the watch
 * off.  Default window 16, factor 0.5.  *count receives the number of solves that fell back so far. */
int hfk_pmg_set_fallback(hfk_pmg* g, int window, double factor);
int hfk_pmg_fallbacks(hfk_pmg* g, int64_t* count);
/* Training data from the plain p-multigrid: for each fine right-hand side b_k ([k][nfine], device) the
 * pre-smoothed residual restricted to the coarse level (bc_k) -- with coarse_smoothed != 0, further the residual of
 * the coarse Chebyshev pre-smoothing of it (the input of the network cycle) -- and, when ec_dev is not NULL, the
 * coarse solution of bc_k to rtol (ec_k). */
int hfk_pmg_coarse_pairs_dev(hfk_pmg* g, const double* b_dev, int nsamples, double* bc_dev, double* ec_dev,
                             int coarse_smoothed, double rtol, int maxiter, void* stream);
/* An auxiliary input channel per lattice vertex (e.g. the normalised log shear modulus); NULL removes it. */
int hfk_net_set_aux(hfk_net* net, const float* aux);
/* Energy training on coarse right-hand sides bc_k (device): Adam on L(w) = w'(A/h)w - 2 w'x, x = r/s, whose
 * minimiser is the coarse solve; one product with the assembled coarse matrix per sample, no exact solves. */
int hfk_pmg_train_coarse_network(hfk_pmg* g, hfk_net* net, const double* bc_dev, int nsamples, int steps, int batch,
                                 double lr, uint64_t seed, double* history, void* stream);
/* The energy loss of hfk_pmg_train_coarse_network summed over the samples bc_k (device) and its weight gradient
 * (host FP32, hfk_net_num_params values; may be NULL), without a training step (no counterpart in the reference). */
int hfk_pmg_coarse_network_energy_gradient(hfk_pmg* g, hfk_net* net, const double* bc_dev, int nsamples, float* grad,
                                           double* loss, void* stream);
/* Mean relative energy-norm error of the network's coarse correction against the coarse solve to 1e-10. */
int hfk_pmg_coarse_network_error(hfk_pmg* g, hfk_net* net, const double* bc_dev, int nsamples, double* err,
                                 void* stream);

/* ---- MatrixFreeSolver solve! (src/solvers.jl:111-158) ------------------------------------------------- */
/* kind 0 = CG (left preconditioned, stops on a non-positive curvature as cg! with linesearch), 1 = flexible CG,
 * 2 = GMRES, 3 = FGMRES (right preconditioned, restarted every `restart` iterations, as gmres!/fgmres! with
 * restart = true; the reference's default is GMRES with restart 50); precond 0 = none, 1 = op's BlockJacobi
 * (updated by hfk_block_jacobi_update), 2 = one V-cycle of pmg.  From x = 0 to ||b - A x||_2 <= atol +
 * rtol ||b||_2.  status = (converged, iterations, indefinite); rnorm = (true residual norm, ||b||).  A failed
 * preconditioner (pmg not ok) returns at once with converged = 0; a stop without the method's own test is decided
 * on the true residual. */
int hfk_krylov_solve(hfk_op* op, hfk_pmg* pmg, int kind, int precond, const double* b, double* x, double rtol,
                     double atol, int maxiter, int restart, int* status, double* rnorm);
int hfk_krylov_solve_dev(hfk_op* op, hfk_pmg* pmg, int kind, int precond, const double* b_dev, double* x_dev,
                         double rtol, double atol, int maxiter, int restart, int* status, double* rnorm, void* stream);

/* ------------------------------------------------------------------------------------------------------------
 * Oseen-BDF2 step on tetrahedra, Taylor-Hood P2-P1 with SUPG/PSPG/LSIC (the weak form of test/common.jl:165-228:
 * volume terms, elementwise Laplacian in the stabilisation load, outlet backflow/Robin/pressure terms): coloured FP64
 * assembly of the residual and the CSC Jacobian, as residual!/jacobian!
 * (src/assembly.jl:909-927) do for that form, with the reference's sparsity (src/assembly.jl:773-899) and colouring
 * (src/coloring.jl:4-28) over the 34 dofs of each cell (velocity and pressure blocks all coupled).
 *   cell_dofs  ncells x 34 signed ids in the host's local order (Gridap's get_cell_dof_ids of the multi-field
 *              space, src/tables.jl:329-342); > 0 free (1-based, velocity and pressure in one numbering), < 0
 *              Dirichlet (index into dvals)
 *   nodeperm   10: the host's P2 table node of the library's node f, where f = 0..3 are the cell's vertices in the
 *              order of `coords` and 4..9 the midpoints of their edges (0,1) (0,2) (0,3) (1,2) (1,3) (2,3)
 *   pnodeperm  4: the host's P1 (pressure) table node of vertex v
 *   ldof       34: ldof[c*10 + n] = local dof of velocity component c at host P2 node n, ldof[30 + n] = local dof
 *              of the pressure at host P1 node n (FieldTable ldof, src/tables.jl:171-237).  NULL for any of the three
 *              arrays means the identity (velocity dofs [c*10 + f], pressure [30 + v])
 *   coords     ncells x 4 x 3 vertex coordinates (affine cells)
 *   qpoints    np x 3 reference coordinates of the quadrature points on the unit tetrahedron, qweights np weights
 *   rho, mu    density and dynamic viscosity; c0 the time-derivative coefficient (1/dt Euler, 1.5/dt BDF2)
 * Point data (hfk_oseen_set_fields): ncells x np x 9 = advecting velocity w, velocity un of the stabilisation
 * parameters, history term f. */
typedef struct hfk_oseen hfk_oseen;
int hfk_oseen_create(hfk_ctx* ctx, int64_t ncells, const int32_t* cell_dofs, const int32_t* nodeperm,
                     const int32_t* pnodeperm, const int32_t* ldof, const double* coords, int np,
                     const double* qpoints, const double* qweights, int64_t nfree, int64_t ndir, double rho, double mu,
                     double c0, hfk_oseen** out);
int hfk_oseen_destroy(hfk_oseen* op);
int hfk_oseen_set_fields(hfk_oseen* op, const double* fields);
/* refresh_dirichlet_values! (src/assembly.jl:393-410) */
int hfk_oseen_set_dirichlet(hfk_oseen* op, const double* dvals);
/* The elementwise Laplacian lapn of un at the points, ncells x np x 3 (test/common.jl:191-195: rk = rho f + mu lapn
 * in the SUPG/PSPG load); NULL removes the term. */
int hfk_oseen_set_laplacian(hfk_oseen* op, const double* lap_points);
/* The step's point data materialised on the device from the FE functions of the two previous steps (the FE-function
 * leaves of test/common.jl:183-196, which src/fefunctions.jl:156+ and src/assembly.jl:936-1052 materialise on the CPU;
 * the time loop of test/test_cfd.jl:47-75 overwrites them in place every step): un = u^n and unm1 = u^{n-1}, each
 * given by its free values (nfree, the operator's numbering; pressure entries are not read) and its Dirichlet values
 * (ndir).  At every quadrature point w = aw un + bw unm1 (advecting velocity: aw = 1, bw = 0 for backward Euler,
 * 2, -1 for BDF2), un, and f = c1 un - c2 unm1 (history); with laplacian != 0 the elementwise Laplacian of un
 * (test/common.jl:191-195), otherwise that term is removed; and w at the dGb points of the outlet faces.  Replaces
 * hfk_oseen_set_fields, hfk_oseen_set_laplacian and the w of hfk_oseen_set_outlet.  _dev: device pointers,
 * asynchronous on `stream`. */
int hfk_oseen_step_fields(hfk_oseen* op, const double* un, const double* un_dir, const double* unm1,
                          const double* unm1_dir, double aw, double bw, double c1, double c2, int laplacian);
int hfk_oseen_step_fields_dev(hfk_oseen* op, const double* un, const double* un_dir, const double* unm1,
                              const double* unm1_dir, double aw, double bw, double c1, double c2, int laplacian,
                              void* stream);
/* Outlet faces (test/common.jl:197-200, 212-213, 219): a += -rho negpart(w.n) u.v dGb + ZA (u.n)(v.n) dGo and
 * l += -Pc v.n dGo.  Per face: its cell (0-based), the outward unit normal (nfaces x 3), the nqo points of dGo as
 * [face][q][4] = reference coordinates in the cell (3) and weight dS, the nqb points of dGb as [face][q][7] =
 * reference coordinates (3), dS, w at the point (3).  1..32 points per measure; nfaces = 0 removes the terms. */
int hfk_oseen_set_outlet(hfk_oseen* op, int64_t nfaces, const int32_t* face_cell, const double* normals, int nqo,
                         const double* xo, int nqb, const double* xb, double Pc, double ZA);
/* rho, mu and c0 rebound for the next assemblies (update(op, res, jac) with the same structure: load and time
 * stepping, e.g. the switch from backward Euler to BDF2, c0 = 1/dt -> 1.5/dt, test/test_cfd.jl:66-75); colouring,
 * pattern and point data are kept. */
int hfk_oseen_set_parameters(hfk_oseen* op, double rho, double mu, double c0);
/* The outlet pressure Pc alone (a number of the integrand, updated every step by a Windkessel model). */
int hfk_oseen_set_outlet_pressure(hfk_oseen* op, double Pc);
/* The volume integrand of the step as a point form compiled at run time with hiprtc (the generated pointwise code of
 * the captured integrand, PointForm, src/codegen.jl:122-198), in place of the built-in SUPG/PSPG/LSIC integrand:
 * `source` (HIP) defines a struct `name` with
 *   static __device__ void residual(const double* s, const double* d, double h, const double* prm, double* co);
 *   static __device__ void tangent(const double* d, double h, const double* prm, double* k);
 * s: the 16 state slots at a point (u (3), du_i/dx_J at 3 + 3i + J, p, dp/dx_J at 13 + J); d: w, un, f, lap (12);
 * h the cell size (6 sqrt(2) |K|)^(1/3); prm: rho, mu, c0 (hfk_oseen_set_parameters), then `params` (at most 5).
 * residual: the coefficient of each of the 16 test slots; tangent: the 50 factors of the factorised tangent (see
 * hfk_oseen_form.hpp) and `static constexpr unsigned long long kNonzero`, bit i set unless factor i is zero.  Code objects
 * are cached by source text; a form that does not compile gives HFK_EARG with the
 * compiler log.  source = NULL restores the built-in integrand.  hfk/formgen.py emits such sources from an
 * integrand written in Python. */
int hfk_oseen_set_form(hfk_oseen* op, const char* source, const char* name, const double* params, int nparams);
/* Compile a point form without a device (hiprtc only): HFK_OK, or HFK_EARG with the compiler log in `log`. */
int hfk_oseen_form_check(const char* source, const char* name, char* log, int64_t log_size);
/* Deterministic assembly (1, the default: the Jacobian colour by colour with plain adds in the reference's colour
 * order, src/assembly.jl:1413-1432; the residual as element rows of one launch summed per dof in ascending (cell,
 * row) order; both bitwise reproducible) or one launch over the cells in mesh order with FP64 atomic adds (0: the
 * CSC values that neighbouring cells share stay in L2; the order of the additions into one entry varies between
 * runs). */
int hfk_oseen_set_deterministic(hfk_oseen* op, int deterministic);
/* info[6]: nfree, ndir, ncells, np, ncolors, nnz (-1 before the pattern) */
int hfk_oseen_info(hfk_oseen* op, int64_t* info);
int hfk_oseen_colors(hfk_oseen* op, int32_t* colors);
/* sparsity_pattern / nonzero_positions (src/assembly.jl:768-899): 1-based CSC and nzpos [cell][ii + 34 jj], ii and jj
 * the host's local dofs (the order of the cell_dofs given to hfk_oseen_create) */
/* Field blocks (block_spec, cell_block_mask, block_keys!: src/assembly.jl:276-323, 825-893): which (test field,
 * trial field) blocks carry a term, mask[fi + 2 fj] with field 0 the velocity and 1 the pressure, the same on every
 * cell, or cell_mask (ncells x 4) per cell; both NULL: every block (the default).  The pattern then holds the free
 * pairs of the set blocks only, and nzpos of a pair outside its cell's blocks is its stored position when another
 * cell stores it, 0 otherwise.  Rebuilds the pattern at its next use (allocate the Jacobian again). */
int hfk_oseen_set_blocks(hfk_oseen* op, const int32_t* mask, const int32_t* cell_mask);
int hfk_oseen_pattern_nnz(hfk_oseen* op, int64_t* nnz);
int hfk_oseen_pattern(hfk_oseen* op, int64_t* colptr, int64_t* rowval, int32_t* nzpos);
/* residual! (src/assembly.jl:909-916): b = a(x, .) - l(.) on the free test dofs, Dirichlet values lifted */
int hfk_oseen_residual(hfk_oseen* op, const double* x, double* b);
int hfk_oseen_residual_dev(hfk_oseen* op, const double* x_dev, double* b_dev, void* stream);
/* jacobian! (src/assembly.jl:918-927): the CSC values (the step is linear; x is not read) */
int hfk_oseen_jacobian(hfk_oseen* op, const double* x, double* nzval);
int hfk_oseen_jacobian_dev(hfk_oseen* op, const double* x_dev, double* nzval_dev, void* stream);
/* residual_and_jacobian! (src/assembly.jl:929-933) in one pass */
int hfk_oseen_residual_jacobian_dev(hfk_oseen* op, const double* x_dev, double* b_dev, double* nzval_dev,
                                    void* stream);

/* ------------------------------------------------------------------------------------------------------------
 * Assembled multi-field systems: a square CSC matrix on the device (the values of jacobian!, e.g.
 * hfk_oseen_jacobian_dev) with block Jacobi over nodal blocks that couple every field's dofs at a mesh node, damped
 * block-Jacobi sweeps as a smoother, and GMRES / FGMRES.  The reference's BlockJacobi (src/preconditioners.jl:
 * 136-171 field_blocks, 284-340 invert_blocks!/ldiv!) keeps one set of nodal blocks per field, which leaves the
 * pressure blocks of a mixed formulation without a pressure diagonal singular; its p-multigrid is single-field
 * for that reason (src/pmultigrid.jl:252-254).  A block here is any group of at most 6 free dofs given by the
 * host: the reference's per-field blocks and blocks coupling velocity and pressure at a node run on the same
 * kernels.  Empty slots act as the identity (as the reference's fixed components).
 */
typedef struct hfk_sys hfk_sys;
/* colptr (n + 1) and rowval (colptr[n] - 1): the CSC pattern, 1-based, rows ascending within a column */
int hfk_sys_create(hfk_ctx* ctx, int64_t n, const int64_t* colptr, const int64_t* rowval, hfk_sys** out);
int hfk_sys_destroy(hfk_sys* sys);
/* nblocks x 6 free dofs (1-based, 0 = an empty slot); every dof in exactly one slot (preconditioners.jl:151-171:
 * the blocks of field_blocks, here over any grouping) */
int hfk_sys_set_blocks(hfk_sys* sys, int64_t nblocks, const int32_t* block_dofs);
/* new values (CSC order, device): the CSR copy for products, then the blocks extracted and inverted
 * (block_diagonal! + invert_blocks!, preconditioners.jl:186-218, 284-304); nsingular = blocks whose inverse is not
 * finite (the reference's `ok` cleared, or its strict error) */
int hfk_sys_update_dev(hfk_sys* sys, const double* nzval_dev, int64_t* nsingular, void* stream);
/* y = A x */
int hfk_sys_matvec_dev(hfk_sys* sys, const double* x_dev, double* y_dev, void* stream);
/* z = Binv r block by block (ldiv!, preconditioners.jl:310-340) */
int hfk_sys_blocks_apply_dev(hfk_sys* sys, const double* r_dev, double* z_dev, void* stream);
/* `sweeps` damped block-Jacobi steps z <- z + omega Binv (r - A z) from z = 0 (0 < omega < 2) */
int hfk_sys_smooth_dev(hfk_sys* sys, int sweeps, double omega, const double* r_dev, double* z_dev, void* stream);
/* host copies of the extracted blocks and their inverses, nblocks x 6 x 6 row-major each (either may be NULL) */
int hfk_sys_blocks(hfk_sys* sys, double* B, double* Binv);
/* Overlapping patches (no counterpart in the reference: the Vanka blocks of a saddle point, a pressure dof with the
 * velocity dofs around it, whose nodal blocks alone are too weakly coupled): the dofs of patch p are
 * dofs[ptr[p] .. ptr[p+1]) (1-based, ptr[0] = 0, 1 to 80 per patch), owned[k] != 0 when patch p writes dof dofs[k];
 * every dof owned by exactly one patch.  The next hfk_sys_update_dev extracts the dense patch matrices and inverts
 * them (singular ones counted in nsingular); the application is restricted additive Schwarz:
 * z[d] = (A_p^-1 r_p)[d] from the patch p that owns d. */
int hfk_sys_set_patches(hfk_sys* sys, int64_t npatch, const int64_t* ptr, const int32_t* dofs, const int8_t* owned);
int hfk_sys_patches_apply_dev(hfk_sys* sys, const double* r_dev, double* z_dev, void* stream);
/* host copy of the patch inverses, patch after patch, each m x m row-major (sum of m^2 doubles) */
int hfk_sys_patch_inverses(hfk_sys* sys, double* Pinv);
/* solve! as MatrixFreeSolver (src/solvers.jl:111-158) on the assembled matrix: kind 2 = GMRES, 3 = FGMRES
 * (right preconditioned, restarted every `restart`); precond 0 = none, 1 = the blocks, 2 = `sweeps` damped sweeps
 * with `omega`, 3 = the patches; from x = 0 to ||b - A x|| <= atol + rtol ||b||; status = (converged, iterations,
 * 0), rnorm = (true residual norm, ||b||) */
int hfk_sys_solve_dev(hfk_sys* sys, int kind, int precond, int sweeps, double omega, const double* b_dev, double* x_dev,
                      double rtol, double atol, int maxiter, int restart, int* status, double* rnorm, void* stream);

#ifdef __cplusplus
}
#endif
#endif
