// Internal declarations of the solver-layer kernels (hfk_solver.hip).  Not part of the C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace hfk {

// device scalars of a CG/FCG iteration
enum : int { kRZ = 0, kPQ = 1, kZQ = 2, kRZN = 3, kRR = 4, kBB = 5, kRZ2 = 6, kNS = 8 };  // kRZ2, kRZ2 + 1: fused PCG

// out[0..] = a0.b0 (, a1.b1 (, a2.b2)): deterministic two-pass reductions on a fixed grid; `part` holds
// dot_partial_doubles() doubles
hipError_t vec_dots(int64_t n, const double* a0, const double* b0, const double* a1, const double* b1, const double* a2,
                    const double* b2, double* part, double* out, hipStream_t s);
size_t dot_partial_doubles();
hipError_t vec_axpby(int64_t n, double a, const double* x, double b, double* y, hipStream_t s);  // y = a x + b y
hipError_t vec_copy(int64_t n, const double* x, double* y, hipStream_t s);
// GMRES: out[k] = V[k] . w for k < K (V: K vectors at stride ldv; part: multi_dot_partial_doubles(K)), and
// w += alpha sum_k coef[k] V[k] (coef on the device)
size_t multi_dot_partial_doubles(int K);
hipError_t multi_dot(int64_t n, const double* V, int64_t ldv, int K, const double* w, double* part, double* out,
                     hipStream_t s);
hipError_t multi_axpy(int64_t n, const double* V, int64_t ldv, int K, const double* coef, double alpha, double* w,
                      hipStream_t s);
// Arnoldi step j on the device: H[i + ldh j] = c1[i] + c2[i] (i <= j), H[j + 1 + ldh j] = sqrt(ww[0]), w /= sqrt(ww[0])
hipError_t arnoldi_column(int64_t n, int j, int ldh, const double* c1, const double* c2, const double* ww, double* H,
                          double* w, hipStream_t s);
// node-block ELL product y = alpha A x + beta y (val FP64, or FP32 when fp32) and its value fill from the CSC
// one block-Jacobi PCG iteration on the node-block ELL matrix (FP64 values, or FP32 values with FP64 accumulation) in
// three launches: rz in sc[slot] (the new rz goes to sc[next], r.r to sc[kRR]); part holds 3 ceil(nb / 256) partial
// sums
hipError_t pcg_fused_iteration(int64_t nb, int W, const int32_t* nbr, const void* val, bool fp32, const int32_t* dofs,
                               const double* Binv, double* x, double* r, double* z, double* p, double* q, double* sc,
                               int slot, int next, double* part, hipStream_t s);
hipError_t bell_spmv(int64_t nb, int W, const int32_t* nbr, const void* val, bool fp32, const int32_t* dofs,
                     const double* x, double* y, double alpha, double beta, hipStream_t s);
hipError_t bell_fill(int64_t nnz, const int32_t* pos, const double* val, double* v64, float* v32, hipStream_t s);
hipError_t vec_mul(int64_t n, const double* a, const double* x, double* y, hipStream_t s);  // y = a .* x (y may be x)
// d = 1/sqrt(diag(A)) of a CSC matrix with sorted rows (1 where the diagonal is missing, non-positive or not finite)
hipError_t csc_diag_rsqrt(int64_t ncols, const int32_t* colptr, const int32_t* rowval, const double* val, double* d,
                          hipStream_t s);
hipError_t cg_update(int64_t n, const double* sc, const double* p, const double* q, double* x, double* r, hipStream_t s);
hipError_t cg_direction(int64_t n, double* sc, const double* z, double* p, int flexible, hipStream_t s);
hipError_t blocks_invert(int64_t nb, const int32_t* dofs, const double* B, double* Binv, int* bad, hipStream_t s);
hipError_t blocks_step(int64_t nb, const int32_t* dofs, const double* Binv, const double* r, double* d, double* x,
                       double cd, double cz, hipStream_t s);
// the same with FP32 inverses (a smoother on FP32 tangents); b[i] = (float) a[i]
hipError_t blocks_step(int64_t nb, const int32_t* dofs, const float* Binv, const double* r, double* d, double* x,
                       double cd, double cz, hipStream_t s);
hipError_t vec_to_fp32(int64_t n, const double* a, float* b, hipStream_t s);
hipError_t csr_blocks(int64_t nb, const int32_t* dofs, const int32_t* colptr, const int32_t* rowval, const double* val,
                      double* B, hipStream_t s);
hipError_t csc_t_spmv(int64_t ncols, const int32_t* colptr, const int32_t* rowval, const double* val, const double* x,
                      double* y, double alpha, double beta, hipStream_t s);
hipError_t transfer_prolongate(int64_t nslot, int NF, int NC, const int32_t* fid, const int32_t* cid, const double* E,
                               const double* xc, double* yf, hipStream_t s);
hipError_t transfer_restrict(int64_t slot0, int64_t nslot, int NF, int NC, const int32_t* fid, const int32_t* cid,
                             const double* E, const double* invmult, const double* yf, double* yc, hipStream_t s);
// yf = P xc with one thread per fine free dof from its owning cell slot (coarse Q1: NC = 8)
hipError_t transfer_prolongate_owner(int64_t nf, int NF, const int64_t* owner, const int32_t* cid, const double* E,
                                     const double* xc, double* yf, hipStream_t s);
hipError_t transfer_inject(int64_t nslot, int NF, int NC, const int32_t* fid, const int32_t* cid, const double* Einj,
                           const double* xf, const double* df, double* xc, double* dc, hipStream_t s);
hipError_t count_nonfinite(int64_t n, const double* x, int* bad, hipStream_t s);
// y = A x for a CSR matrix with 0-based columns (rp: nrows + 1 offsets)
hipError_t csr_spmv(int64_t nrows, const int64_t* rp, const int32_t* col, const double* val, const double* x, double* y,
                    hipStream_t s);
// the same with `lanes` (1, 4, 8 or 16) threads per row; csr_lanes picks them from the mean row length
int csr_lanes(int64_t nrows, int64_t nnz);
hipError_t csr_spmv_lanes(int64_t nrows, const int64_t* rp, const int32_t* col, const double* val, const double* x,
                          double* y, int lanes, hipStream_t s);
// out[i] = src[i] > 0 ? xf[src[i] - 1] : df[-src[i] - 1] (nodal injection between nested meshes)
hipError_t inject_map(int64_t n, const int32_t* src, const double* xf, const double* df, double* out, hipStream_t s);
// y = A x with a dense n x n matrix (row-major)
hipError_t dense_matvec(int64_t n, const double* A, const double* x, double* y, hipStream_t s);
// Ainv = inverse of the symmetric part of an n x n CSC matrix (0-based), by Gauss-Jordan without pivoting on the
// device (W: n*n work, rk, ck: n each; *bad set when a pivot is not positive or not finite; *bad must start at 0)
hipError_t dense_spd_inverse(int64_t n, const int32_t* colptr, const int32_t* rowval, const double* val, double* W,
                             double* rk, double* ck, double* Ainv, int* bad, hipStream_t s);

}  // namespace hfk
