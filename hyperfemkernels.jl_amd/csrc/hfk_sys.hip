// Assembled multi-field systems on the device: a square CSC matrix (the values of jacobian!, e.g. the Oseen P2-P1
// Jacobian) with a CSR copy of its values for y = A x, block Jacobi over nodal blocks that couple the dofs of every
// field at a mesh node, damped block-Jacobi sweeps as a smoother, overlapping patches (Vanka: a pressure dof with
// the velocity dofs around it) applied as restricted additive Schwarz, and GMRES / FGMRES over them.
//
// The reference's BlockJacobi (src/preconditioners.jl:136-171, 284-340) keeps one set of blocks per field (the
// components of that field at a node); a mixed formulation without a pressure diagonal leaves its pressure blocks
// singular, and the p-multigrid is single-field for that reason (src/pmultigrid.jl:252-254).  Here a block is any
// group of at most six free dofs given by the host (the Python host groups the velocity and the pressure of a
// node), so the same kernels serve the reference's per-field blocks and coupled ones.  Rows and columns of an empty
// slot are the identity, as the reference's fixed components; a block whose inverse is not finite is counted.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "hfk.h"
#include "hfk_internal.hpp"
#include "hfk_krylov.hpp"
#include "hfk_op.hpp"
#include "hfk_solver.hpp"

using namespace hfk;

namespace {

constexpr int kBS = 6;         // dofs per block (padded)
constexpr int kRowLanes = 16;  // lanes per row of the CSR product

inline dim3 grid_n(int64_t n, int threads = 256) {
  return dim3((unsigned)std::min<int64_t>((n + threads - 1) / threads, 1 << 20));
}

__global__ void k_gather_values(int64_t nnz, const int32_t* __restrict__ perm, const double* __restrict__ csc,
                                double* __restrict__ csr) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < nnz; k += (int64_t)gridDim.x * blockDim.x)
    csr[k] = csc[perm[k]];
}

// y = A x (or y = b - A x when b is given): kRowLanes lanes per row, each summing the entries k = lane, lane + 16,
// ... of its row, then a butterfly over the 16 lanes (a fixed order: the result does not depend on the launch)
__global__ __launch_bounds__(256) void k_csr_rows(int64_t n, const int64_t* __restrict__ rp,
                                                  const int32_t* __restrict__ col, const double* __restrict__ val,
                                                  const double* __restrict__ x, const double* __restrict__ b,
                                                  double* __restrict__ y) {
  const int lane = threadIdx.x & (kRowLanes - 1);
  const int64_t rows_per_launch = (int64_t)gridDim.x * (blockDim.x / kRowLanes);
  for (int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kRowLanes; i < n; i += rows_per_launch) {
    double v = 0.0;
    for (int64_t k = rp[i] + lane; k < rp[i + 1]; k += kRowLanes) v += val[k] * x[col[k]];
#pragma unroll
    for (int off = kRowLanes / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kRowLanes);
    if (lane == 0) y[i] = b ? b[i] - v : v;
  }
}

// B[blk][i][j] = A[d_i][d_j] (CSR search of row d_i for column d_j; 0 outside the pattern), identity on empty slots
__global__ void k_block_extract(int64_t nb, const int32_t* __restrict__ dofs, const int64_t* __restrict__ rp,
                                const int32_t* __restrict__ col, const double* __restrict__ val, double* __restrict__ B) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < nb * kBS * kBS;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t blk = t / (kBS * kBS);
    const int e = (int)(t - blk * kBS * kBS), i = e / kBS, j = e - (e / kBS) * kBS;
    const int32_t di = dofs[blk * kBS + i], dj = dofs[blk * kBS + j];
    double v = i == j ? 1.0 : 0.0;
    if (di > 0 && dj > 0) {
      v = 0.0;
      int64_t lo = rp[di - 1], hi = rp[di] - 1;
      const int32_t c = dj - 1;
      while (lo <= hi) {
        const int64_t mid = (lo + hi) >> 1;
        const int32_t cm = col[mid];
        if (cm == c) {
          v = val[mid];
          break;
        }
        if (cm < c) lo = mid + 1;
        else hi = mid - 1;
      }
    }
    B[t] = v;
  }
}

// Binv = B^-1 per block: Gauss-Jordan with partial pivoting on the 6 x 6 (row-major); a block with a non-finite
// entry is left as it is and a block whose inverse is not finite (a zero pivot included) is counted in *bad
__global__ __launch_bounds__(256) void k_block_invert(int64_t nb, const double* __restrict__ B, double* __restrict__ Binv,
                                                      int* __restrict__ bad) {
  for (int64_t blk = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; blk < nb; blk += (int64_t)gridDim.x * blockDim.x) {
    double a[kBS][kBS], r[kBS][kBS];
    bool finite = true;
#pragma unroll
    for (int i = 0; i < kBS; ++i)
#pragma unroll
      for (int j = 0; j < kBS; ++j) {
        a[i][j] = B[blk * kBS * kBS + i * kBS + j];
        r[i][j] = i == j ? 1.0 : 0.0;
        finite = finite && isfinite(a[i][j]);
      }
    if (!finite) {
#pragma unroll
      for (int e = 0; e < kBS * kBS; ++e) Binv[blk * kBS * kBS + e] = B[blk * kBS * kBS + e];
      atomicAdd(bad, 1);
      continue;
    }
#pragma unroll
    for (int k = 0; k < kBS; ++k) {
      int p = k;
      double best = fabs(a[k][k]);
#pragma unroll
      for (int i = k + 1; i < kBS; ++i)
        if (fabs(a[i][k]) > best) {
          best = fabs(a[i][k]);
          p = i;
        }
#pragma unroll
      for (int i = k + 1; i < kBS; ++i)
        if (i == p) {
#pragma unroll
          for (int j = 0; j < kBS; ++j) {
            const double t1 = a[k][j], t2 = r[k][j];
            a[k][j] = a[i][j];
            r[k][j] = r[i][j];
            a[i][j] = t1;
            r[i][j] = t2;
          }
        }
      const double inv = 1.0 / a[k][k];  // a zero pivot gives inf: the block is counted below
#pragma unroll
      for (int j = 0; j < kBS; ++j) {
        a[k][j] *= inv;
        r[k][j] *= inv;
      }
#pragma unroll
      for (int i = 0; i < kBS; ++i) {
        if (i == k) continue;
        const double f = a[i][k];
#pragma unroll
        for (int j = 0; j < kBS; ++j) {
          a[i][j] -= f * a[k][j];
          r[i][j] -= f * r[k][j];
        }
      }
    }
    bool ok = true;
#pragma unroll
    for (int i = 0; i < kBS; ++i)
#pragma unroll
      for (int j = 0; j < kBS; ++j) {
        Binv[blk * kBS * kBS + i * kBS + j] = r[i][j];
        ok = ok && isfinite(r[i][j]);
      }
    if (!ok) atomicAdd(bad, 1);
  }
}

// z[d_i] = alpha sum_j Binv[i][j] r[d_j] (+ z[d_i] when accumulate): one thread per (block, slot)
__global__ void k_block_apply(int64_t nb, const int32_t* __restrict__ dofs, const double* __restrict__ Binv,
                              const double* __restrict__ r, double alpha, int accumulate, double* __restrict__ z) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < nb * kBS; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t blk = t / kBS;
    const int i = (int)(t - blk * kBS);
    const int32_t di = dofs[t];
    if (di <= 0) continue;
    double v = 0.0;
#pragma unroll
    for (int j = 0; j < kBS; ++j) {
      const int32_t dj = dofs[blk * kBS + j];
      if (dj > 0) v += Binv[t * kBS + j] * r[dj - 1];
    }
    z[di - 1] = accumulate ? z[di - 1] + alpha * v : alpha * v;
  }
}

// ---- overlapping patches (Vanka), restricted additive Schwarz ------------------------------------------------
constexpr int kMaxPatch = 80;  // dofs per patch: the dense inverse of one patch stays in 64 KB of LDS

// the dense matrix of every patch: P[off_p + i m_p + j] = A[d_i][d_j] (0 outside the pattern)
__global__ void k_patch_extract(int64_t total, int64_t npatch, const int64_t* __restrict__ off,
                                const int64_t* __restrict__ ptr, const int32_t* __restrict__ dofs,
                                const int64_t* __restrict__ rp, const int32_t* __restrict__ col,
                                const double* __restrict__ val, double* __restrict__ P) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    int64_t lo = 0, hi = npatch - 1;  // the patch holding entry t: off[p] <= t < off[p + 1]
    while (lo < hi) {
      const int64_t mid = (lo + hi + 1) >> 1;
      if (off[mid] <= t) lo = mid;
      else hi = mid - 1;
    }
    const int64_t p = lo, m = ptr[p + 1] - ptr[p], e = t - off[p];
    const int32_t di = dofs[ptr[p] + e / m], dj = dofs[ptr[p] + e % m];
    double v = 0.0;
    int64_t a = rp[di - 1], z = rp[di] - 1;
    while (a <= z) {
      const int64_t mid = (a + z) >> 1;
      const int32_t cm = col[mid];
      if (cm == dj - 1) {
        v = val[mid];
        break;
      }
      if (cm < dj - 1) a = mid + 1;
      else z = mid - 1;
    }
    P[t] = v;
  }
}

// in-place Gauss-Jordan inverse of each patch matrix with partial (row) pivoting, one 64-lane work-group per patch
// in LDS; the column interchanges are undone at the end.  A zero pivot or a non-finite inverse is counted in *bad.
// The inverse goes back to P (row-major), and its owned rows, which are all the application reads, to Q: patch p's
// o_p owned rows column by column, Q[qoff_p + j o_p + r] = Pinv_p[orow_r][j], so that lanes r read consecutively.
__global__ __launch_bounds__(64) void k_patch_invert(int64_t npatch, const int64_t* __restrict__ off,
                                                     const int64_t* __restrict__ ptr, const int64_t* __restrict__ optr,
                                                     const int32_t* __restrict__ orow, const int64_t* __restrict__ qoff,
                                                     double* __restrict__ P, double* __restrict__ Q,
                                                     int* __restrict__ bad) {
  __shared__ double M[kMaxPatch * kMaxPatch];
  __shared__ double F[kMaxPatch];
  __shared__ int piv[kMaxPatch];
  __shared__ int fail;
  const int64_t p = blockIdx.x;
  if (p >= npatch) return;
  const int m = (int)(ptr[p + 1] - ptr[p]);
  double* G = P + off[p];
  const int lane = threadIdx.x;
  for (int e = lane; e < m * m; e += 64) M[e] = G[e];
  if (lane == 0) fail = 0;
  __syncthreads();
  for (int k = 0; k < m; ++k) {
    // pivot: the largest |M[i][k]|, i >= k (ties: the smallest i), by a butterfly over the lanes
    double best = -1.0;
    int arg = m;
    for (int i = k + lane; i < m; i += 64) {
      const double v = fabs(M[i * m + k]);
      if (v > best) {
        best = v;
        arg = i;
      }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const double ob = __shfl_xor(best, o);
      const int oa = __shfl_xor(arg, o);
      if (ob > best || (ob == best && oa < arg)) {
        best = ob;
        arg = oa;
      }
    }
    if (lane == 0) {
      piv[k] = arg;
      if (!(best > 0.0)) fail = 1;
    }
    if (arg != k && arg < m)
      for (int j = lane; j < m; j += 64) {
        const double t = M[k * m + j];
        M[k * m + j] = M[arg * m + j];
        M[arg * m + j] = t;
      }
    __syncthreads();
    const double inv = 1.0 / M[k * m + k];
    __syncthreads();
    for (int j = lane; j < m; j += 64) M[k * m + j] = (j == k ? 1.0 : M[k * m + j]) * inv;
    for (int i = lane; i < m; i += 64) F[i] = M[i * m + k];
    __syncthreads();
    for (int e = lane; e < m * m; e += 64) {
      const int i = e / m, j = e - (e / m) * m;
      if (i == k) continue;
      M[e] = (j == k ? 0.0 : M[e]) - M[k * m + j] * F[i];
    }
    __syncthreads();
  }
  for (int k = m - 1; k >= 0; --k) {  // undo the row interchanges as column interchanges, last first
    const int q = piv[k];
    if (q != k && q < m)
      for (int i = lane; i < m; i += 64) {
        const double t = M[i * m + k];
        M[i * m + k] = M[i * m + q];
        M[i * m + q] = t;
      }
    __syncthreads();
  }
  bool ok = !fail;
  for (int e = lane; e < m * m; e += 64) {
    G[e] = M[e];
    ok = ok && isfinite(M[e]);
  }
  const int64_t o0 = optr[p];
  const int o = (int)(optr[p + 1] - o0);
  for (int e = lane; e < o * m; e += 64) {
    const int j = e / o, r = e - (e / o) * o;
    Q[qoff[p] + e] = M[orow[o0 + r] * m + j];
  }
  if (__any(!ok) && lane == 0) atomicAdd(bad, 1);
}

// restricted additive Schwarz: z[d_i] = alpha (Pinv_p r_p)_i for the slots i that patch p owns (every dof is owned
// by one patch, so each entry of z is written once); one 64-lane work-group per patch, a lane per owned row
__global__ __launch_bounds__(64) void k_patch_apply(int64_t npatch, const int64_t* __restrict__ ptr,
                                                    const int32_t* __restrict__ dofs, const int64_t* __restrict__ optr,
                                                    const int32_t* __restrict__ orow, const int64_t* __restrict__ qoff,
                                                    const double* __restrict__ Q, const double* __restrict__ r,
                                                    double alpha, double* __restrict__ z) {
  __shared__ double R[kMaxPatch];
  const int64_t p = blockIdx.x;
  if (p >= npatch) return;
  const int64_t b0 = ptr[p];
  const int m = (int)(ptr[p + 1] - b0);
  const int lane = threadIdx.x;
  for (int i = lane; i < m; i += 64) R[i] = r[dofs[b0 + i] - 1];
  __syncthreads();
  const int64_t o0 = optr[p];
  const int o = (int)(optr[p + 1] - o0);
  const double* G = Q + qoff[p];
  for (int k = lane; k < o; k += 64) {
    double v = 0.0;
    for (int j = 0; j < m; ++j) v += G[(int64_t)j * o + k] * R[j];
    z[dofs[b0 + orow[o0 + k]] - 1] = alpha * v;
  }
}

}  // namespace

struct hfk_sys {
  hfk_ctx* ctx = nullptr;
  int64_t n = 0, nnz = 0, nblocks = 0;
  DevBuf rp, col, perm, val, dofs, B, Binv, bad, t;
  // overlapping patches: dofs of patch p in pdofs[pptr[p] .. pptr[p + 1]), their dense inverses at poff[p]
  // owned rows: orow[optr[p] .. optr[p + 1]) the owned slots of patch p, their inverse rows at qoff[p] in pown
  int64_t npatch = 0, ptotal = 0;
  DevBuf pptr, pdofs, powned, poff, pinv, poptr, porow, pqoff, pown;
  bool values = false;
  hipStream_t stream(void* s) const { return s ? static_cast<hipStream_t>(s) : ctx->stream; }
};

namespace {

int sys_bind(hfk_sys* A) {
  if (!A || !A->ctx) return set_err(nullptr, HFK_EARG, "HyperFEMKernels: null assembled system");
  HIPCHK(A->ctx, hipSetDevice(A->ctx->device));
  return HFK_OK;
}

int sys_matvec(hfk_sys* A, const double* x, const double* b, double* y, hipStream_t s) {
  if (!A->values) return set_err(A->ctx, HFK_ESTATE, "HyperFEMKernels: assembled system used before its values were set");
  const int64_t threads = A->n * kRowLanes;
  hipLaunchKernelGGL(k_csr_rows, grid_n(threads), dim3(256), 0, s, A->n, A->rp.as<int64_t>(), A->col.as<int32_t>(),
                     A->val.as<double>(), x, b, y);
  HIPCHK(A->ctx, hipGetLastError());
  return HFK_OK;
}

int sys_blocks(hfk_sys* A, const double* r, double alpha, bool acc, double* z, hipStream_t s) {
  if (A->nblocks == 0) return set_err(A->ctx, HFK_ESTATE, "HyperFEMKernels: no blocks set on the assembled system");
  if (!A->values) return set_err(A->ctx, HFK_ESTATE, "HyperFEMKernels: assembled system used before its values were set");
  hipLaunchKernelGGL(k_block_apply, grid_n(A->nblocks * kBS), dim3(256), 0, s, A->nblocks, A->dofs.as<int32_t>(),
                     A->Binv.as<double>(), r, alpha, acc ? 1 : 0, z);
  HIPCHK(A->ctx, hipGetLastError());
  return HFK_OK;
}

int sys_patches(hfk_sys* A, const double* r, double alpha, double* z, hipStream_t s) {
  if (A->npatch == 0) return set_err(A->ctx, HFK_ESTATE, "HyperFEMKernels: no patches set on the assembled system");
  if (!A->values) return set_err(A->ctx, HFK_ESTATE, "HyperFEMKernels: assembled system used before its values were set");
  hipLaunchKernelGGL(k_patch_apply, dim3((unsigned)A->npatch), dim3(64), 0, s, A->npatch, A->pptr.as<int64_t>(),
                     A->pdofs.as<int32_t>(), A->poptr.as<int64_t>(), A->porow.as<int32_t>(), A->pqoff.as<int64_t>(),
                     A->pown.as<double>(), r, alpha, z);
  HIPCHK(A->ctx, hipGetLastError());
  return HFK_OK;
}

// `sweeps` damped block-Jacobi steps z <- z + omega Binv (r - A z) from z = 0: a fixed linear operator of r
int sys_sweeps(hfk_sys* A, int sweeps, double omega, const double* r, double* z, hipStream_t s) {
  int rc = sys_blocks(A, r, omega, false, z, s);
  for (int k = 1; k < sweeps && !rc; ++k) {
    if ((rc = sys_matvec(A, z, r, A->t.as<double>(), s))) return rc;  // t = r - A z
    rc = sys_blocks(A, A->t.as<double>(), omega, true, z, s);
  }
  return rc;
}

}  // namespace

extern "C" {

int hfk_sys_create(hfk_ctx* ctx, int64_t n, const int64_t* colptr, const int64_t* rowval, hfk_sys** out) {
  if (!ctx || !out || !colptr || (n > 0 && !rowval))
    return set_err(ctx, HFK_EARG, "HyperFEMKernels: hfk_sys_create: null argument");
  *out = nullptr;
  if (n <= 0 || n >= (int64_t)1 << 31) return set_err(ctx, HFK_EARG, "HyperFEMKernels: hfk_sys_create: bad size");
  if (colptr[0] != 1) return set_err(ctx, HFK_EARG, "HyperFEMKernels: hfk_sys_create: colptr must start at 1");
  for (int64_t j = 0; j < n; ++j)
    if (colptr[j + 1] < colptr[j]) return set_err(ctx, HFK_EARG, "HyperFEMKernels: hfk_sys_create: colptr decreases");
  const int64_t nnz = colptr[n] - 1;
  if (nnz >= (int64_t)1 << 31) return set_err(ctx, HFK_EARG, "HyperFEMKernels: hfk_sys_create: more than 2^31 - 1 entries");
  auto* A = new hfk_sys();
  A->ctx = ctx;
  A->n = n;
  A->nnz = nnz;
  auto fail = [&](int rc) {
    delete A;
    return rc;
  };
  if (hipSetDevice(ctx->device) != hipSuccess) return fail(set_err(ctx, HFK_EHIP, "HyperFEMKernels: hipSetDevice failed"));
  // CSR of the same matrix, built on the device: rp, the columns of every row (ascending) and perm[k] = the CSC
  // position of CSR entry k
  int bad = 0;
  int rc = device_csc_to_csr(ctx, ctx->stream, n, colptr, rowval, A->rp, A->col, A->perm, &bad);
  if (rc) return fail(rc);
  if (bad & 1) return fail(set_err(ctx, HFK_EARG, "HyperFEMKernels: hfk_sys_create: row index out of range"));
  if (bad & 2) return fail(set_err(ctx, HFK_EARG, "HyperFEMKernels: hfk_sys_create: rows of a column must ascend"));
  if (A->val.alloc(std::max<int64_t>(nnz, 1) * sizeof(double)) != hipSuccess || A->t.alloc(n * sizeof(double)) != hipSuccess ||
      A->bad.alloc(sizeof(int)) != hipSuccess)
    return fail(set_err(ctx, HFK_EHIP, "HyperFEMKernels: hfk_sys_create: device allocation failed"));
  *out = A;
  return HFK_OK;
}

int hfk_sys_destroy(hfk_sys* A) {
  if (A) {
    (void)hipSetDevice(A->ctx->device);
    delete A;
  }
  return HFK_OK;
}

int hfk_sys_set_blocks(hfk_sys* A, int64_t nblocks, const int32_t* block_dofs) {
  int rc = sys_bind(A);
  if (rc) return rc;
  if (nblocks < 0 || (nblocks > 0 && !block_dofs))
    return set_err(A->ctx, HFK_EARG, "HyperFEMKernels: hfk_sys_set_blocks: bad arguments");
  std::vector<uint8_t> seen(A->n, 0);
  for (int64_t b = 0; b < nblocks; ++b) {
    int used = 0;
    for (int i = 0; i < kBS; ++i) {
      const int32_t d = block_dofs[b * kBS + i];
      if (d == 0) continue;
      if (d < 0 || d > A->n) return set_err(A->ctx, HFK_EARG, "HyperFEMKernels: hfk_sys_set_blocks: dof out of range");
      if (seen[d - 1]) return set_err(A->ctx, HFK_EARG, "HyperFEMKernels: hfk_sys_set_blocks: a dof in two block slots");
      seen[d - 1] = 1;
      ++used;
    }
    if (used == 0) return set_err(A->ctx, HFK_EARG, "HyperFEMKernels: hfk_sys_set_blocks: an empty block");
  }
  if (nblocks > 0 && std::find(seen.begin(), seen.end(), 0) != seen.end())
    return set_err(A->ctx, HFK_EARG, "HyperFEMKernels: hfk_sys_set_blocks: the blocks must cover every dof");
  A->nblocks = nblocks;
  if (nblocks == 0) return HFK_OK;
  HIPCHK(A->ctx, A->dofs.alloc(nblocks * kBS * sizeof(int32_t)));
  HIPCHK(A->ctx, A->B.alloc(nblocks * kBS * kBS * sizeof(double)));
  HIPCHK(A->ctx, A->Binv.alloc(nblocks * kBS * kBS * sizeof(double)));
  HIPCHK(A->ctx, hipMemcpy(A->dofs.p, block_dofs, nblocks * kBS * sizeof(int32_t), hipMemcpyHostToDevice));
  A->values = false;  // the blocks of the current values are extracted by the next update
  return HFK_OK;
}

int hfk_sys_update_dev(hfk_sys* A, const double* nzval_dev, int64_t* nsingular, void* stream) {
  int rc = sys_bind(A);
  if (rc) return rc;
  if (!nzval_dev) return set_err(A->ctx, HFK_EARG, "HyperFEMKernels: hfk_sys_update_dev: null values");
  const hipStream_t s = A->stream(stream);
  hipLaunchKernelGGL(k_gather_values, grid_n(A->nnz), dim3(256), 0, s, A->nnz, A->perm.as<int32_t>(), nzval_dev,
                     A->val.as<double>());
  HIPCHK(A->ctx, hipGetLastError());
  int bad = 0;
  if (A->nblocks > 0) {
    hipLaunchKernelGGL(k_block_extract, grid_n(A->nblocks * kBS * kBS), dim3(256), 0, s, A->nblocks,
                       A->dofs.as<int32_t>(), A->rp.as<int64_t>(), A->col.as<int32_t>(), A->val.as<double>(),
                       A->B.as<double>());
    HIPCHK(A->ctx, hipGetLastError());
    HIPCHK(A->ctx, hipMemsetAsync(A->bad.p, 0, sizeof(int), s));
    hipLaunchKernelGGL(k_block_invert, grid_n(A->nblocks), dim3(256), 0, s, A->nblocks, A->B.as<double>(),
                       A->Binv.as<double>(), A->bad.as<int>());
    HIPCHK(A->ctx, hipGetLastError());
    HIPCHK(A->ctx, hipMemcpyAsync(&bad, A->bad.p, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(A->ctx, hipStreamSynchronize(s));
  }
  if (A->npatch > 0) {
    hipLaunchKernelGGL(k_patch_extract, grid_n(A->ptotal), dim3(256), 0, s, A->ptotal, A->npatch,
                       A->poff.as<int64_t>(), A->pptr.as<int64_t>(), A->pdofs.as<int32_t>(), A->rp.as<int64_t>(),
                       A->col.as<int32_t>(), A->val.as<double>(), A->pinv.as<double>());
    HIPCHK(A->ctx, hipGetLastError());
    HIPCHK(A->ctx, hipMemsetAsync(A->bad.p, 0, sizeof(int), s));
    hipLaunchKernelGGL(k_patch_invert, dim3((unsigned)A->npatch), dim3(64), 0, s, A->npatch, A->poff.as<int64_t>(),
                       A->pptr.as<int64_t>(), A->poptr.as<int64_t>(), A->porow.as<int32_t>(), A->pqoff.as<int64_t>(),
                       A->pinv.as<double>(), A->pown.as<double>(), A->bad.as<int>());
    HIPCHK(A->ctx, hipGetLastError());
    int pbad = 0;
    HIPCHK(A->ctx, hipMemcpyAsync(&pbad, A->bad.p, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(A->ctx, hipStreamSynchronize(s));
    bad += pbad;
  }
  if (nsingular) *nsingular = bad;
  A->values = true;
  return HFK_OK;
}

int hfk_sys_set_patches(hfk_sys* A, int64_t npatch, const int64_t* ptr, const int32_t* dofs, const int8_t* owned) {
  int rc = sys_bind(A);
  if (rc) return rc;
  if (npatch < 0 || (npatch > 0 && (!ptr || !dofs || !owned)) || npatch >= ((int64_t)1 << 31))
    return set_err(A->ctx, HFK_EARG, "HyperFEMKernels: hfk_sys_set_patches: bad arguments");
  std::vector<int64_t> off(npatch + 1, 0), optr(npatch + 1, 0), qoff(npatch + 1, 0);
  std::vector<int32_t> orow;
  std::vector<uint8_t> own(A->n, 0);
  if (npatch > 0 && ptr[0] != 0) return set_err(A->ctx, HFK_EARG, "HyperFEMKernels: hfk_sys_set_patches: ptr[0] must be 0");
  for (int64_t p = 0; p < npatch; ++p) {
    const int64_t m = ptr[p + 1] - ptr[p];
    if (m < 1 || m > kMaxPatch)
      return set_err(A->ctx, HFK_EARG, "HyperFEMKernels: hfk_sys_set_patches: a patch must hold 1 to 80 dofs");
    std::vector<int32_t> seen(dofs + ptr[p], dofs + ptr[p + 1]);
    std::sort(seen.begin(), seen.end());
    if (std::adjacent_find(seen.begin(), seen.end()) != seen.end())
      return set_err(A->ctx, HFK_EARG, "HyperFEMKernels: hfk_sys_set_patches: a dof twice in one patch");
    for (int64_t k = ptr[p]; k < ptr[p + 1]; ++k) {
      if (dofs[k] < 1 || dofs[k] > A->n) return set_err(A->ctx, HFK_EARG, "HyperFEMKernels: hfk_sys_set_patches: dof out of range");
      if (owned[k]) {
        if (own[dofs[k] - 1]) return set_err(A->ctx, HFK_EARG, "HyperFEMKernels: hfk_sys_set_patches: a dof owned twice");
        own[dofs[k] - 1] = 1;
        orow.push_back((int32_t)(k - ptr[p]));
      }
    }
    off[p + 1] = off[p] + m * m;
    optr[p + 1] = (int64_t)orow.size();
    qoff[p + 1] = qoff[p] + (optr[p + 1] - optr[p]) * m;
  }
  if (npatch > 0 && std::find(own.begin(), own.end(), 0) != own.end())
    return set_err(A->ctx, HFK_EARG, "HyperFEMKernels: hfk_sys_set_patches: every dof must be owned by one patch");
  A->npatch = npatch;
  A->ptotal = off[npatch];
  if (npatch == 0) return HFK_OK;
  const int64_t nd = ptr[npatch];
  HIPCHK(A->ctx, A->pptr.alloc((npatch + 1) * sizeof(int64_t)));
  HIPCHK(A->ctx, A->poff.alloc((npatch + 1) * sizeof(int64_t)));
  HIPCHK(A->ctx, A->pdofs.alloc(nd * sizeof(int32_t)));
  HIPCHK(A->ctx, A->powned.alloc(nd));
  HIPCHK(A->ctx, A->pinv.alloc(A->ptotal * sizeof(double)));
  HIPCHK(A->ctx, A->poptr.alloc((npatch + 1) * sizeof(int64_t)));
  HIPCHK(A->ctx, A->pqoff.alloc((npatch + 1) * sizeof(int64_t)));
  HIPCHK(A->ctx, A->porow.alloc(std::max<size_t>(orow.size(), 1) * sizeof(int32_t)));
  HIPCHK(A->ctx, A->pown.alloc(std::max<int64_t>(qoff[npatch], 1) * sizeof(double)));
  HIPCHK(A->ctx, hipMemcpy(A->poptr.p, optr.data(), (npatch + 1) * sizeof(int64_t), hipMemcpyHostToDevice));
  HIPCHK(A->ctx, hipMemcpy(A->pqoff.p, qoff.data(), (npatch + 1) * sizeof(int64_t), hipMemcpyHostToDevice));
  HIPCHK(A->ctx, hipMemcpy(A->porow.p, orow.data(), orow.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  HIPCHK(A->ctx, hipMemcpy(A->pptr.p, ptr, (npatch + 1) * sizeof(int64_t), hipMemcpyHostToDevice));
  HIPCHK(A->ctx, hipMemcpy(A->poff.p, off.data(), (npatch + 1) * sizeof(int64_t), hipMemcpyHostToDevice));
  HIPCHK(A->ctx, hipMemcpy(A->pdofs.p, dofs, nd * sizeof(int32_t), hipMemcpyHostToDevice));
  HIPCHK(A->ctx, hipMemcpy(A->powned.p, owned, nd, hipMemcpyHostToDevice));
  A->values = false;  // the patches of the current values are extracted by the next update
  return HFK_OK;
}

int hfk_sys_patch_inverses(hfk_sys* A, double* Pinv) {
  int rc = sys_bind(A);
  if (rc) return rc;
  if (!A->values || A->npatch == 0) return set_err(A->ctx, HFK_ESTATE, "HyperFEMKernels: hfk_sys_patch_inverses: no patches extracted");
  HIPCHK(A->ctx, hipStreamSynchronize(A->ctx->stream));
  HIPCHK(A->ctx, hipMemcpy(Pinv, A->pinv.p, A->ptotal * sizeof(double), hipMemcpyDeviceToHost));
  return HFK_OK;
}

int hfk_sys_patches_apply_dev(hfk_sys* A, const double* r_dev, double* z_dev, void* stream) {
  int rc = sys_bind(A);
  if (rc) return rc;
  if (!r_dev || !z_dev) return set_err(A->ctx, HFK_EARG, "HyperFEMKernels: hfk_sys_patches_apply_dev: null vector");
  return sys_patches(A, r_dev, 1.0, z_dev, A->stream(stream));
}

int hfk_sys_matvec_dev(hfk_sys* A, const double* x_dev, double* y_dev, void* stream) {
  int rc = sys_bind(A);
  if (rc) return rc;
  if (!x_dev || !y_dev) return set_err(A->ctx, HFK_EARG, "HyperFEMKernels: hfk_sys_matvec_dev: null vector");
  return sys_matvec(A, x_dev, nullptr, y_dev, A->stream(stream));
}

int hfk_sys_blocks_apply_dev(hfk_sys* A, const double* r_dev, double* z_dev, void* stream) {
  int rc = sys_bind(A);
  if (rc) return rc;
  if (!r_dev || !z_dev) return set_err(A->ctx, HFK_EARG, "HyperFEMKernels: hfk_sys_blocks_apply_dev: null vector");
  return sys_blocks(A, r_dev, 1.0, false, z_dev, A->stream(stream));
}

int hfk_sys_smooth_dev(hfk_sys* A, int sweeps, double omega, const double* r_dev, double* z_dev, void* stream) {
  int rc = sys_bind(A);
  if (rc) return rc;
  if (!r_dev || !z_dev || sweeps < 1 || !(omega > 0.0 && omega < 2.0))
    return set_err(A->ctx, HFK_EARG, "HyperFEMKernels: hfk_sys_smooth_dev: bad arguments (sweeps >= 1, 0 < omega < 2)");
  return sys_sweeps(A, sweeps, omega, r_dev, z_dev, A->stream(stream));
}

int hfk_sys_blocks(hfk_sys* A, double* B, double* Binv) {
  int rc = sys_bind(A);
  if (rc) return rc;
  if (!A->values || A->nblocks == 0) return set_err(A->ctx, HFK_ESTATE, "HyperFEMKernels: hfk_sys_blocks: no blocks extracted");
  HIPCHK(A->ctx, hipStreamSynchronize(A->ctx->stream));
  if (B) HIPCHK(A->ctx, hipMemcpy(B, A->B.p, A->nblocks * kBS * kBS * sizeof(double), hipMemcpyDeviceToHost));
  if (Binv) HIPCHK(A->ctx, hipMemcpy(Binv, A->Binv.p, A->nblocks * kBS * kBS * sizeof(double), hipMemcpyDeviceToHost));
  return HFK_OK;
}

int hfk_sys_solve_dev(hfk_sys* A, int kind, int precond, int sweeps, double omega, const double* b_dev, double* x_dev,
                      double rtol, double atol, int maxiter, int restart, int* status, double* rnorm, void* stream) {
  int rc = sys_bind(A);
  if (rc) return rc;
  if (!b_dev || !x_dev || !status || !rnorm) return set_err(A->ctx, HFK_EARG, "HyperFEMKernels: hfk_sys_solve_dev: null argument");
  if (kind != 2 && kind != 3)
    return set_err(A->ctx, HFK_EARG, "HyperFEMKernels: hfk_sys_solve_dev: kind must be 2 (GMRES) or 3 (FGMRES)");
  if (precond < 0 || precond > 3 || restart < 1 || maxiter < 0 ||
      (precond == 2 && (sweeps < 1 || !(omega > 0.0 && omega < 2.0))))
    return set_err(A->ctx, HFK_EARG, "HyperFEMKernels: hfk_sys_solve_dev: bad preconditioner or iteration arguments");
  if (!A->values) return set_err(A->ctx, HFK_ESTATE, "HyperFEMKernels: assembled system used before its values were set");
  if ((precond == 1 || precond == 2) && A->nblocks == 0)
    return set_err(A->ctx, HFK_ESTATE, "HyperFEMKernels: no blocks set on the assembled system");
  if (precond == 3 && A->npatch == 0) return set_err(A->ctx, HFK_ESTATE, "HyperFEMKernels: no patches set on the assembled system");
  const hipStream_t s = A->stream(stream);
  const int64_t n = A->n;
  auto Ap = [&](const double* v, double* y) -> int { return sys_matvec(A, v, nullptr, y, s); };
  auto M = [&](const double* r, double* z) -> int {
    if (precond == 0) {
      HIPCHK(A->ctx, vec_copy(n, r, z, s));
      return HFK_OK;
    }
    if (precond == 1) return sys_blocks(A, r, 1.0, false, z, s);
    if (precond == 3) return sys_patches(A, r, 1.0, z, s);
    return sys_sweeps(A, sweeps, omega, r, z, s);
  };
  NoWatch watch;
  return gmres_core(A->ctx, n, Ap, M, watch, kind == 3, false, b_dev, x_dev, rtol, atol, maxiter, restart, status,
                    rnorm, s);
}

}  // extern "C"
