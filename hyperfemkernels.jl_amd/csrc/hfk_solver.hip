// Device kernels of the solver layer: vector updates with deterministic reductions, the nodal block Jacobi
// (invert/apply), the fused Chebyshev step, the p-multigrid transfers and the CSR product of the assembled coarse
// level.  Host orchestration is in hfk_pmg.hip.
#include <hip/hip_runtime.h>

#include <algorithm>

#include <cmath>

#include "hfk_solver.hpp"

namespace hfk {
namespace {

constexpr int kRedBlocks = 512;  // fixed grid of the reductions: the summation order never depends on the device
constexpr int kRedThreads = 256;

__device__ __forceinline__ double block_sum(double v, double* sh) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (t < s) sh[t] += sh[t + s];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

// partial[b] = sum over the grid-stride chunk of block b of a[i]*b[i] (up to 3 dot products at once)
__global__ __launch_bounds__(kRedThreads) void k_dot_partial(int64_t n, const double* __restrict__ a0,
                                                             const double* __restrict__ b0, const double* __restrict__ a1,
                                                             const double* __restrict__ b1, const double* __restrict__ a2,
                                                             const double* __restrict__ b2, double* __restrict__ part) {
  __shared__ double sh[kRedThreads];
  double s0 = 0, s1 = 0, s2 = 0;
  for (int64_t i = blockIdx.x * (int64_t)kRedThreads + threadIdx.x; i < n; i += (int64_t)kRedBlocks * kRedThreads) {
    s0 += a0[i] * b0[i];
    if (a1) s1 += a1[i] * b1[i];
    if (a2) s2 += a2[i] * b2[i];
  }
  s0 = block_sum(s0, sh);
  if (a1) s1 = block_sum(s1, sh);
  if (a2) s2 = block_sum(s2, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = s0;
    part[kRedBlocks + blockIdx.x] = s1;
    part[2 * kRedBlocks + blockIdx.x] = s2;
  }
}

// out[k] = sum_b part[k*kRedBlocks + b], in a fixed tree order
__global__ __launch_bounds__(kRedThreads) void k_dot_final(const double* __restrict__ part, int ndots,
                                                           double* __restrict__ out) {
  __shared__ double sh[kRedThreads];
  for (int k = 0; k < ndots; ++k) {
    double v = part[k * kRedBlocks + threadIdx.x] + part[k * kRedBlocks + threadIdx.x + kRedThreads];
    v = block_sum(v, sh);
    if (threadIdx.x == 0) out[k] = v;
  }
}

// Gram-Schmidt helpers of GMRES: partial sums of the K dot products V[k] . w (k = 8 tile + t), tile by tile;
// part[k][kRedBlocks]
__global__ __launch_bounds__(kRedThreads) void k_multi_dot_partial(int64_t n, const double* __restrict__ V, int64_t ldv,
                                                                   int K, const double* __restrict__ w,
                                                                   double* __restrict__ part) {
  __shared__ double sh[kRedThreads];
  const int k0 = 8 * blockIdx.y;
  double s[8];
#pragma unroll
  for (int t = 0; t < 8; ++t) s[t] = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)kRedThreads + threadIdx.x; i < n; i += (int64_t)kRedBlocks * kRedThreads) {
    const double wi = w[i];
#pragma unroll
    for (int t = 0; t < 8; ++t)
      if (k0 + t < K) s[t] += V[(size_t)(k0 + t) * ldv + i] * wi;
  }
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    if (k0 + t >= K) break;  // uniform over the block
    const double v = block_sum(s[t], sh);
    if (threadIdx.x == 0) part[(size_t)(k0 + t) * kRedBlocks + blockIdx.x] = v;
  }
}

// out[k] = sum_b part[k][b] (fixed tree order), one block per dot product
__global__ __launch_bounds__(kRedThreads) void k_multi_dot_final(const double* __restrict__ part, double* __restrict__ out) {
  __shared__ double sh[kRedThreads];
  const int k = blockIdx.x;
  double v = part[(size_t)k * kRedBlocks + threadIdx.x] + part[(size_t)k * kRedBlocks + threadIdx.x + kRedThreads];
  v = block_sum(v, sh);
  if (threadIdx.x == 0) out[k] = v;
}

// w += alpha sum_k coef[k] V[k]
__global__ void k_multi_axpy(int64_t n, const double* __restrict__ V, int64_t ldv, int K, const double* __restrict__ coef,
                             double alpha, double* __restrict__ w) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int k = 0; k < K; ++k) s += coef[k] * V[(size_t)k * ldv + i];
    w[i] += alpha * s;
  }
}

// Arnoldi step j (classical Gram-Schmidt applied twice): column j of the Hessenberg matrix, H[i + ldh j] =
// c1[i] + c2[i] for i <= j and H[j + 1 + ldh j] = sqrt(ww), kept on the device
__global__ void k_arnoldi_column(int j, int ldh, const double* __restrict__ c1, const double* __restrict__ c2,
                                 const double* __restrict__ ww, double* __restrict__ H) {
  const int i = threadIdx.x;
  if (i <= j) H[i + (size_t)ldh * j] = c1[i] + c2[i];
  if (i == 0) H[j + 1 + (size_t)ldh * j] = sqrt(ww[0]);
}

// w /= sqrt(ww) (a zero or non-finite norm leaves w non-finite: the host truncates the Arnoldi process there)
__global__ void k_scale_inv_sqrt(int64_t n, const double* __restrict__ ww, double* __restrict__ w) {
  const double sc = 1.0 / sqrt(ww[0]);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    w[i] = sc * w[i];
}

__global__ void k_axpby(int64_t n, double a, const double* __restrict__ x, double b, double* __restrict__ y) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = a * x[i] + b * y[i];
}

__global__ void k_mul(int64_t n, const double* __restrict__ a, const double* x, double* y) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = a[i] * x[i];
}

// d[c] = 1 / sqrt(A[c][c]) for a positive finite diagonal entry, 1 otherwise (CSC, sorted rows)
__global__ void k_csc_diag_rsqrt(int64_t ncols, const int32_t* __restrict__ colptr, const int32_t* __restrict__ rowval,
                                 const double* __restrict__ val, double* __restrict__ d) {
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < ncols; c += (int64_t)gridDim.x * blockDim.x) {
    int lo = colptr[c], hi = colptr[c + 1] - 1;
    double a = 0.0;
    while (lo <= hi) {
      const int mid = (lo + hi) >> 1;
      const int r = rowval[mid];
      if (r == (int)c) {
        a = val[mid];
        break;
      }
      if (r < (int)c) lo = mid + 1;
      else hi = mid - 1;
    }
    d[c] = (a > 0.0 && isfinite(a)) ? 1.0 / sqrt(a) : 1.0;
  }
}

__global__ void k_copy(int64_t n, const double* __restrict__ x, double* __restrict__ y) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) y[i] = x[i];
}

// CG step on device scalars s[]: alpha = s[RZ]/s[PQ];  x += alpha p;  r -= alpha q
__global__ void k_cg_update(int64_t n, const double* __restrict__ s, const double* __restrict__ p,
                            const double* __restrict__ q, double* __restrict__ x, double* __restrict__ r) {
  const double alpha = s[kRZ] / s[kPQ];
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    x[i] += alpha * p[i];
    r[i] -= alpha * q[i];
  }
}

// p = z + beta p: beta = s[RZN]/s[RZ] (CG) or -s[ZQ]/s[PQ] (flexible CG); then s[RZ] = s[RZN] (one thread)
__global__ void k_cg_direction(int64_t n, double* __restrict__ s, const double* __restrict__ z, double* __restrict__ p,
                               int flexible) {
  const double beta = flexible ? -s[kZQ] / s[kPQ] : s[kRZN] / s[kRZ];
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = z[i] + beta * p[i];
}
__global__ void k_scalar_copy(double* s, int dst, int src) { s[dst] = s[src]; }

// --- nodal blocks ------------------------------------------------------------------------------------------
__device__ __forceinline__ double det3d(const double* a) {
  return a[0] * (a[4] * a[8] - a[7] * a[5]) - a[3] * (a[1] * a[8] - a[7] * a[2]) + a[6] * (a[1] * a[5] - a[4] * a[2]);
}

// _invert_blocks! (src/preconditioners.jl:292-307): fixed rows/columns -> identity; a non-finite block is kept
// as it is and counted in bad[0].  The inverses are stored entry-major, Binv[(i + 3j) * nb + k], so that the
// block step reads each entry with one coalesced access per wave.
__global__ void k_invert_blocks(int64_t nb, const int32_t* __restrict__ dofs, const double* __restrict__ B,
                                double* __restrict__ Binv, int* __restrict__ bad) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < nb; k += (int64_t)gridDim.x * blockDim.x) {
    double M[9];
    bool fin = true;
    for (int j = 0; j < 3; ++j)
      for (int i = 0; i < 3; ++i) {
        const bool fixed = dofs[3 * k + i] == 0 || dofs[3 * k + j] == 0;
        M[i + 3 * j] = fixed ? (i == j ? 1.0 : 0.0) : B[9 * k + i + 3 * j];
        fin = fin && isfinite(M[i + 3 * j]);
      }
    double R[9];
    if (fin) {
      const double id = 1.0 / det3d(M);
      R[0] = (M[4] * M[8] - M[7] * M[5]) * id;
      R[3] = -(M[3] * M[8] - M[6] * M[5]) * id;
      R[6] = (M[3] * M[7] - M[6] * M[4]) * id;
      R[1] = -(M[1] * M[8] - M[7] * M[2]) * id;
      R[4] = (M[0] * M[8] - M[6] * M[2]) * id;
      R[7] = -(M[0] * M[7] - M[6] * M[1]) * id;
      R[2] = (M[1] * M[5] - M[4] * M[2]) * id;
      R[5] = -(M[0] * M[5] - M[3] * M[2]) * id;
      R[8] = (M[0] * M[4] - M[3] * M[1]) * id;
    } else {
      for (int e = 0; e < 9; ++e) R[e] = M[e];
    }
    bool ok = true;
    for (int e = 0; e < 9; ++e) {
      Binv[(size_t)e * nb + k] = R[e];
      ok = ok && isfinite(R[e]);
    }
    if (!ok) atomicAdd(bad, 1);
  }
}

// _apply_blocks! (src/preconditioners.jl:327-342) fused with the Chebyshev direction update
// (src/smoothers.jl:130-141):  z = Binv r on the block;  d = cd d + cz z;  x += d (if x)
// With cd = 0, cz = 1 and x = null this is ldiv!(y = d, blocks, r).
template <class BT>
__global__ void k_block_step(int64_t nb, const int32_t* __restrict__ dofs, const BT* __restrict__ Binv,
                             const double* __restrict__ r, double* __restrict__ d, double* __restrict__ x, double cd,
                             double cz) {
  // every load of a block is issued before the arithmetic (ids, then r, d, x and the inverse), so that their
  // latencies overlap
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < nb; k += (int64_t)gridDim.x * blockDim.x) {
    int id[3];
    double rv[3], dv[3], xv[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) id[j] = dofs[3 * k + j];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const bool f = id[j] > 0;
      const int64_t o = f ? id[j] - 1 : 0;
      rv[j] = f ? r[o] : 0.0;
      dv[j] = f && cd != 0.0 ? d[o] : 0.0;
      xv[j] = f && x ? x[o] : 0.0;
    }
    double b[9];
#pragma unroll
    for (int e = 0; e < 9; ++e) b[e] = static_cast<double>(Binv[(size_t)e * nb + k]);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      if (id[i] <= 0) continue;
      double z = 0;
#pragma unroll
      for (int j = 0; j < 3; ++j)
        if (id[j] > 0) z += b[i + 3 * j] * rv[j];
      const double dn = (cd == 0.0 ? 0.0 : cd * dv[i]) + cz * z;
      d[id[i] - 1] = dn;
      if (x) x[id[i] - 1] = xv[i] + dn;
    }
  }
}

// nodal blocks of an assembled matrix (CSC of A, 0-based): B[i + 3j] = A[dof_i, dof_j]
__global__ void k_csr_blocks(int64_t nb, const int32_t* __restrict__ dofs, const int32_t* __restrict__ colptr,
                             const int32_t* __restrict__ rowval, const double* __restrict__ val, double* __restrict__ B) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < nb; k += (int64_t)gridDim.x * blockDim.x) {
    for (int j = 0; j < 3; ++j)
      for (int i = 0; i < 3; ++i) {
        const int ci = dofs[3 * k + i], cj = dofs[3 * k + j];
        double v = 0;
        if (ci > 0 && cj > 0) {
          int lo = colptr[cj - 1], hi = colptr[cj];
          while (lo < hi) {  // first row >= ci
            const int mid = (lo + hi) >> 1;
            if (rowval[mid] < ci - 1) lo = mid + 1; else hi = mid;
          }
          if (lo < colptr[cj] && rowval[lo] == ci - 1) v = val[lo];
        }
        B[9 * k + i + 3 * j] = v;
      }
  }
}

// y = alpha * A' x + beta * y over the columns of the CSC of A (= rows of A' ; A symmetric at the coarse level),
// 16 lanes per column
__global__ __launch_bounds__(256) void k_csc_t_spmv(int64_t ncols, const int32_t* __restrict__ colptr,
                                                    const int32_t* __restrict__ rowval, const double* __restrict__ val,
                                                    const double* __restrict__ x, double* __restrict__ y, double alpha,
                                                    double beta) {
  const int lane = threadIdx.x & 15;
  for (int64_t c = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 4; c < ncols;
       c += ((int64_t)gridDim.x * blockDim.x) >> 4) {
    double s = 0;
    for (int p = colptr[c] + lane; p < colptr[c + 1]; p += 16) s += val[p] * x[rowval[p]];
    for (int o = 8; o > 0; o >>= 1) s += __shfl_xor(s, o, 16);
    if (lane == 0) y[c] = alpha * s + (beta == 0.0 ? 0.0 : beta * y[c]);
  }
}

// y = alpha A x + beta y with A in node-block ELL form: block row i (the up to 3 free dofs of node block i) has W
// slots of 3x3 blocks; slot k couples it with block nbr[k nb + i] (padding slots point at i and hold zeros) and
// holds A[dof(i,ci), dof(j,cj)] at val[(9k + ci + 3cj) nb + i]: one thread per block row, coalesced values, one
// index per 9 values (the scalar CSC above reads one index per value)
template <class TV>
__global__ __launch_bounds__(256) void k_bell_spmv(int64_t nb, int W, const int32_t* __restrict__ nbr,
                                                   const TV* __restrict__ val, const int32_t* __restrict__ dofs,
                                                   const double* __restrict__ x, double* __restrict__ y, double alpha,
                                                   double beta) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nb; i += (int64_t)gridDim.x * blockDim.x) {
    double s0 = 0, s1 = 0, s2 = 0;
    for (int k = 0; k < W; ++k) {
      const int64_t j = nbr[(size_t)k * nb + i];
      const int d0 = dofs[3 * j], d1 = dofs[3 * j + 1], d2 = dofs[3 * j + 2];
      const double x0 = d0 > 0 ? x[d0 - 1] : 0.0, x1 = d1 > 0 ? x[d1 - 1] : 0.0, x2 = d2 > 0 ? x[d2 - 1] : 0.0;
      const TV* v = val + (size_t)9 * k * nb + i;
      s0 += (double)v[0] * x0 + (double)v[3 * nb] * x1 + (double)v[6 * nb] * x2;
      s1 += (double)v[nb] * x0 + (double)v[4 * nb] * x1 + (double)v[7 * nb] * x2;
      s2 += (double)v[2 * nb] * x0 + (double)v[5 * nb] * x1 + (double)v[8 * nb] * x2;
    }
    const int o[3] = {dofs[3 * i], dofs[3 * i + 1], dofs[3 * i + 2]};
    const double sv[3] = {s0, s1, s2};
#pragma unroll
    for (int c = 0; c < 3; ++c)
      if (o[c] > 0) y[o[c] - 1] = alpha * sv[c] + (beta == 0.0 ? 0.0 : beta * y[o[c] - 1]);
  }
}

// ---- block-Jacobi PCG on the node-block ELL matrix in three launches per iteration ---------------------------
// (1) q = A p per block row, with per-work-group partials of p.q;  (2) alpha = rz / sum(p.q partials), x += alpha p,
// r -= alpha q, z = Binv r per node block, with per-work-group partials of r.z and r.r;  (3) beta = sum(r.z
// partials) / rz, p = z + beta p, the new rz into the other of two device slots and r.r for the convergence test.
// One thread per node block, G = ceil(nb / 256) work-groups; every work-group sums the partials of the previous
// launch itself, in the same fixed order, so alpha and beta agree across work-groups and runs.  Replaces the nine
// launches (product, two dot products with their final sums, update, block step, direction, scalar copy) of the
// generic loop.
__device__ __forceinline__ double sum_partials(const double* __restrict__ part, int G, double* sh) {
  double v = 0.0;
  for (int b = threadIdx.x; b < G; b += blockDim.x) v += part[b];
  return block_sum(v, sh);
}

template <class TV>
__global__ __launch_bounds__(256) void k_pcg_spmv(int64_t nb, int W, const int32_t* __restrict__ nbr,
                                                  const TV* __restrict__ val, const int32_t* __restrict__ dofs,
                                                  const double* __restrict__ p, double* __restrict__ q,
                                                  double* __restrict__ part) {
  __shared__ double sh[256];
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  double pq = 0.0;
  if (i < nb) {
    double s0 = 0, s1 = 0, s2 = 0;
    for (int k = 0; k < W; ++k) {
      const int64_t j = nbr[(size_t)k * nb + i];
      const int d0 = dofs[3 * j], d1 = dofs[3 * j + 1], d2 = dofs[3 * j + 2];
      const double x0 = d0 > 0 ? p[d0 - 1] : 0.0, x1 = d1 > 0 ? p[d1 - 1] : 0.0, x2 = d2 > 0 ? p[d2 - 1] : 0.0;
      const TV* v = val + (size_t)9 * k * nb + i;
      s0 += (double)v[0] * x0 + (double)v[3 * nb] * x1 + (double)v[6 * nb] * x2;
      s1 += (double)v[nb] * x0 + (double)v[4 * nb] * x1 + (double)v[7 * nb] * x2;
      s2 += (double)v[2 * nb] * x0 + (double)v[5 * nb] * x1 + (double)v[8 * nb] * x2;
    }
    const int o[3] = {dofs[3 * i], dofs[3 * i + 1], dofs[3 * i + 2]};
    const double sv[3] = {s0, s1, s2};
#pragma unroll
    for (int c = 0; c < 3; ++c)
      if (o[c] > 0) {
        q[o[c] - 1] = sv[c];
        pq += p[o[c] - 1] * sv[c];
      }
  }
  pq = block_sum(pq, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = pq;
}

__global__ __launch_bounds__(256) void k_pcg_update(int64_t nb, int G, const int32_t* __restrict__ dofs,
                                                    const double* __restrict__ Binv, const double* __restrict__ p,
                                                    const double* __restrict__ q, double* __restrict__ x,
                                                    double* __restrict__ r, double* __restrict__ z,
                                                    const double* __restrict__ sc, int slot, double* __restrict__ part) {
  __shared__ double sh[256];
  const double alpha = sc[slot] / sum_partials(part, G, sh);
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  double rz = 0.0, rr = 0.0;
  if (i < nb) {
    int id[3];
    double rv[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      id[j] = dofs[3 * i + j];
      rv[j] = 0.0;
      if (id[j] > 0) {
        x[id[j] - 1] += alpha * p[id[j] - 1];
        rv[j] = r[id[j] - 1] - alpha * q[id[j] - 1];
        r[id[j] - 1] = rv[j];
      }
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      if (id[a] <= 0) continue;
      double zz = 0.0;
#pragma unroll
      for (int j = 0; j < 3; ++j)
        if (id[j] > 0) zz += Binv[(size_t)(a + 3 * j) * nb + i] * rv[j];
      z[id[a] - 1] = zz;
      rz += rv[a] * zz;
      rr += rv[a] * rv[a];
    }
  }
  rz = block_sum(rz, sh);
  rr = block_sum(rr, sh);
  if (threadIdx.x == 0) {
    part[G + blockIdx.x] = rz;
    part[2 * G + blockIdx.x] = rr;
  }
}

__global__ __launch_bounds__(256) void k_pcg_direction(int64_t nb, int G, const int32_t* __restrict__ dofs,
                                                       const double* __restrict__ z, double* __restrict__ p,
                                                       double* __restrict__ sc, int slot, int next,
                                                       const double* __restrict__ part) {
  __shared__ double sh[256];
  const double rzn = sum_partials(part + G, G, sh);
  const double beta = rzn / sc[slot];
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < nb) {
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const int id = dofs[3 * i + j];
      if (id > 0) p[id - 1] = z[id - 1] + beta * p[id - 1];
    }
  }
  if (blockIdx.x == 0) {  // uniform over the work-group
    const double rr = sum_partials(part + 2 * G, G, sh);
    if (threadIdx.x == 0) {
      sc[next] = rzn;
      sc[kRR] = rr;
    }
  }
}

// the block-ELL values from the CSC values: v64[pos[p]] = val[p] (and the FP32 copy when v32 is given)
__global__ void k_bell_fill(int64_t nnz, const int32_t* __restrict__ pos, const double* __restrict__ val,
                            double* __restrict__ v64, float* __restrict__ v32) {
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < nnz; p += (int64_t)gridDim.x * blockDim.x) {
    v64[pos[p]] = val[p];
    if (v32) v32[pos[p]] = (float)val[p];
  }
}

// --- transfers (src/transfers.jl:188-221, 295-313) ---------------------------------------------------------
// fine ids: [slot][c][tf] (the fine operator's batch layout); coarse ids of the same cells: [slot][c][tc]
__global__ void k_prolongate(int64_t nslot, int NF, int NC, const int32_t* __restrict__ fid,
                             const int32_t* __restrict__ cid, const double* __restrict__ E, const double* __restrict__ xc,
                             double* __restrict__ yf) {
  const int64_t n = nslot * 3 * NF;
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < n; g += (int64_t)gridDim.x * blockDim.x) {
    const int id = fid[g];
    if (id <= 0) continue;
    const int64_t sc = g / NF;  // slot*3 + c
    const int tf = (int)(g - sc * NF);
    const int32_t* ci = cid + sc * NC;
    double v = 0;
    for (int tc = 0; tc < NC; ++tc) {
      const double e = E[tf * NC + tc];
      const int j = ci[tc];
      if (e != 0.0 && j > 0) v += e * xc[j - 1];
    }
    yf[id - 1] = v;  // insertion: the cells sharing a fine dof write the same value
  }
}

// The same with one thread per (slot, component): the NC coarse values are loaded once and every fine node of the
// cell is formed from them in the order above (bitwise the same values), instead of NF threads each loading them.
template <int NC>
__global__ void k_prolongate_cell(int64_t nslot, int NF, const int32_t* __restrict__ fid,
                                  const int32_t* __restrict__ cid, const double* __restrict__ E,
                                  const double* __restrict__ xc, double* __restrict__ yf) {
  const int64_t n = nslot * 3;
  for (int64_t sc = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; sc < n; sc += (int64_t)gridDim.x * blockDim.x) {
    const int32_t* ci = cid + sc * NC;
    int jc[NC];
    double xv[NC];
#pragma unroll
    for (int tc = 0; tc < NC; ++tc) jc[tc] = ci[tc];
#pragma unroll
    for (int tc = 0; tc < NC; ++tc) xv[tc] = jc[tc] > 0 ? xc[jc[tc] - 1] : 0.0;
    const int32_t* fi = fid + sc * NF;
    for (int tf = 0; tf < NF; ++tf) {
      const int id = fi[tf];
      if (id <= 0) continue;
      double v = 0;
#pragma unroll
      for (int tc = 0; tc < NC; ++tc) {
        const double e = E[tf * NC + tc];
        if (e != 0.0 && jc[tc] > 0) v += e * xv[tc];
      }
      yf[id - 1] = v;  // insertion: the cells sharing a fine dof write the same value
    }
  }
}

// The same with one thread per fine free dof, from the cell slot that owns it (owner = (slot * 3 + c) * NF + tf,
// -1 for a dof no cell holds): each fine value is written once, the sum in the order above (bitwise the same)
template <int NC>
__global__ void k_prolongate_owner(int64_t nf, int NF, const int64_t* __restrict__ owner,
                                   const int32_t* __restrict__ cid, const double* __restrict__ E,
                                   const double* __restrict__ xc, double* __restrict__ yf) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nf; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t o = owner[i];
    double v = 0;
    if (o >= 0) {
      const int64_t sc = o / NF;
      const int tf = (int)(o - sc * NF);
      const int32_t* ci = cid + sc * NC;
#pragma unroll
      for (int tc = 0; tc < NC; ++tc) {
        const double e = E[tf * NC + tc];
        const int j = ci[tc];
        if (e != 0.0 && j > 0) v += e * xc[j - 1];
      }
    }
    yf[i] = v;
  }
}

// restriction over the batches of one colour of the fine operator (cells of a colour share no fine free dof,
// hence no coarse free dof): yc[j] += sum_tf E[tf, tc] yf[id] / mult[id]
__global__ void k_restrict(int64_t slot0, int64_t nslot, int NF, int NC, const int32_t* __restrict__ fid,
                           const int32_t* __restrict__ cid, const double* __restrict__ E,
                           const double* __restrict__ invmult, const double* __restrict__ yf, double* __restrict__ yc) {
  const int64_t n = nslot * 3 * NC;
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < n; g += (int64_t)gridDim.x * blockDim.x) {
    const int64_t sc = slot0 * 3 + g / NC;
    const int tc = (int)(g % NC);
    const int j = cid[sc * NC + tc];
    if (j <= 0) continue;
    const int32_t* fi = fid + sc * NF;
    double v = 0;
    for (int tf = 0; tf < NF; ++tf) {
      const double e = E[tf * NC + tc];
      const int id = fi[tf];
      if (e != 0.0 && id > 0) v += e * yf[id - 1] * invmult[id - 1];
    }
    yc[j - 1] += v;
  }
}

// The same with one thread per (slot, component): every fine value of the cell is loaded once and added to the NC
// coarse sums in the order above (bitwise the same values), instead of NC threads each loading all of them.
template <int NC>
__global__ void k_restrict_cell(int64_t slot0, int64_t nslot, int NF, const int32_t* __restrict__ fid,
                                const int32_t* __restrict__ cid, const double* __restrict__ E,
                                const double* __restrict__ invmult, const double* __restrict__ yf,
                                double* __restrict__ yc) {
  const int64_t n = nslot * 3;
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < n; g += (int64_t)gridDim.x * blockDim.x) {
    const int64_t sc = slot0 * 3 + g;
    const int32_t* ci = cid + sc * NC;
    const int32_t* fi = fid + sc * NF;
    double v[NC];
#pragma unroll
    for (int tc = 0; tc < NC; ++tc) v[tc] = 0;
    for (int tf = 0; tf < NF; ++tf) {
      const int id = fi[tf];
      if (id <= 0) continue;
      const double yv = yf[id - 1], im = invmult[id - 1];
#pragma unroll
      for (int tc = 0; tc < NC; ++tc) {
        const double e = E[tf * NC + tc];
        if (e != 0.0) v[tc] += e * yv * im;
      }
    }
#pragma unroll
    for (int tc = 0; tc < NC; ++tc) {
      const int j = ci[tc];
      if (j > 0) yc[j - 1] += v[tc];
    }
  }
}

// nodal interpolation of the fine state (free xf, Dirichlet df) at the coarse nodes
__global__ void k_inject(int64_t nslot, int NF, int NC, const int32_t* __restrict__ fid, const int32_t* __restrict__ cid,
                         const double* __restrict__ Einj, const double* __restrict__ xf, const double* __restrict__ df,
                         double* __restrict__ xc, double* __restrict__ dc) {
  const int64_t n = nslot * 3 * NC;
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < n; g += (int64_t)gridDim.x * blockDim.x) {
    const int j = cid[g];
    if (j == 0) continue;
    const int64_t sc = g / NC;
    const int tc = (int)(g - sc * NC);
    const int32_t* fi = fid + sc * NF;
    double v = 0;
    for (int tf = 0; tf < NF; ++tf) {
      const double w = Einj[tc * NF + tf];
      if (w == 0.0) continue;
      const int id = fi[tf];
      v += w * (id > 0 ? xf[id - 1] : (id < 0 ? df[-id - 1] : 0.0));
    }
    if (j > 0) xc[j - 1] = v; else dc[-j - 1] = v;
  }
}

__global__ void k_count_nonfinite(int64_t n, const double* __restrict__ x, int* __restrict__ bad) {
  int local = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    local += isfinite(x[i]) ? 0 : 1;
  if (local) atomicAdd(bad, local);
}

inline dim3 grid_for(int64_t n, int threads = 256) {
  int64_t b = (n + threads - 1) / threads;
  if (b < 1) b = 1;
  if (b > 4096) b = 4096;
  return dim3((unsigned)b);
}

}  // namespace

hipError_t vec_dots(int64_t n, const double* a0, const double* b0, const double* a1, const double* b1, const double* a2,
                    const double* b2, double* part, double* out, hipStream_t s) {
  const int nd = a2 ? 3 : (a1 ? 2 : 1);
  hipLaunchKernelGGL(k_dot_partial, dim3(kRedBlocks), dim3(kRedThreads), 0, s, n, a0, b0, a1, b1, a2, b2, part);
  hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(kRedThreads), 0, s, part, nd, out);
  return hipGetLastError();
}
size_t dot_partial_doubles() { return 3 * kRedBlocks; }
size_t multi_dot_partial_doubles(int K) { return (size_t)K * kRedBlocks; }
hipError_t multi_dot(int64_t n, const double* V, int64_t ldv, int K, const double* w, double* part, double* out,
                     hipStream_t s) {
  hipLaunchKernelGGL(k_multi_dot_partial, dim3(kRedBlocks, (K + 7) / 8), dim3(kRedThreads), 0, s, n, V, ldv, K, w, part);
  hipLaunchKernelGGL(k_multi_dot_final, dim3(K), dim3(kRedThreads), 0, s, part, out);
  return hipGetLastError();
}
hipError_t multi_axpy(int64_t n, const double* V, int64_t ldv, int K, const double* coef, double alpha, double* w,
                      hipStream_t s) {
  hipLaunchKernelGGL(k_multi_axpy, grid_for(n), dim3(256), 0, s, n, V, ldv, K, coef, alpha, w);
  return hipGetLastError();
}

hipError_t arnoldi_column(int64_t n, int j, int ldh, const double* c1, const double* c2, const double* ww, double* H,
                          double* w, hipStream_t s) {
  hipLaunchKernelGGL(k_arnoldi_column, dim3(1), dim3((unsigned)(((j + 1 + 63) / 64) * 64)), 0, s, j, ldh, c1, c2, ww, H);
  hipLaunchKernelGGL(k_scale_inv_sqrt, grid_for(n), dim3(256), 0, s, n, ww, w);
  return hipGetLastError();
}

hipError_t vec_axpby(int64_t n, double a, const double* x, double b, double* y, hipStream_t s) {
  hipLaunchKernelGGL(k_axpby, grid_for(n), dim3(256), 0, s, n, a, x, b, y);
  return hipGetLastError();
}
hipError_t vec_mul(int64_t n, const double* a, const double* x, double* y, hipStream_t s) {
  hipLaunchKernelGGL(k_mul, grid_for(n), dim3(256), 0, s, n, a, x, y);
  return hipGetLastError();
}
hipError_t csc_diag_rsqrt(int64_t ncols, const int32_t* colptr, const int32_t* rowval, const double* val, double* d,
                          hipStream_t s) {
  hipLaunchKernelGGL(k_csc_diag_rsqrt, grid_for(ncols), dim3(256), 0, s, ncols, colptr, rowval, val, d);
  return hipGetLastError();
}
hipError_t vec_copy(int64_t n, const double* x, double* y, hipStream_t s) {
  hipLaunchKernelGGL(k_copy, grid_for(n), dim3(256), 0, s, n, x, y);
  return hipGetLastError();
}
hipError_t cg_update(int64_t n, const double* sc, const double* p, const double* q, double* x, double* r, hipStream_t s) {
  hipLaunchKernelGGL(k_cg_update, grid_for(n), dim3(256), 0, s, n, sc, p, q, x, r);
  return hipGetLastError();
}
hipError_t cg_direction(int64_t n, double* sc, const double* z, double* p, int flexible, hipStream_t s) {
  hipLaunchKernelGGL(k_cg_direction, grid_for(n), dim3(256), 0, s, n, sc, z, p, flexible);
  hipLaunchKernelGGL(k_scalar_copy, dim3(1), dim3(1), 0, s, sc, kRZ, kRZN);
  return hipGetLastError();
}
hipError_t blocks_invert(int64_t nb, const int32_t* dofs, const double* B, double* Binv, int* bad, hipStream_t s) {
  hipLaunchKernelGGL(k_invert_blocks, grid_for(nb), dim3(256), 0, s, nb, dofs, B, Binv, bad);
  return hipGetLastError();
}
hipError_t blocks_step(int64_t nb, const int32_t* dofs, const double* Binv, const double* r, double* d, double* x,
                       double cd, double cz, hipStream_t s) {
  hipLaunchKernelGGL(k_block_step<double>, grid_for(nb), dim3(256), 0, s, nb, dofs, Binv, r, d, x, cd, cz);
  return hipGetLastError();
}
hipError_t blocks_step(int64_t nb, const int32_t* dofs, const float* Binv, const double* r, double* d, double* x,
                       double cd, double cz, hipStream_t s) {
  hipLaunchKernelGGL(k_block_step<float>, grid_for(nb), dim3(256), 0, s, nb, dofs, Binv, r, d, x, cd, cz);
  return hipGetLastError();
}
__global__ void k_to_fp32(int64_t n, const double* __restrict__ a, float* __restrict__ b) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    b[i] = static_cast<float>(a[i]);
}
hipError_t vec_to_fp32(int64_t n, const double* a, float* b, hipStream_t s) {
  hipLaunchKernelGGL(k_to_fp32, grid_for(n), dim3(256), 0, s, n, a, b);
  return hipGetLastError();
}
hipError_t csr_blocks(int64_t nb, const int32_t* dofs, const int32_t* colptr, const int32_t* rowval, const double* val,
                      double* B, hipStream_t s) {
  hipLaunchKernelGGL(k_csr_blocks, grid_for(nb), dim3(256), 0, s, nb, dofs, colptr, rowval, val, B);
  return hipGetLastError();
}
hipError_t csc_t_spmv(int64_t ncols, const int32_t* colptr, const int32_t* rowval, const double* val, const double* x,
                      double* y, double alpha, double beta, hipStream_t s) {
  hipLaunchKernelGGL(k_csc_t_spmv, grid_for(ncols * 16), dim3(256), 0, s, ncols, colptr, rowval, val, x, y, alpha, beta);
  return hipGetLastError();
}
hipError_t bell_spmv(int64_t nb, int W, const int32_t* nbr, const void* val, bool fp32, const int32_t* dofs,
                     const double* x, double* y, double alpha, double beta, hipStream_t s) {
  if (fp32)
    hipLaunchKernelGGL(k_bell_spmv<float>, grid_for(nb), dim3(256), 0, s, nb, W, nbr, static_cast<const float*>(val), dofs,
                       x, y, alpha, beta);
  else
    hipLaunchKernelGGL(k_bell_spmv<double>, grid_for(nb), dim3(256), 0, s, nb, W, nbr, static_cast<const double*>(val),
                       dofs, x, y, alpha, beta);
  return hipGetLastError();
}
hipError_t pcg_fused_iteration(int64_t nb, int W, const int32_t* nbr, const void* val, bool fp32, const int32_t* dofs,
                               const double* Binv, double* x, double* r, double* z, double* p, double* q, double* sc,
                               int slot, int next, double* part, hipStream_t s) {
  const int G = (int)((nb + 255) / 256);
  if (fp32)
    hipLaunchKernelGGL(k_pcg_spmv<float>, dim3((unsigned)G), dim3(256), 0, s, nb, W, nbr, static_cast<const float*>(val),
                       dofs, p, q, part);
  else
    hipLaunchKernelGGL(k_pcg_spmv<double>, dim3((unsigned)G), dim3(256), 0, s, nb, W, nbr, static_cast<const double*>(val),
                       dofs, p, q, part);
  hipLaunchKernelGGL(k_pcg_update, dim3((unsigned)G), dim3(256), 0, s, nb, G, dofs, Binv, p, q, x, r, z, sc, slot, part);
  hipLaunchKernelGGL(k_pcg_direction, dim3((unsigned)G), dim3(256), 0, s, nb, G, dofs, z, p, sc, slot, next, part);
  return hipGetLastError();
}
hipError_t bell_fill(int64_t nnz, const int32_t* pos, const double* val, double* v64, float* v32, hipStream_t s) {
  hipLaunchKernelGGL(k_bell_fill, grid_for(nnz), dim3(256), 0, s, nnz, pos, val, v64, v32);
  return hipGetLastError();
}
hipError_t transfer_prolongate(int64_t nslot, int NF, int NC, const int32_t* fid, const int32_t* cid, const double* E,
                               const double* xc, double* yf, hipStream_t s) {
  if (NC == 8) hipLaunchKernelGGL(k_prolongate_cell<8>, grid_for(nslot * 3), dim3(256), 0, s, nslot, NF, fid, cid, E, xc, yf);
  else if (NC == 27) hipLaunchKernelGGL(k_prolongate_cell<27>, grid_for(nslot * 3), dim3(256), 0, s, nslot, NF, fid, cid, E, xc, yf);
  else hipLaunchKernelGGL(k_prolongate, grid_for(nslot * 3 * NF), dim3(256), 0, s, nslot, NF, NC, fid, cid, E, xc, yf);
  return hipGetLastError();
}
hipError_t transfer_prolongate_owner(int64_t nf, int NF, const int64_t* owner, const int32_t* cid, const double* E,
                                     const double* xc, double* yf, hipStream_t s) {
  if (nf <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_prolongate_owner<8>, grid_for(nf), dim3(256), 0, s, nf, NF, owner, cid, E, xc, yf);
  return hipGetLastError();
}
hipError_t transfer_restrict(int64_t slot0, int64_t nslot, int NF, int NC, const int32_t* fid, const int32_t* cid,
                             const double* E, const double* invmult, const double* yf, double* yc, hipStream_t s) {
  if (nslot <= 0) return hipSuccess;
  if (NC == 8) hipLaunchKernelGGL(k_restrict_cell<8>, grid_for(nslot * 3), dim3(256), 0, s, slot0, nslot, NF, fid, cid, E, invmult, yf, yc);
  else if (NC == 27) hipLaunchKernelGGL(k_restrict_cell<27>, grid_for(nslot * 3), dim3(256), 0, s, slot0, nslot, NF, fid, cid, E, invmult, yf, yc);
  else hipLaunchKernelGGL(k_restrict, grid_for(nslot * 3 * NC), dim3(256), 0, s, slot0, nslot, NF, NC, fid, cid, E, invmult, yf, yc);
  return hipGetLastError();
}
hipError_t transfer_inject(int64_t nslot, int NF, int NC, const int32_t* fid, const int32_t* cid, const double* Einj,
                           const double* xf, const double* df, double* xc, double* dc, hipStream_t s) {
  hipLaunchKernelGGL(k_inject, grid_for(nslot * 3 * NC), dim3(256), 0, s, nslot, NF, NC, fid, cid, Einj, xf, df, xc, dc);
  return hipGetLastError();
}
hipError_t count_nonfinite(int64_t n, const double* x, int* bad, hipStream_t s) {
  hipLaunchKernelGGL(k_count_nonfinite, grid_for(n), dim3(256), 0, s, n, x, bad);
  return hipGetLastError();
}


namespace {

__global__ void k_csr_spmv(int64_t nrows, const int64_t* __restrict__ rp, const int32_t* __restrict__ col,
                           const double* __restrict__ val, const double* __restrict__ x, double* __restrict__ y) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nrows; i += (int64_t)gridDim.x * blockDim.x) {
    double v = 0.0;
    for (int64_t k = rp[i]; k < rp[i + 1]; ++k) v += val[k] * x[col[k]];
    y[i] = v;
  }
}

// The same with L lanes per row (rows with many entries): lane l sums the entries rp[i] + l, rp[i] + l + L, ... in
// order, then the L partial sums are added by a fixed butterfly (deterministic); the L lanes read L consecutive
// entries per step
template <int L>
__global__ void k_csr_spmv_vec(int64_t nrows, const int64_t* __restrict__ rp, const int32_t* __restrict__ col,
                               const double* __restrict__ val, const double* __restrict__ x, double* __restrict__ y) {
  const int lane = threadIdx.x & (L - 1);
  const int64_t stride = (int64_t)gridDim.x * (blockDim.x / L);
  for (int64_t i = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / L; i < nrows; i += stride) {
    double v = 0.0;
    const int64_t e = rp[i + 1];
    for (int64_t k = rp[i] + lane; k < e; k += L) v += val[k] * x[col[k]];
#pragma unroll
    for (int o = L / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, L);
    if (lane == 0) y[i] = v;
  }
}

__global__ void k_inject_map(int64_t n, const int32_t* __restrict__ src, const double* __restrict__ xf,
                             const double* __restrict__ df, double* __restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t k = src[i];
    out[i] = k > 0 ? xf[k - 1] : df[-k - 1];
  }
}

// one work-group per row: a fixed-order tree over the row (deterministic)
__global__ void k_dense_matvec(int64_t n, const double* __restrict__ A, const double* __restrict__ x,
                               double* __restrict__ y) {
  __shared__ double red[256];
  const int64_t i = blockIdx.x;
  double v = 0.0;
  for (int64_t j = threadIdx.x; j < n; j += blockDim.x) v += A[i * n + j] * x[j];
  red[threadIdx.x] = v;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) y[i] = red[0];
}

int grid_of(int64_t n) { return (int)std::min<int64_t>((n + 255) / 256, 8192); }

}  // namespace

hipError_t csr_spmv(int64_t nrows, const int64_t* rp, const int32_t* col, const double* val, const double* x, double* y,
                    hipStream_t s) {
  if (nrows <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_csr_spmv, dim3(grid_of(nrows)), dim3(256), 0, s, nrows, rp, col, val, x, y);
  return hipGetLastError();
}
int csr_lanes(int64_t nrows, int64_t nnz) {
  const double avg = nrows > 0 ? (double)nnz / (double)nrows : 0.0;
  return avg <= 6 ? 1 : avg <= 12 ? 4 : avg <= 24 ? 8 : 16;
}
hipError_t csr_spmv_lanes(int64_t nrows, const int64_t* rp, const int32_t* col, const double* val, const double* x,
                          double* y, int lanes, hipStream_t s) {
  if (nrows <= 0) return hipSuccess;
  const dim3 g(grid_of(nrows * lanes)), t(256);
  switch (lanes) {
    case 1: hipLaunchKernelGGL(k_csr_spmv, g, t, 0, s, nrows, rp, col, val, x, y); break;
    case 4: hipLaunchKernelGGL(k_csr_spmv_vec<4>, g, t, 0, s, nrows, rp, col, val, x, y); break;
    case 8: hipLaunchKernelGGL(k_csr_spmv_vec<8>, g, t, 0, s, nrows, rp, col, val, x, y); break;
    case 16: hipLaunchKernelGGL(k_csr_spmv_vec<16>, g, t, 0, s, nrows, rp, col, val, x, y); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t inject_map(int64_t n, const int32_t* src, const double* xf, const double* df, double* out, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_inject_map, dim3(grid_of(n)), dim3(256), 0, s, n, src, xf, df, out);
  return hipGetLastError();
}

// the dense symmetric part of a CSC matrix (0-based): each entry adds half its value at (r, c) and at (c, r); every
// position receives at most two additions (a CSC pattern holds no duplicates), so the sum is order independent
__global__ void k_csc_dense_sym(int64_t n, const int32_t* __restrict__ colptr, const int32_t* __restrict__ rowval,
                                const double* __restrict__ val, double* __restrict__ A) {
  const int64_t c = blockIdx.x;
  for (int64_t p = colptr[c] + threadIdx.x; p < colptr[c + 1]; p += blockDim.x) {
    const int64_t r = rowval[p];
    const double v = 0.5 * val[p];
    atomicAdd(&A[r * n + c], v);
    atomicAdd(&A[c * n + r], v);
  }
}

// Gauss-Jordan inversion in place, one pivot per pair of launches: save row and column k, then
//   a_kk <- 1/p, a_kj <- a_kj/p, a_ik <- -a_ik/p, a_ij <- a_ij - a_ik (a_kj/p)   (p = a_kk, old values)
// No pivoting: the matrices are symmetric positive definite; a pivot that is not positive (or not finite) sets *bad.
__global__ void k_gj_save(int64_t n, int64_t k, const double* __restrict__ A, double* __restrict__ rk,
                          double* __restrict__ ck) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) {
    rk[i] = A[k * n + i];
    ck[i] = A[i * n + k];
  }
}

__global__ void k_gj_step(int64_t n, int64_t k, double* __restrict__ A, const double* __restrict__ rk,
                          const double* __restrict__ ck, int* __restrict__ bad) {
  const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t i = blockIdx.y;
  if (j >= n) return;
  const double p = rk[k];
  double* a = A + i * n + j;
  if (i == k && j == k) {
    if (!(p > 0) || !isfinite(p)) *bad = 1;
    *a = 1.0 / p;
  } else if (i == k) {
    *a = rk[j] / p;
  } else if (j == k) {
    *a = -ck[i] / p;
  } else {
    *a -= ck[i] * (rk[j] / p);
  }
}

__global__ void k_sym_part(int64_t n, const double* __restrict__ X, double* __restrict__ Y) {
  const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t i = blockIdx.y;
  if (j < n) Y[i * n + j] = 0.5 * (X[i * n + j] + X[j * n + i]);
}

hipError_t dense_spd_inverse(int64_t n, const int32_t* colptr, const int32_t* rowval, const double* val, double* W,
                             double* rk, double* ck, double* Ainv, int* bad, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipError_t e;
  if ((e = hipMemsetAsync(W, 0, (size_t)n * n * sizeof(double), s)) != hipSuccess) return e;
  hipLaunchKernelGGL(k_csc_dense_sym, dim3((unsigned)n), dim3(64), 0, s, n, colptr, rowval, val, W);
  const dim3 grid((unsigned)((n + 255) / 256), (unsigned)n);
  for (int64_t k = 0; k < n; ++k) {
    hipLaunchKernelGGL(k_gj_save, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, k, W, rk, ck);
    hipLaunchKernelGGL(k_gj_step, grid, dim3(256), 0, s, n, k, W, rk, ck, bad);
  }
  hipLaunchKernelGGL(k_sym_part, grid, dim3(256), 0, s, n, W, Ainv);
  return hipGetLastError();
}

hipError_t dense_matvec(int64_t n, const double* A, const double* x, double* y, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_dense_matvec, dim3((unsigned)n), dim3(256), 0, s, n, A, x, y);
  return hipGetLastError();
}

}  // namespace hfk
