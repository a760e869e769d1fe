// Device kernels of libhfk for gfx950 (CDNA4, wave64).
//
// Data layout (HBM).  Cells are processed in batches of C = 256 / P^3 cells (9 at Q2/P=3, 4 at Q3/P=4); the
// batches of one colour are contiguous, so one launch per colour is race-free (src/coloring.jl:1-3).  Per batch:
//   ids      [cell][c][t]          int32, tensor node order t = i + Q j + Q^2 k, 0 on padding cells
//   coords   [cell][v][i]          the 8 vertices of the trilinear map
//   tangents [entry][cell*P^3 + q] entry-major, so the 256 threads of the point phase read each entry of their
//                                  point with one fully coalesced access per wave (64 contiguous doubles)
// The stored tangent is D = dV iJt' T iJt (reference coordinates): the point phase needs neither the geometry
// nor dV, and the apply reads nothing per point but D (src/matrixfree.jl:366-375 multiplies T, iJt and dV at
// every point).
//
// Work-group structure (256 threads, one batch): gather -> 3 interpolation sweeps (LDS) -> point phase (one
// thread per (cell, point)) -> 3 contraction sweeps (LDS) -> scatter.  The sweeps are the 3D sum factorisation
// of src/tensorproduct.jl:125-159 (tp_interpolate!) and :189-224 (tp_contract!), one thread per line of the
// Q x Q x Q (or P x P x P) cube of one (cell, component).
#include <hip/hip_runtime.h>

#include <type_traits>

#include <cmath>
#include <cstdlib>

#include "hfk_internal.hpp"
#include "hfk_point.hpp"

namespace hfk {
namespace {

// ---------------------------------------------------------------------------------------------------------
// Batch of a work-group in the single-launch (atomic) mode.  The hardware hands work-group b to XCD b % 8;
// work-groups j = b / 8 of one XCD take the colours of one neighbourhood in turn: colour c = j % ncol of
// neighbourhood g = xcd * gx + j / ncol (the g-th batch of every colour covers neighbouring cells, which share
// nodes), so the dx entries they gather and the y entries they add to stay in that XCD's L2.  -1: idle.
__device__ __forceinline__ int interleaved_batch(const int* __restrict__ colfc, int ncol, int gx) {
  const int b = blockIdx.x, xcd = b & 7, j = b >> 3;
  const int c = j % ncol, g = xcd * gx + j / ncol;
  return g < colfc[2 * c + 1] ? colfc[2 * c] + g : -1;
}

// mul!(y, mf, dx) (src/matrixfree.jl:341-396): gather dx (0 at Dirichlet dofs), interpolate, sv = D su,
// contract, add to the free rows of y.  Deterministic mode (ATOMIC = false): the batches of one colour per
// launch, plain read-modify-writes, the reference's colour order of the sums.  ATOMIC: every batch in one
// launch, FP64 atomic adds (the order of the additions into one y entry varies from run to run).
// the stored tangents are read once per product and never reused within it: non-temporal loads keep them from
// evicting the gathered dx and the scattered y from the caches
template <class T>
__device__ __forceinline__ T ld_stream(const T* p) {
  return __builtin_nontemporal_load(p);
}

template <int Q, int P, class TS, int LAY, bool ATOMIC>
__global__ __launch_bounds__(kThreads) void k_apply(OpDev op, const void* __restrict__ tangents,
                                                    const double* __restrict__ dx, double* __restrict__ y,
                                                    int batch0, double alpha, const int* __restrict__ colfc,
                                                    int ncol, int gx) {
  using Dm = Dims<Q, P>;
  // FP32 tangents (the smoothers' storages): the sweeps and the point product run in FP32 too, which halves the
  // LDS and register footprint of the work-group; y accumulates in FP64
  using CT = typename std::conditional<std::is_same<TS, float>::value, float, double>::type;
  constexpr bool SYM = LAY == kLaySym, FAC = LAY == kLayFac || LAY == kLayAfac, AFAC = LAY == kLayAfac;
  constexpr int C = Dm::C, SA = Dm::SA, SB = Dm::SB, NSLOT = Dm::NSLOT, NE = lay_entries(LAY);
  __shared__ CT lds[Dm::LDS];
  CT* A = lds;
  CT* B = lds + C * SA;
  const int tid = threadIdx.x;
  const int batch = ATOMIC ? interleaved_batch(colfc, ncol, gx) : batch0 + blockIdx.x;
  if (ATOMIC && batch < 0) return;  // uniform over the work-group
  const int32_t* ids = op.ids + (size_t)batch * C * 3 * Dm::NN;
  const TS* Dp = static_cast<const TS*>(tangents) + (size_t)batch * NE * NSLOT + tid;
  // FP32 packed storage: the point's 45 entries are loaded before the gather and the interpolation sweeps (45
  // registers), so that their latency overlaps those phases instead of following their barriers
  // (the 17 factors of the factorised layout likewise, in either precision)
  constexpr bool PRE = (SYM && std::is_same<TS, float>::value) || FAC;
  constexpr int NPRE = PRE ? NE : 1;
  CT dpre[NPRE];
  if (PRE && tid < NSLOT) {
#pragma unroll
    for (int e = 0; e < NPRE; ++e) dpre[e] = static_cast<CT>(ld_stream(Dp + (size_t)e * NSLOT));
  }
  for (int it = tid; it < C * 3 * Dm::NN; it += kThreads) {
    const int id = ids[it];
    const int cell = it / (3 * Dm::NN);
    A[cell * SA + (it - cell * 3 * Dm::NN)] = id > 0 ? static_cast<CT>(dx[id - 1]) : CT(0);
  }
  // FP32 factorised storages: the y entries this thread adds to in the last contraction sweep (the x-line of item
  // tid) are loaded now.  No other work-group of this colour touches them and the previous colour's launch has
  // completed, so their values are final, and their latency overlaps the sweeps instead of following the last
  // barrier (afac32 0.30 -> 0.29 ms, fac32 0.311 -> 0.30 ms at 64^3 Q2; no gain in FP64).
  constexpr bool EY = !ATOMIC && FAC && std::is_same<TS, float>::value && C * 3 * Q * Q <= kThreads;
  int eid[Q];
  double ey[Q];
  if (EY && tid < C * 3 * Q * Q) {
    const int cell = tid / (3 * Q * Q), r = tid - cell * 3 * Q * Q, c = r / (Q * Q), jk = r - c * Q * Q;
    const int32_t* idl = ids + cell * 3 * Dm::NN + c * Dm::NN + Q * jk;
#pragma unroll
    for (int i = 0; i < Q; ++i) eid[i] = idl[i];
#pragma unroll
    for (int i = 0; i < Q; ++i) ey[i] = eid[i] > 0 ? y[eid[i] - 1] : 0.0;
  }
  __syncthreads();
  interp_sweeps<Q, P, CT>(A, B, op.tab, tid);
  if (tid < NSLOT) {
    const int cell = tid / Dm::NP, q = tid - cell * Dm::NP;
    CT* s = B + cell * SB + q * 9;
    CT su[9], sv[9];
#pragma unroll
    for (int j = 0; j < 9; ++j) {
      su[j] = s[j];
      sv[j] = 0.0;
    }
    if constexpr (FAC) {
      // sv[K+3c] = sum_M G_KM su[M+3c] + b sum_k H_Kk X_ck + l H_Kc tr(X),  X_ck = sum_M H_Mc su[M+3k]
      CT G[6], bm, bl;
      const CT* H;
      if constexpr (AFAC) {  // G = g iJt'iJt, l = lam w_q det J from the cell's geometry
        const double* cg = op.cellgeo + ((size_t)batch * C + cell) * 8;
#pragma unroll
        for (int e = 0; e < 6; ++e) G[e] = dpre[10] * static_cast<CT>(cg[e]);
        H = dpre;
        bm = dpre[9];
        bl = static_cast<CT>(op.lam * point_weight<P>(op.tab, q) * cg[6]);
      } else {
#pragma unroll
        for (int e = 0; e < 6; ++e) G[e] = dpre[e];
        H = dpre + 6;
        bm = dpre[15];
        bl = dpre[16];
      }
      CT X[9];
#pragma unroll
      for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          CT v = 0;
#pragma unroll
          for (int M = 0; M < 3; ++M) v += H[M + 3 * c] * su[M + 3 * k];
          X[c + 3 * k] = v;
        }
      const CT tr = X[0] + X[4] + X[8];
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int K = 0; K < 3; ++K) {
          CT g = 0, h = 0;
#pragma unroll
          for (int M = 0; M < 3; ++M) g += G[K <= M ? sym_index(K, M) : sym_index(M, K)] * su[M + 3 * c];
#pragma unroll
          for (int k = 0; k < 3; ++k) h += H[K + 3 * k] * X[c + 3 * k];
          sv[K + 3 * c] = g + bm * h + bl * H[K + 3 * c] * tr;
        }
    } else if constexpr (PRE) {
#pragma unroll
      for (int j = 0; j < 9; ++j)
#pragma unroll
        for (int i = 0; i <= j; ++i) {
          const CT d = dpre[sym_index(i, j)];
          sv[i] += d * su[j];
          if (i != j) sv[j] += d * su[i];
        }
    } else if constexpr (!SYM) {
#pragma unroll
      for (int j = 0; j < 9; ++j)
#pragma unroll
        for (int i = 0; i < 9; ++i) sv[i] += static_cast<CT>(ld_stream(Dp + (size_t)(i + 9 * j) * NSLOT)) * su[j];
    } else {
#pragma unroll
      for (int j = 0; j < 9; ++j)
#pragma unroll
        for (int i = 0; i <= j; ++i) {
          const CT d = static_cast<CT>(ld_stream(Dp + (size_t)sym_index(i, j) * NSLOT));
          sv[i] += d * su[j];
          if (i != j) sv[j] += d * su[i];
        }
    }
#pragma unroll
    for (int i = 0; i < 9; ++i) s[i] = sv[i];
  }
  contract_sweeps<Q, P, CT>(A, B, op.tab, tid, [&](int cell, int c, int t0, const CT* r) {
    const int32_t* idl = ids + cell * 3 * Dm::NN + c * Dm::NN + t0;
    if (ATOMIC) {
#pragma unroll
      for (int i = 0; i < Q; ++i) {
        const int id = idl[i];
        if (id > 0) unsafeAtomicAdd(y + id - 1, alpha * static_cast<double>(r[i]));  // hardware FP64 add in L2
      }
    } else if (EY) {
#pragma unroll
      for (int i = 0; i < Q; ++i)
        if (eid[i] > 0) y[eid[i] - 1] = ey[i] + alpha * static_cast<double>(r[i]);
    } else {
      scatter_line<Q>(idl, r, alpha, y);  // race-free: the cells of one colour share no free dof
    }
  });
}

// ---------------------------------------------------------------------------------------------------------
// BlockJacobi from the stored tangents (_blocks_from_tangents!, src/preconditioners.jl:234-263), the cells of
// one colour: B_n[ci,cj] += sum_q sum_K sum_M dN_n/dxi_K D_q[(K,ci),(M,cj)] dN_n/dxi_M.  One thread per
// (cell, node); the factorised storages form B_n from their factors directly, the others stage the tangents of the
// current point of the batch's cells in LDS.
template <int Q, int P, class TS, int LAY>
__global__ __launch_bounds__(kThreads) void k_blocks(OpDev op, const void* __restrict__ tangents,
                                                     const int32_t* __restrict__ block_of, double* __restrict__ Bk,
                                                     int batch0) {
  using Dm = Dims<Q, P>;
  constexpr bool SYM = LAY == kLaySym, FAC = LAY == kLayFac || LAY == kLayAfac, AFAC = LAY == kLayAfac;
  constexpr int C = Dm::C, NN = Dm::NN, NP = Dm::NP, NSLOT = Dm::NSLOT, NE = lay_entries(LAY);
  const int tid = threadIdx.x;
  const int batch = batch0 + blockIdx.x;
  const bool active = tid < C * NN;
  const int cell = tid / NN, t = tid - cell * NN;
  const int i = t % Q, j = (t / Q) % Q, k = t / (Q * Q);
  const TS* Dg = static_cast<const TS*>(tangents) + (size_t)batch * NE * NSLOT;
  double acc[9];
#pragma unroll
  for (int e = 0; e < 9; ++e) acc[e] = 0.0;
  if constexpr (FAC) {
    // The factorised tangent D[(K,c),(M,k)] = G_KM d_ck + b H_Kk H_Mc + l H_Kc H_Mk contracted with the node's
    // gradient g on both sides: B_n = sum_q (g'G g) I + (b + l) u u', u_c = sum_K g_K H_Kc.  The batch's factors
    // are staged in LDS once (entry-major, as stored: coalesced), then every (cell, node) thread walks its points.
    __shared__ double Fl[NE * NSLOT];
    for (int it = tid; it < NE * NSLOT; it += kThreads) Fl[it] = static_cast<double>(ld_stream(Dg + it));
    __syncthreads();
    if (active) {
      double S[6] = {0, 0, 0, 0, 0, 0}, dj = 0;
      if constexpr (AFAC) {  // the cell's metric iJt'iJt and |det J|
        const double* cg = op.cellgeo + ((size_t)batch * C + cell) * 8;
#pragma unroll
        for (int e = 0; e < 6; ++e) S[e] = cg[e];
        dj = cg[6];
      }
      for (int q = 0; q < NP; ++q) {
        const int a = q % P, b = (q / P) % P, cc = q / (P * P);
        const double na = op.tab.N[a * Q + i], nb = op.tab.N[b * Q + j], nc = op.tab.N[cc * Q + k];
        const double gr[3] = {op.tab.dN[a * Q + i] * nb * nc, na * op.tab.dN[b * Q + j] * nc, na * nb * op.tab.dN[cc * Q + k]};
        const double* f = Fl + cell * NP + q;  // factor e of this point: f[e * NSLOT]
        double Gm[6], bb;
        constexpr int HO = AFAC ? 0 : 6;  // first factor of H
        if constexpr (AFAC) {
          const double g = f[10 * NSLOT];
#pragma unroll
          for (int e = 0; e < 6; ++e) Gm[e] = g * S[e];
          bb = f[9 * NSLOT] + op.lam * point_weight<P>(op.tab, q) * dj;
        } else {
#pragma unroll
          for (int e = 0; e < 6; ++e) Gm[e] = f[e * NSLOT];
          bb = f[15 * NSLOT] + f[16 * NSLOT];
        }
        double u[3], gg = 0;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          double v = 0;
#pragma unroll
          for (int K = 0; K < 3; ++K) v += gr[K] * f[(HO + K + 3 * c) * NSLOT];
          u[c] = v;
        }
#pragma unroll
        for (int K = 0; K < 3; ++K)
#pragma unroll
          for (int M = 0; M < 3; ++M) gg += gr[K] * Gm[K <= M ? sym_index(K, M) : sym_index(M, K)] * gr[M];
#pragma unroll
        for (int cj = 0; cj < 3; ++cj)
#pragma unroll
          for (int ci = 0; ci < 3; ++ci) acc[ci + 3 * cj] += (ci == cj ? gg : 0.0) + bb * u[ci] * u[cj];
      }
    }
  } else {  // full or symmetric storage: the batch's tangents at one point staged in LDS
    __shared__ double Dl[C * 81];
    for (int q = 0; q < NP; ++q) {
      __syncthreads();
      for (int it = tid; it < C * 81; it += kThreads) {
        const int c2 = it / 81, e = it - c2 * 81, ii = e % 9, jj = e / 9;
        const int se = SYM ? (ii <= jj ? sym_index(ii, jj) : sym_index(jj, ii)) : e;
        Dl[it] = static_cast<double>(ld_stream(Dg + (size_t)se * NSLOT + c2 * NP + q));
      }
      __syncthreads();
      if (active) {
        const int a = q % P, b = (q / P) % P, cc = q / (P * P);
        const double na = op.tab.N[a * Q + i], nb = op.tab.N[b * Q + j], nc = op.tab.N[cc * Q + k];
        const double G[3] = {op.tab.dN[a * Q + i] * nb * nc, na * op.tab.dN[b * Q + j] * nc, na * nb * op.tab.dN[cc * Q + k]};
        const double* D = Dl + cell * 81;
#pragma unroll
        for (int cj = 0; cj < 3; ++cj)
#pragma unroll
          for (int ci = 0; ci < 3; ++ci) {
            double s = 0;
#pragma unroll
            for (int M = 0; M < 3; ++M) {
              double u = 0;
#pragma unroll
              for (int K = 0; K < 3; ++K) u += G[K] * D[(K + 3 * ci) + 9 * (M + 3 * cj)];
              s += u * G[M];
            }
            acc[ci + 3 * cj] += s;
          }
      }
    }
  }
  if (active) {
    const int blk = block_of[((size_t)batch * C + cell) * NN + t];
    if (blk >= 0) {
#pragma unroll
      for (int e = 0; e < 9; ++e) Bk[(size_t)blk * 9 + e] += acc[e];  // race-free within a colour
    }
  }
}

// ---------------------------------------------------------------------------------------------------------
// jacobian!: element matrices from the prepared FP64 tangents, per cell of a colour, summed into the nonzeros
// through nzpos (element_jacobian_mm!/scatter_jacobian!, src/assembly.jl:1316-1408).
// Ke[(n,c),(m,k)] = sum_q sum_K sum_M dN_n/dxi_K(q) D_q[(K,c),(M,k)] dN_m/dxi_M(q), in two steps per point:
//   H[K][c][(m,k)] = sum_M D_q[(K,c),(M,k)] dN_m/dxi_M(q);   Ke[(n,c),(m,k)] += sum_K dN_n/dxi_K(q) H[K][c][(m,k)]
// The cell's tangents at all its points are loaded into LDS first (one round of loads instead of one per point).
// SPLIT > 1: the cell's columns (m,k) are shared out over SPLIT work-groups (each forms H for its columns only),
// for colours with too few cells to occupy the GPU with one work-group per cell.  The sums of every entry are
// formed in the same order whatever the split.
template <int Q, int P, int NT, int SPLIT>
__global__ __launch_bounds__(NT) void k_assemble(OpDev op, const double* __restrict__ D, double* __restrict__ nzval,
                                                 int batch0) {
  using Dm = Dims<Q, P>;
  constexpr int C = Dm::C, NN = Dm::NN, ND = 3 * NN, NSLOT = Dm::NSLOT, NP = Dm::NP;
  constexpr int NCOL = (ND + SPLIT - 1) / SPLIT;  // columns of one work-group, at most
  // Each thread accumulates 3 x 3 blocks of entries: 3 consecutive rows (n, c) of one component c by 3
  // consecutive columns, so that the 9 gradient values and 9 H values it reads per point serve 27 products.
  constexpr int RB = (NN + 2) / 3;        // row blocks per component
  constexpr int NRB = 3 * RB;             // row blocks
  constexpr int NCB = (NCOL + 2) / 3;     // column blocks, at most
  constexpr int NBLK = (NRB * NCB + NT - 1) / NT;  // blocks per thread
  __shared__ double Dl[NP * 81];  // Dl[q * 81 + e]: entry e of the cell's tangent at point q
  __shared__ double H[9 * NCOL];
  __shared__ double G[3 * NN];  // dN_t/dxi_K at the current point
  const int tid = threadIdx.x;
  const int part = SPLIT > 1 ? (int)(blockIdx.x % SPLIT) : 0;
  const int cb = (int)(blockIdx.x / SPLIT);
  const int batch = batch0 + cb / C, cell = cb % C;
  const int col0 = part * NCOL, ncol = ND - col0 < NCOL ? ND - col0 : NCOL;
  const int32_t* ids = op.ids + ((size_t)batch * C + cell) * ND;
  bool pad = true;
  for (int j = 0; j < ND; ++j) pad = pad && ids[j] == 0;
  if (pad) return;  // padding cell of the last batch of a colour (uniform over the work-group)
  {
    const double* Dc = D + (size_t)batch * 81 * NSLOT + cell * NP;
    for (int it = tid; it < 81 * NP; it += NT) {
      const int e = it / NP, q = it - e * NP;  // consecutive threads read consecutive points of one entry
      Dl[q * 81 + e] = Dc[(size_t)e * NSLOT + q];
    }
  }
  const int ncb = (ncol + 2) / 3;
  double acc[NBLK][3][3];
#pragma unroll
  for (int k = 0; k < NBLK; ++k)
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) acc[k][i][j] = 0.0;
  for (int q = 0; q < NP; ++q) {
    const int a = q % P, b = (q / P) % P, cc = q / (P * P);
    for (int it = tid; it < 3 * NN; it += NT) {
      const int t = it / 3, K = it - 3 * t;
      const int i = t % Q, j = (t / Q) % Q, k = t / (Q * Q);
      const double na = op.tab.N[a * Q + i], nb = op.tab.N[b * Q + j], nc = op.tab.N[cc * Q + k];
      G[it] = K == 0 ? op.tab.dN[a * Q + i] * nb * nc : K == 1 ? na * op.tab.dN[b * Q + j] * nc : na * nb * op.tab.dN[cc * Q + k];
    }
    __syncthreads();
    const double* Dq = Dl + q * 81;
    for (int it = tid; it < 9 * ncol; it += NT) {
      const int Kc = it % 9, mk = col0 + it / 9;  // H[Kc + 9 (mk - col0)], Kc = K + 3c, mk = m + NN k (tensor order)
      const int k = mk / NN, m = mk - k * NN;
      double sum = 0;
#pragma unroll
      for (int M = 0; M < 3; ++M) sum += Dq[Kc + 9 * (M + 3 * k)] * G[3 * m + M];
      H[it] = sum;
    }
    __syncthreads();
#pragma unroll
    for (int kb = 0; kb < NBLK; ++kb) {
      const int blk = tid + kb * NT;
      if (blk < NRB * ncb) {
        const int rb = blk % NRB, cbk = blk / NRB;
        const int c = rb / RB, n0 = 3 * (rb - c * RB), l0 = 3 * cbk;
        double g[3][3], h[3][3];  // g[i][K] = G of row n0 + i, h[K][j] = H of column l0 + j
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int K = 0; K < 3; ++K) g[i][K] = n0 + i < NN ? G[3 * (n0 + i) + K] : 0.0;
#pragma unroll
        for (int K = 0; K < 3; ++K)
#pragma unroll
          for (int j = 0; j < 3; ++j) h[K][j] = l0 + j < ncol ? H[(K + 3 * c) + 9 * (l0 + j)] : 0.0;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int j = 0; j < 3; ++j) {
            double sum = 0;
#pragma unroll
            for (int K = 0; K < 3; ++K) sum += g[i][K] * h[K][j];
            acc[kb][i][j] += sum;
          }
      }
    }
    __syncthreads();
  }
  const int32_t* pos = op.nzpos + ((size_t)batch * C + cell) * ND * ND;
#pragma unroll
  for (int kb = 0; kb < NBLK; ++kb) {
    const int blk = tid + kb * NT;
    if (blk < NRB * ncb) {
      const int rb = blk % NRB, cbk = blk / NRB;
      const int c = rb / RB, n0 = 3 * (rb - c * RB), l0 = 3 * cbk;
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          if (n0 + i >= NN || l0 + j >= ncol) continue;
          const int ii = op.loc_of[c * NN + n0 + i], jj = op.loc_of[col0 + l0 + j];
          const int p = pos[ii + ND * jj];
          if (p > 0) nzval[p - 1] += acc[kb][i][j];  // race-free: the cells of one colour share no free test row
        }
    }
  }
}

__global__ void k_fill(double* __restrict__ y, int64_t n, double v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) y[i] = v;
}

// dS of Gauss point (a, b) of the reference-cube face xi_d = side (other directions e1 < e2) under the trilinear map
__device__ __forceinline__ double face_dS(const Tab& tab, const double* __restrict__ X, int d, int side, int e1, int e2,
                                          int a, int b) {
  double xi[3];
  xi[d] = side;
  xi[e1] = tab.xq[a];
  xi[e2] = tab.xq[b];
  double g1[3] = {0.0, 0.0, 0.0}, g2[3] = {0.0, 0.0, 0.0};
  for (int v = 0; v < 8; ++v) {
    double mv[3], dv[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const int bit = (v >> k) & 1;
      mv[k] = bit ? xi[k] : 1.0 - xi[k];
      dv[k] = bit ? 1.0 : -1.0;
    }
    const double d1 = dv[e1] * mv[e2] * mv[d], d2 = dv[e2] * mv[e1] * mv[d];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      g1[k] += d1 * X[v * 3 + k];
      g2[k] += d2 * X[v * 3 + k];
    }
  }
  const double n0 = g1[1] * g2[2] - g1[2] * g2[1], n1 = g1[2] * g2[0] - g1[0] * g2[2], n2 = g1[0] * g2[1] - g1[1] * g2[0];
  return tab.wq[a] * tab.wq[b] * sqrt(n0 * n0 + n1 * n1 + n2 * n2);
}

// tensor node of face node m = i1 + Q i2 (i1 along e1, i2 along e2) of the face xi_d = side
__device__ __forceinline__ int face_tnode(int Q, int d, int side, int e1, int e2, int m) {
  int idx[3];
  idx[d] = side ? Q - 1 : 0;
  idx[e1] = m % Q;
  idx[e2] = m / Q;
  return idx[0] + Q * (idx[1] + Q * idx[2]);
}

__device__ __forceinline__ void face_dirs(int lf, int& d, int& side, int& e1, int& e2) {
  d = lf >> 1;
  side = lf & 1;
  e1 = d == 0 ? 1 : 0;
  e2 = d == 2 ? 1 : 2;
}

// Robin (spring) term k int u.v dGamma on faces of the cells.  Per face the scalar face mass matrix
// M[m1][m2] = k sum_q dS N_m1 N_m2 over the face's Q^2 nodes (the same for every component), computed once; the
// residual, Jv, the node blocks, the assembled values and the energy are built from it.  Launch groups hold faces of
// different cells of one colour (no shared dof).
__global__ __launch_bounds__(256) void k_face_mass(OpDev op, const int64_t* __restrict__ face_slot,
                                                   const int32_t* __restrict__ face_local, const double* __restrict__ kf,
                                                   double* __restrict__ M) {
  __shared__ double dS[16];
  const int64_t f = blockIdx.x;
  const int Q = op.Q, P = op.P, Q2 = Q * Q;
  int d, side, e1, e2;
  face_dirs(face_local[f], d, side, e1, e2);
  const int tid = threadIdx.x;
  if (tid < P * P) dS[tid] = face_dS(op.tab, op.coords + face_slot[f] * 24, d, side, e1, e2, tid % P, tid / P);
  __syncthreads();
  if (tid >= Q2 * Q2) return;
  const int m1 = tid / Q2, m2 = tid - m1 * Q2;
  const int i1 = m1 % Q, j1 = m1 / Q, i2 = m2 % Q, j2 = m2 / Q;
  double s = 0.0;
  for (int b = 0; b < P; ++b)
    for (int a = 0; a < P; ++a)
      s += dS[a + P * b] * op.tab.N[a * Q + i1] * op.tab.N[b * Q + j1] * op.tab.N[a * Q + i2] * op.tab.N[b * Q + j2];
  M[f * Q2 * Q2 + tid] = kf[f] * s;
}

__device__ __forceinline__ double dof_value(int32_t id, const double* __restrict__ uf, const double* __restrict__ ud) {
  return id > 0 ? uf[id - 1] : (id < 0 && ud ? ud[-id - 1] : 0.0);
}

// y[row] += alpha sum_m2 M[m1][m2] u[(c, m2)] on the free rows (c, m1) of the faces of one group; u from the free
// values uf and the Dirichlet values ud (null: zero)
__global__ __launch_bounds__(64) void k_face_apply(OpDev op, const int64_t* __restrict__ face_slot,
                                                   const int32_t* __restrict__ face_local, const double* __restrict__ M,
                                                   const int32_t* __restrict__ faces, const double* __restrict__ uf,
                                                   const double* __restrict__ ud, double* __restrict__ y, double alpha) {
  const int64_t f = faces[blockIdx.x];
  const int Q = op.Q, Q2 = Q * Q, NN = Q * Q2;
  int d, side, e1, e2;
  face_dirs(face_local[f], d, side, e1, e2);
  const int tid = threadIdx.x;
  if (tid >= 3 * Q2) return;
  const int c = tid / Q2, m1 = tid - c * Q2;
  const int32_t* ids = op.ids + face_slot[f] * 3 * NN + c * NN;
  const int32_t row = ids[face_tnode(Q, d, side, e1, e2, m1)];
  if (row <= 0) return;
  const double* Mr = M + f * Q2 * Q2 + m1 * Q2;
  double s = 0.0;
  for (int m2 = 0; m2 < Q2; ++m2) s += Mr[m2] * dof_value(ids[face_tnode(Q, d, side, e1, e2, m2)], uf, ud);
  y[row - 1] += alpha * s;
}

// the diagonal of the node blocks: B[blk][c + 3c] += M[m1][m1]
__global__ __launch_bounds__(64) void k_face_blocks(OpDev op, const int64_t* __restrict__ face_slot,
                                                    const int32_t* __restrict__ face_local, const double* __restrict__ M,
                                                    const int32_t* __restrict__ faces, const int32_t* __restrict__ block_of,
                                                    double* __restrict__ B) {
  const int64_t f = faces[blockIdx.x];
  const int Q = op.Q, Q2 = Q * Q, NN = Q * Q2;
  int d, side, e1, e2;
  face_dirs(face_local[f], d, side, e1, e2);
  const int m1 = threadIdx.x;
  if (m1 >= Q2) return;
  const int blk = block_of[face_slot[f] * NN + face_tnode(Q, d, side, e1, e2, m1)];
  if (blk < 0) return;
  const double v = M[f * Q2 * Q2 + m1 * Q2 + m1];
#pragma unroll
  for (int c = 0; c < 3; ++c) B[(size_t)blk * 9 + 4 * c] += v;
}

// the assembled values: nzval[pos(c m1, c m2)] += M[m1][m2]
__global__ __launch_bounds__(256) void k_face_assemble(OpDev op, const int64_t* __restrict__ face_slot,
                                                       const int32_t* __restrict__ face_local, const double* __restrict__ M,
                                                       const int32_t* __restrict__ faces, double* __restrict__ nzval) {
  const int64_t f = faces[blockIdx.x];
  const int Q = op.Q, Q2 = Q * Q, NN = Q * Q2, ND = 3 * NN;
  int d, side, e1, e2;
  face_dirs(face_local[f], d, side, e1, e2);
  const int32_t* pos = op.nzpos + face_slot[f] * ND * ND;
  for (int it = threadIdx.x; it < 3 * Q2 * Q2; it += 256) {
    const int c = it / (Q2 * Q2), r = it - c * Q2 * Q2, m1 = r / Q2, m2 = r - m1 * Q2;
    const int ii = op.loc_of[c * NN + face_tnode(Q, d, side, e1, e2, m1)];
    const int jj = op.loc_of[c * NN + face_tnode(Q, d, side, e1, e2, m2)];
    const int p = pos[ii + ND * jj];
    if (p > 0) nzval[p - 1] += M[f * Q2 * Q2 + r];
  }
}

// the energy of the springs: part[f] = 1/2 sum_c sum_m1,m2 u_c,m1 M[m1][m2] u_c,m2 (free and Dirichlet values)
__global__ __launch_bounds__(kThreads) void k_face_energy(OpDev op, const int64_t* __restrict__ face_slot,
                                                          const int32_t* __restrict__ face_local,
                                                          const double* __restrict__ M, const double* __restrict__ x,
                                                          double* __restrict__ part) {
  __shared__ double red[kThreads];
  const int64_t f = blockIdx.x;
  const int Q = op.Q, Q2 = Q * Q, NN = Q * Q2;
  int d, side, e1, e2;
  face_dirs(face_local[f], d, side, e1, e2);
  const int tid = threadIdx.x;
  double e = 0.0;
  if (tid < 3 * Q2) {
    const int c = tid / Q2, m1 = tid - c * Q2;
    const int32_t* ids = op.ids + face_slot[f] * 3 * NN + c * NN;
    const double* Mr = M + f * Q2 * Q2 + m1 * Q2;
    double s = 0.0;
    for (int m2 = 0; m2 < Q2; ++m2) s += Mr[m2] * dof_value(ids[face_tnode(Q, d, side, e1, e2, m2)], x, op.dvals);
    e = 0.5 * dof_value(ids[face_tnode(Q, d, side, e1, e2, m1)], x, op.dvals) * s;
  }
  e = block_reduce<0>(e, red);
  if (tid == 0) part[f] = e;
}

// Boundary traction (a Neumann load on faces of the cells): one work-group per face, one thread per (component,
// node of the face); the load of that dof, the sum over the P x P Gauss points of the face of dS t_c N_node, is added
// into f (free ids at f[id - 1], Dirichlet ids at f[nfree - id - 1]).  The faces of one launch belong to different
// cells of one colour, so no two of them share a dof.  Local face 2d + s is the reference-cube face xi_d = s; e1 < e2 are the
// other directions, dS = w_a w_b |dx/dxi_e1 x dx/dxi_e2| with the trilinear map.
__global__ __launch_bounds__(64) void k_face_load(OpDev op, const int64_t* __restrict__ face_slot,
                                                  const int32_t* __restrict__ face_local,
                                                  const double* __restrict__ traction, const int32_t* __restrict__ faces,
                                                  int64_t nfree, double* __restrict__ f) {
  const int64_t fi = faces[blockIdx.x];
  const int64_t slot = face_slot[fi];
  int d, side, e1, e2;
  face_dirs(face_local[fi], d, side, e1, e2);
  const int Q = op.Q, P = op.P, NN = Q * Q * Q;
  const int tid = threadIdx.x;
  if (tid >= 3 * Q * Q) return;
  const int c = tid / (Q * Q), m = tid - c * Q * Q, i1 = m % Q, i2 = m / Q;
  const double* X = op.coords + slot * 24;
  double acc = 0.0;
  for (int b = 0; b < P; ++b)
    for (int a = 0; a < P; ++a) acc += face_dS(op.tab, X, d, side, e1, e2, a, b) * op.tab.N[a * Q + i1] * op.tab.N[b * Q + i2];
  const int32_t id = op.ids[slot * 3 * NN + c * NN + face_tnode(Q, d, side, e1, e2, m)];
  const double v = acc * traction[fi * 3 + c];
  if (id > 0) f[id - 1] += v;
  else if (id < 0) f[nfree - id - 1] += v;
}

__global__ void k_sub_scaled(double* __restrict__ y, const double* __restrict__ x, double a, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] -= a * x[i];
}

// scale * sum a_i b_i per work-group into part[blockIdx.x], a fixed order for a fixed grid
__global__ __launch_bounds__(kThreads) void k_dot_partials(const double* __restrict__ a, const double* __restrict__ b,
                                                           int64_t n, double scale, double* __restrict__ part) {
  __shared__ double red[kThreads];
  double s = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads)
    s += a[i] * b[i];
  s = block_reduce<0>(s, red);
  if (threadIdx.x == 0) part[blockIdx.x] = scale * s;
}

// Expand the stored tangents of every batch to D[(batch*C + cell)*np + q][81] (FP64), for inspection.
template <int P, class TS, int LAY>
__global__ void k_expand(OpDev op, const void* __restrict__ tangents, double* __restrict__ out, int64_t nbatches) {
  constexpr bool SYM = LAY == kLaySym;
  constexpr int NP = P * P * P, C = cells_per_batch(P), NSLOT = C * NP, NE = lay_entries(LAY);
  const int64_t n = nbatches * NSLOT;
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < n; g += (int64_t)gridDim.x * blockDim.x) {
    const int64_t batch = g / NSLOT;
    const int slot = (int)(g - batch * NSLOT);
    const TS* Dp = static_cast<const TS*>(tangents) + batch * NE * NSLOT + slot;
    for (int j = 0; j < 9; ++j)
      for (int i = 0; i < 9; ++i) {
        if (LAY == kLayFac) {
          out[g * 81 + i + 9 * j] = fac_entry(Dp, NSLOT, i, j);
          continue;
        }
        if (LAY == kLayAfac) {
          out[g * 81 + i + 9 * j] = afac_entry(Dp, NSLOT, op.cellgeo + (batch * C + slot / NP) * 8,
                                               point_weight<P>(op.tab, slot % NP), op.lam, i, j);
          continue;
        }
        const int e = SYM ? (i <= j ? sym_index(i, j) : sym_index(j, i)) : i + 9 * j;
        out[g * 81 + i + 9 * j] = static_cast<double>(Dp[(size_t)e * NSLOT]);
      }
  }
}

// ---------------------------------------------------------------------------------------------------------
// dispatch helpers
template <int Q, int P, class Law>
hipError_t prepare_qp(const OpDev& op, int storage, void* T, const double* x, BatchRange r, hipStream_t s) {
  const dim3 g(r.count), b(kThreads);
  switch (storage) {
    case 0: hipLaunchKernelGGL((k_prepare<Q, P, Law, double, kLayFull>), g, b, 0, s, op, T, x, r.first); break;
    case 1: hipLaunchKernelGGL((k_prepare<Q, P, Law, double, kLaySym>), g, b, 0, s, op, T, x, r.first); break;
    case 2: hipLaunchKernelGGL((k_prepare<Q, P, Law, float, kLayFull>), g, b, 0, s, op, T, x, r.first); break;
    case 3: hipLaunchKernelGGL((k_prepare<Q, P, Law, float, kLaySym>), g, b, 0, s, op, T, x, r.first); break;
    default:
      if constexpr (std::is_same<Law, LawSVK>::value) {
        return hipErrorInvalidValue;  // no factorised form (refused on the host before)
      } else {
        if (storage == 4) hipLaunchKernelGGL((k_prepare<Q, P, Law, double, kLayFac>), g, b, 0, s, op, T, x, r.first);
        else if (storage == 5) hipLaunchKernelGGL((k_prepare<Q, P, Law, float, kLayFac>), g, b, 0, s, op, T, x, r.first);
        else if (storage == 6) hipLaunchKernelGGL((k_prepare<Q, P, Law, double, kLayAfac>), g, b, 0, s, op, T, x, r.first);
        else hipLaunchKernelGGL((k_prepare<Q, P, Law, float, kLayAfac>), g, b, 0, s, op, T, x, r.first);
      }
      break;
  }
  return hipGetLastError();
}

template <int Q, int P>
hipError_t prepare_law(const OpDev& op, int storage, void* T, const double* x, BatchRange r, hipStream_t s) {
  switch (op.law) {
    case 0: return prepare_qp<Q, P, LawSVK>(op, storage, T, x, r, s);
    case 1: return prepare_qp<Q, P, LawNeoHooke>(op, storage, T, x, r, s);
    default: return prepare_qp<Q, P, LawLinear>(op, storage, T, x, r, s);
  }
}

template <int Q, int P, bool ATOMIC>
hipError_t apply_qpa(const OpDev& op, int storage, const void* T, const double* dx, double* y, BatchRange r,
                     hipStream_t s, double alpha, const ApplyMap* m) {
  const dim3 g(ATOMIC ? m->grid : r.count), b(kThreads);
  const int* fc = ATOMIC ? m->colfc : nullptr;
  const int nc = ATOMIC ? m->ncol : 0, gx = ATOMIC ? m->gx : 0;
  switch (storage) {
    case 0: hipLaunchKernelGGL((k_apply<Q, P, double, kLayFull, ATOMIC>), g, b, 0, s, op, T, dx, y, r.first, alpha, fc, nc, gx); break;
    case 1: hipLaunchKernelGGL((k_apply<Q, P, double, kLaySym, ATOMIC>), g, b, 0, s, op, T, dx, y, r.first, alpha, fc, nc, gx); break;
    case 2: hipLaunchKernelGGL((k_apply<Q, P, float, kLayFull, ATOMIC>), g, b, 0, s, op, T, dx, y, r.first, alpha, fc, nc, gx); break;
    case 3: hipLaunchKernelGGL((k_apply<Q, P, float, kLaySym, ATOMIC>), g, b, 0, s, op, T, dx, y, r.first, alpha, fc, nc, gx); break;
    case 4: hipLaunchKernelGGL((k_apply<Q, P, double, kLayFac, ATOMIC>), g, b, 0, s, op, T, dx, y, r.first, alpha, fc, nc, gx); break;
    case 5: hipLaunchKernelGGL((k_apply<Q, P, float, kLayFac, ATOMIC>), g, b, 0, s, op, T, dx, y, r.first, alpha, fc, nc, gx); break;
    case 6: hipLaunchKernelGGL((k_apply<Q, P, double, kLayAfac, ATOMIC>), g, b, 0, s, op, T, dx, y, r.first, alpha, fc, nc, gx); break;
    default: hipLaunchKernelGGL((k_apply<Q, P, float, kLayAfac, ATOMIC>), g, b, 0, s, op, T, dx, y, r.first, alpha, fc, nc, gx); break;
  }
  return hipGetLastError();
}

template <int Q, int P>
hipError_t apply_qp(const OpDev& op, int storage, const void* T, const double* dx, double* y, BatchRange r,
                    hipStream_t s, double alpha, const ApplyMap* m) {
  return m ? apply_qpa<Q, P, true>(op, storage, T, dx, y, r, s, alpha, m)
           : apply_qpa<Q, P, false>(op, storage, T, dx, y, r, s, alpha, m);
}

template <int Q, int P>
hipError_t blocks_qp(const OpDev& op, int storage, const void* T, const int32_t* block_of, double* B, BatchRange r,
                     hipStream_t s) {
  const dim3 g(r.count), b(kThreads);
  switch (storage) {
    case 0: hipLaunchKernelGGL((k_blocks<Q, P, double, kLayFull>), g, b, 0, s, op, T, block_of, B, r.first); break;
    case 1: hipLaunchKernelGGL((k_blocks<Q, P, double, kLaySym>), g, b, 0, s, op, T, block_of, B, r.first); break;
    case 2: hipLaunchKernelGGL((k_blocks<Q, P, float, kLayFull>), g, b, 0, s, op, T, block_of, B, r.first); break;
    case 3: hipLaunchKernelGGL((k_blocks<Q, P, float, kLaySym>), g, b, 0, s, op, T, block_of, B, r.first); break;
    case 4: hipLaunchKernelGGL((k_blocks<Q, P, double, kLayFac>), g, b, 0, s, op, T, block_of, B, r.first); break;
    case 5: hipLaunchKernelGGL((k_blocks<Q, P, float, kLayFac>), g, b, 0, s, op, T, block_of, B, r.first); break;
    case 6: hipLaunchKernelGGL((k_blocks<Q, P, double, kLayAfac>), g, b, 0, s, op, T, block_of, B, r.first); break;
    default: hipLaunchKernelGGL((k_blocks<Q, P, float, kLayAfac>), g, b, 0, s, op, T, block_of, B, r.first); break;
  }
  return hipGetLastError();
}

template <int Q, int P>
hipError_t residual_law(const OpDev& op, const double* x, double* bvec, BatchRange r, hipStream_t s) {
  const dim3 g(r.count), b(kThreads);
  switch (op.law) {
    case 0: hipLaunchKernelGGL((k_residual<Q, P, LawSVK>), g, b, 0, s, op, x, bvec, r.first); break;
    case 1: hipLaunchKernelGGL((k_residual<Q, P, LawNeoHooke>), g, b, 0, s, op, x, bvec, r.first); break;
    default: hipLaunchKernelGGL((k_residual<Q, P, LawLinear>), g, b, 0, s, op, x, bvec, r.first); break;
  }
  return hipGetLastError();
}

template <int Q, int P>
hipError_t energy_law(const OpDev& op, const double* x, double* part, int64_t nbatches, hipStream_t s) {
  const dim3 g((unsigned)nbatches), b(kThreads);
  switch (op.law) {
    case 0: hipLaunchKernelGGL((k_energy<Q, P, LawSVK>), g, b, 0, s, op, x, part); break;
    case 1: hipLaunchKernelGGL((k_energy<Q, P, LawNeoHooke>), g, b, 0, s, op, x, part); break;
    default: hipLaunchKernelGGL((k_energy<Q, P, LawLinear>), g, b, 0, s, op, x, part); break;
  }
  return hipGetLastError();
}

template <int Q, int P>
hipError_t assemble_qp(const OpDev& op, const double* D, double* nzval, BatchRange r, hipStream_t s) {
  constexpr int C = Dims<Q, P>::C;
  constexpr int NT = Q >= 4 ? 1024 : 256;
  // Q2 and Q3 cells: the columns of a cell always in 4 work-groups (with all 81 x 81 or 192 x 192 entries in one,
  // the accumulators take every register: one wave per SIMD, and spills at Q3); Q1 cells: split only when a colour
  // has too few cells to occupy the GPU with one work-group per cell
  const int64_t cells = (int64_t)r.count * C;
  if constexpr (Q >= 3) {
    hipLaunchKernelGGL((k_assemble<Q, P, NT, 4>), dim3(r.count * C * 4), dim3(NT), 0, s, op, D, nzval, r.first);
  } else {
    if (cells < 1024)
      hipLaunchKernelGGL((k_assemble<Q, P, NT, 4>), dim3(r.count * C * 4), dim3(NT), 0, s, op, D, nzval, r.first);
    else if (cells < 2048)
      hipLaunchKernelGGL((k_assemble<Q, P, NT, 2>), dim3(r.count * C * 2), dim3(NT), 0, s, op, D, nzval, r.first);
    else
      hipLaunchKernelGGL((k_assemble<Q, P, NT, 1>), dim3(r.count * C), dim3(NT), 0, s, op, D, nzval, r.first);
  }
  return hipGetLastError();
}

// The (Q, P) pairs with kernels: Q1, Q2, Q3 with P = Q (degree 2 order) and P = Q + 1; Q1 with the points of
// the finer levels (the coarse level of p-multigrid keeps the fine quadrature, src/pmultigrid.jl:248-251).
#define HFK_DISPATCH_QP(Q_, P_, CALL)                                   \
  do {                                                                  \
    const int q__ = (Q_), p__ = (P_);                                   \
    if (q__ == 2 && p__ == 2) return CALL(2, 2);                        \
    if (q__ == 2 && p__ == 3) return CALL(2, 3);                        \
    if (q__ == 2 && p__ == 4) return CALL(2, 4);                        \
    if (q__ == 3 && p__ == 3) return CALL(3, 3);                        \
    if (q__ == 3 && p__ == 4) return CALL(3, 4);                        \
    if (q__ == 4 && p__ == 4) return CALL(4, 4);                        \
    return hipErrorInvalidValue;                                        \
  } while (0)

}  // namespace

bool supported_qp(int Q, int P) {
  return (Q == 2 && (P == 2 || P == 3 || P == 4)) || (Q == 3 && (P == 3 || P == 4)) || (Q == 4 && P == 4);
}

hipError_t launch_prepare(const OpDev& op, int storage, void* T, const double* x, BatchRange r, hipStream_t s) {
  if (r.count <= 0) return hipSuccess;
#define CALL(q, p) prepare_law<q, p>(op, storage, T, x, r, s)
  HFK_DISPATCH_QP(op.Q, op.P, CALL);
#undef CALL
}

hipError_t launch_apply(const OpDev& op, int storage, const void* T, const double* dx, double* y, BatchRange r,
                        hipStream_t s, double alpha, const ApplyMap* m) {
  if (!m && r.count <= 0) return hipSuccess;
#define CALL(q, p) apply_qp<q, p>(op, storage, T, dx, y, r, s, alpha, m)
  HFK_DISPATCH_QP(op.Q, op.P, CALL);
#undef CALL
}

hipError_t launch_blocks(const OpDev& op, int storage, const void* T, const int32_t* block_of, double* B, BatchRange r,
                         hipStream_t s) {
  if (r.count <= 0) return hipSuccess;
#define CALL(q, p) blocks_qp<q, p>(op, storage, T, block_of, B, r, s)
  HFK_DISPATCH_QP(op.Q, op.P, CALL);
#undef CALL
}

hipError_t launch_residual(const OpDev& op, const double* x, double* b, BatchRange r, hipStream_t s) {
  if (r.count <= 0) return hipSuccess;
#define CALL(q, p) residual_law<q, p>(op, x, b, r, s)
  HFK_DISPATCH_QP(op.Q, op.P, CALL);
#undef CALL
}

hipError_t launch_energy(const OpDev& op, const double* x, double* part, int64_t nbatches, hipStream_t s) {
  if (nbatches <= 0) return hipSuccess;
#define CALL(q, p) energy_law<q, p>(op, x, part, nbatches, s)
  HFK_DISPATCH_QP(op.Q, op.P, CALL);
#undef CALL
}

hipError_t launch_face_load(const OpDev& op, const int64_t* face_slot, const int32_t* face_local, const double* traction,
                            const int32_t* faces, int64_t nfaces, int64_t nfree, double* f, hipStream_t s) {
  if (nfaces <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_face_load, dim3((unsigned)nfaces), dim3(64), 0, s, op, face_slot, face_local, traction, faces,
                     nfree, f);
  return hipGetLastError();
}

hipError_t launch_face_mass(const OpDev& op, const int64_t* face_slot, const int32_t* face_local, const double* kf,
                            int64_t nfaces, double* M, hipStream_t s) {
  if (nfaces <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_face_mass, dim3((unsigned)nfaces), dim3(256), 0, s, op, face_slot, face_local, kf, M);
  return hipGetLastError();
}

hipError_t launch_face_apply(const OpDev& op, const int64_t* face_slot, const int32_t* face_local, const double* M,
                             const int32_t* faces, int64_t n, const double* uf, const double* ud, double* y, double alpha,
                             hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_face_apply, dim3((unsigned)n), dim3(64), 0, s, op, face_slot, face_local, M, faces, uf, ud, y,
                     alpha);
  return hipGetLastError();
}

hipError_t launch_face_blocks(const OpDev& op, const int64_t* face_slot, const int32_t* face_local, const double* M,
                              const int32_t* faces, int64_t n, const int32_t* block_of, double* B, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_face_blocks, dim3((unsigned)n), dim3(64), 0, s, op, face_slot, face_local, M, faces, block_of, B);
  return hipGetLastError();
}

hipError_t launch_face_assemble(const OpDev& op, const int64_t* face_slot, const int32_t* face_local, const double* M,
                                const int32_t* faces, int64_t n, double* nzval, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_face_assemble, dim3((unsigned)n), dim3(256), 0, s, op, face_slot, face_local, M, faces, nzval);
  return hipGetLastError();
}

hipError_t launch_face_energy(const OpDev& op, const int64_t* face_slot, const int32_t* face_local, const double* M,
                              int64_t nfaces, const double* x, double* part, hipStream_t s) {
  if (nfaces <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_face_energy, dim3((unsigned)nfaces), dim3(kThreads), 0, s, op, face_slot, face_local, M, x, part);
  return hipGetLastError();
}

hipError_t launch_sub_scaled(double* y, const double* x, double a, int64_t n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  int64_t blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_sub_scaled, dim3((unsigned)blocks), dim3(256), 0, s, y, x, a, n);
  return hipGetLastError();
}

hipError_t launch_dot_partials(const double* a, const double* b, int64_t n, double scale, double* part, int nblocks,
                               hipStream_t s) {
  hipLaunchKernelGGL(k_dot_partials, dim3((unsigned)nblocks), dim3(kThreads), 0, s, a, b, n, scale, part);
  return hipGetLastError();
}

hipError_t launch_reduce_partials(const double* part, int64_t n, int mode, double* out, hipStream_t s) {
  const dim3 g(1), b(kThreads);
  if (mode == 0) hipLaunchKernelGGL(k_reduce_partials<0>, g, b, 0, s, part, n, out);
  else if (mode == 1) hipLaunchKernelGGL(k_reduce_partials<1>, g, b, 0, s, part, n, out);
  else hipLaunchKernelGGL(k_reduce_partials<2>, g, b, 0, s, part, n, out);
  return hipGetLastError();
}

hipError_t launch_assemble(const OpDev& op, const double* D, double* nzval, BatchRange r, hipStream_t s) {
  if (r.count <= 0) return hipSuccess;
#define CALL(q, p) assemble_qp<q, p>(op, D, nzval, r, s)
  HFK_DISPATCH_QP(op.Q, op.P, CALL);
#undef CALL
}

hipError_t launch_fill(double* y, int64_t n, double v, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  int64_t blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_fill, dim3((unsigned)blocks), dim3(256), 0, s, y, n, v);
  return hipGetLastError();
}

hipError_t launch_expand_tangents(const OpDev& op, int storage, const void* T, double* D, int64_t nbatches,
                                  hipStream_t s) {
  const dim3 g(1024), b(256);
#define EXP(P_)                                                                                           \
  switch (storage) {                                                                                      \
    case 0: hipLaunchKernelGGL((k_expand<P_, double, kLayFull>), g, b, 0, s, op, T, D, nbatches); break;   \
    case 1: hipLaunchKernelGGL((k_expand<P_, double, kLaySym>), g, b, 0, s, op, T, D, nbatches); break;    \
    case 2: hipLaunchKernelGGL((k_expand<P_, float, kLayFull>), g, b, 0, s, op, T, D, nbatches); break;    \
    case 3: hipLaunchKernelGGL((k_expand<P_, float, kLaySym>), g, b, 0, s, op, T, D, nbatches); break;     \
    case 4: hipLaunchKernelGGL((k_expand<P_, double, kLayFac>), g, b, 0, s, op, T, D, nbatches); break;    \
    case 5: hipLaunchKernelGGL((k_expand<P_, float, kLayFac>), g, b, 0, s, op, T, D, nbatches); break;     \
    case 6: hipLaunchKernelGGL((k_expand<P_, double, kLayAfac>), g, b, 0, s, op, T, D, nbatches); break;   \
    default: hipLaunchKernelGGL((k_expand<P_, float, kLayAfac>), g, b, 0, s, op, T, D, nbatches); break;   \
  }
  if (op.P == 2) {
    EXP(2)
  } else if (op.P == 3) {
    EXP(3)
  } else if (op.P == 4) {
    EXP(4)
  } else {
    return hipErrorInvalidValue;
  }
#undef EXP
  return hipGetLastError();
}

}  // namespace hfk
