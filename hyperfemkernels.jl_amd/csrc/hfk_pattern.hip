// Sparsity pattern and cell positions of an assembled operator, built on the device.
//
// The reference's algorithm (sparsity / full_keys! / block_keys! / cell_positions!, src/assembly.jl:773-899, with
// the cross blocks of skeleton_keys!, src/skeleton.jl:355-390): one key per (cell, test dof, trial dof) pair of free
// dofs, restricted to the field blocks set on the cell, with the column in the high bits and the row in the low
// bits; the keys sorted and made unique (rocPRIM radix sort and unique); the CSC column pointers found as the first
// key of every column, the rows taken from the low bits; the stored position of every cell pair found by binary
// search in its column.  All of it is integer work: the arrays equal a host sort of the same keys bit for bit.
//
// Sizes: ncells * nd^2 keys (plus nd^2 per facet neighbour pair), two key buffers of 8 bytes each and rocPRIM's
// temporary storage, all freed on return.  At 40^3 x 6 P2-P1 tetrahedra (34 dofs per cell) that is 444M keys,
// 7.1 GB, far inside one GPU's HBM.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_select.hpp>

#include <algorithm>
#include <cstdint>
#include <vector>

#include "hfk_op.hpp"

namespace hfk {
namespace {

constexpr int kPT = 256;

inline unsigned grid_of(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + kPT - 1) / kPT); }

inline int bit_width(uint64_t v) {
  int b = 0;
  while (v >> b) ++b;
  return std::max(b, 1);
}

// keys of the free pairs of every cell: thread p = (cell * nd + jj) * nd + ii, row dof ii, column dof jj
__global__ __launch_bounds__(kPT) void k_cell_keys(const int32_t* __restrict__ dofs, int nd, int64_t ncells,
                                                   const int8_t* __restrict__ field_of, int nf,
                                                   const uint8_t* __restrict__ mask, int rb, uint64_t sentinel,
                                                   uint64_t* __restrict__ keys) {
  const int64_t p = (int64_t)blockIdx.x * kPT + threadIdx.x;
  const int64_t per = (int64_t)nd * nd;
  if (p >= ncells * per) return;
  const int64_t cell = p / per;
  const int r = (int)(p - cell * per), jj = r / nd, ii = r - jj * nd;
  const int32_t row = dofs[cell * nd + ii], col = dofs[cell * nd + jj];
  bool ok = row > 0 && col > 0;
  if (ok && mask) ok = mask[cell * nf * nf + field_of[ii] + nf * field_of[jj]] != 0;
  keys[p] = ok ? (((uint64_t)(col - 1) << rb) | (uint64_t)(row - 1)) : sentinel;
}

// keys of the cross blocks: pair k = (cell, facet neighbour) couples the neighbour's free rows with the cell's free
// columns; thread p = (k * nd + jj) * nd + ii
__global__ __launch_bounds__(kPT) void k_pair_keys(const int32_t* __restrict__ dofs, int nd, int64_t npairs,
                                                   const int32_t* __restrict__ pairs, int rb, uint64_t sentinel,
                                                   uint64_t* __restrict__ keys) {
  const int64_t p = (int64_t)blockIdx.x * kPT + threadIdx.x;
  const int64_t per = (int64_t)nd * nd;
  if (p >= npairs * per) return;
  const int64_t k = p / per;
  const int r = (int)(p - k * per), jj = r / nd, ii = r - jj * nd;
  const int32_t row = dofs[(int64_t)pairs[2 * k + 1] * nd + ii], col = dofs[(int64_t)pairs[2 * k] * nd + jj];
  keys[p] = (row > 0 && col > 0) ? (((uint64_t)(col - 1) << rb) | (uint64_t)(row - 1)) : sentinel;
}

// first position of column c (0..ncols) in the sorted unique keys
__global__ __launch_bounds__(kPT) void k_colptr(const uint64_t* __restrict__ u, int64_t nnz, int64_t ncols, int rb,
                                                int64_t* __restrict__ colptr) {
  const int64_t c = (int64_t)blockIdx.x * kPT + threadIdx.x;
  if (c > ncols) return;
  const uint64_t key = (uint64_t)c << rb;
  int64_t lo = 0, hi = nnz;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (u[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  colptr[c] = lo;
}

__global__ __launch_bounds__(kPT) void k_rows(const uint64_t* __restrict__ u, int64_t nnz, uint64_t rmask,
                                              int32_t* __restrict__ rowval) {
  const int64_t k = (int64_t)blockIdx.x * kPT + threadIdx.x;
  if (k < nnz) rowval[k] = (int32_t)(u[k] & rmask);
}

// nzpos[cell][ii + nd jj] = 1 + stored position of (row dof ii, column dof jj), 0 when a dof is not free or the
// pair is not stored (a pair outside the cell's blocks that no other cell stores)
__global__ __launch_bounds__(kPT) void k_nzpos(const int32_t* __restrict__ dofs, int nd, int64_t ncells,
                                               const int64_t* __restrict__ colptr, const int32_t* __restrict__ rowval,
                                               int32_t* __restrict__ nzpos) {
  const int64_t p = (int64_t)blockIdx.x * kPT + threadIdx.x;
  const int64_t per = (int64_t)nd * nd;
  if (p >= ncells * per) return;
  const int64_t cell = p / per;
  const int r = (int)(p - cell * per), jj = r / nd, ii = r - jj * nd;
  const int32_t row = dofs[cell * nd + ii], col = dofs[cell * nd + jj];
  int32_t pos = 0;
  if (row > 0 && col > 0) {
    int64_t lo = colptr[col - 1], hi = colptr[col];
    const int64_t end = hi;
    const int32_t key = row - 1;
    while (lo < hi) {
      const int64_t mid = (lo + hi) >> 1;
      if (rowval[mid] < key) lo = mid + 1;
      else hi = mid;
    }
    if (lo < end && rowval[lo] == key) pos = (int32_t)(lo + 1);
  }
  nzpos[p] = pos;
}

__global__ __launch_bounds__(kPT) void k_rows64(const int32_t* __restrict__ rowval, int64_t nnz, int64_t base,
                                                int64_t* __restrict__ out) {
  const int64_t k = (int64_t)blockIdx.x * kPT + threadIdx.x;
  if (k < nnz) out[k] = (int64_t)rowval[k] + base;
}

__global__ __launch_bounds__(kPT) void k_colptr_base(const int64_t* __restrict__ colptr, int64_t n, int64_t base,
                                                     int64_t* __restrict__ out) {
  const int64_t k = (int64_t)blockIdx.x * kPT + threadIdx.x;
  if (k < n) out[k] = colptr[k] + base;
}

// nzpos in another local order: out[cell][perm[ii] + nd perm[jj]] = in[cell][ii + nd jj]
__global__ __launch_bounds__(kPT) void k_nzpos_perm(const int32_t* __restrict__ in, int nd, int64_t ncells,
                                                    const int32_t* __restrict__ perm, int32_t* __restrict__ out) {
  const int64_t p = (int64_t)blockIdx.x * kPT + threadIdx.x;
  const int64_t per = (int64_t)nd * nd;
  if (p >= ncells * per) return;
  const int64_t cell = p / per;
  const int r = (int)(p - cell * per), jj = r / nd, ii = r - jj * nd;
  out[cell * per + perm[ii] + nd * perm[jj]] = in[p];
}

// (row, column) keys of a CSC pattern given 1-based (colptr, rowval): one thread per column; the value of key k is
// its CSC position.  bad |= 1 for a row out of range, 2 for rows that do not ascend within a column.
__global__ __launch_bounds__(kPT) void k_csc_keys(int64_t n, const int64_t* __restrict__ colptr,
                                                  const int64_t* __restrict__ rowval, int cb,
                                                  uint64_t* __restrict__ keys, int32_t* __restrict__ vals,
                                                  int* __restrict__ bad) {
  const int64_t j = (int64_t)blockIdx.x * kPT + threadIdx.x;
  if (j >= n) return;
  int flags = 0;
  for (int64_t k = colptr[j] - 1; k < colptr[j + 1] - 1; ++k) {
    const int64_t r = rowval[k];
    if (r < 1 || r > n) flags |= 1;
    if (k > colptr[j] - 1 && r <= rowval[k - 1]) flags |= 2;
    keys[k] = ((uint64_t)(r - 1) << cb) | (uint64_t)j;
    vals[k] = (int32_t)k;
  }
  if (flags) atomicOr(bad, flags);
}

__global__ __launch_bounds__(kPT) void k_csr_cols(const uint64_t* __restrict__ u, int64_t nnz, uint64_t cmask,
                                                  int32_t* __restrict__ col) {
  const int64_t k = (int64_t)blockIdx.x * kPT + threadIdx.x;
  if (k < nnz) col[k] = (int32_t)(u[k] & cmask);
}

template <class T>
hipError_t upload(DevBuf& b, const T* h, size_t n, hipStream_t s) {
  hipError_t e = b.alloc(std::max<size_t>(n, 1) * sizeof(T));
  if (e != hipSuccess || n == 0) return e;
  return hipMemcpyAsync(b.p, h, n * sizeof(T), hipMemcpyHostToDevice, s);
}

}  // namespace

int device_pattern(hfk_ctx* ctx, hipStream_t s, const int32_t* cell_dofs, int nd, int64_t ncells, int64_t ncols,
                   const BlockInfo& blocks, const CellNeighbours* nbr, DevPattern& out) {
  if (nd <= 0 || ncells < 0 || ncols < 0 || ncols >= ((int64_t)1 << 31))
    return set_err(ctx, HFK_EDIM, "HyperFEMKernels: pattern: bad sizes");
  const int64_t per = (int64_t)nd * nd;
  std::vector<int32_t> pairs;  // (cell, neighbour) of every facet neighbour pair
  if (nbr && !nbr->start.empty())
    for (int64_t cell = 0; cell < ncells; ++cell)
      for (int64_t k = nbr->start[cell]; k < nbr->start[cell + 1]; ++k) {
        pairs.push_back((int32_t)cell);
        pairs.push_back(nbr->cells[k]);
      }
  const int64_t npairs = (int64_t)pairs.size() / 2;
  const int64_t ncell_keys = ncells * per, nkeys = ncell_keys + npairs * per;
  const int rb = bit_width((uint64_t)std::max<int64_t>(ncols, 1));  // rows and columns are < ncols
  const int cb = bit_width((uint64_t)ncols);                         // the sentinel's column ncols fits cb bits
  const uint64_t sentinel = (uint64_t)ncols << rb;                   // sorts after every stored key
  DevBuf dofs_d, field_d, mask_d, pairs_d, keys, sorted, tmp, count_d;
  HIPCHK(ctx, upload(dofs_d, cell_dofs, (size_t)ncells * nd, s));
  if (blocks.mask) {
    HIPCHK(ctx, upload(field_d, blocks.field_of, (size_t)nd, s));
    HIPCHK(ctx, upload(mask_d, blocks.mask, (size_t)ncells * blocks.nf * blocks.nf, s));
  }
  HIPCHK(ctx, upload(pairs_d, pairs.data(), pairs.size(), s));
  HIPCHK(ctx, keys.alloc(std::max<int64_t>(nkeys, 1) * sizeof(uint64_t)));
  HIPCHK(ctx, sorted.alloc(std::max<int64_t>(nkeys, 1) * sizeof(uint64_t)));
  if (ncell_keys > 0)
    hipLaunchKernelGGL(k_cell_keys, dim3(grid_of(ncell_keys)), dim3(kPT), 0, s, dofs_d.as<int32_t>(), nd, ncells,
                       field_d.as<int8_t>(), blocks.nf, mask_d.as<uint8_t>(), rb, sentinel, keys.as<uint64_t>());
  if (npairs > 0)
    hipLaunchKernelGGL(k_pair_keys, dim3(grid_of(npairs * per)), dim3(kPT), 0, s, dofs_d.as<int32_t>(), nd, npairs,
                       pairs_d.as<int32_t>(), rb, sentinel, keys.as<uint64_t>() + ncell_keys);
  HIPCHK(ctx, hipGetLastError());
  int64_t nnz = 0;
  if (nkeys > 0) {
    // sort!, then unique! (the unique keys go back into `keys`)
    size_t tb_sort = 0, tb_uniq = 0;
    HIPCHK(ctx, rocprim::radix_sort_keys(nullptr, tb_sort, keys.as<uint64_t>(), sorted.as<uint64_t>(), (size_t)nkeys,
                                         0, rb + cb, s));
    HIPCHK(ctx, rocprim::unique(nullptr, tb_uniq, sorted.as<uint64_t>(), keys.as<uint64_t>(), count_d.as<size_t>(),
                                (size_t)nkeys, rocprim::equal_to<uint64_t>(), s));
    HIPCHK(ctx, tmp.alloc(std::max<size_t>(std::max(tb_sort, tb_uniq), 1)));
    HIPCHK(ctx, count_d.alloc(sizeof(size_t)));
    HIPCHK(ctx, rocprim::radix_sort_keys(tmp.p, tb_sort, keys.as<uint64_t>(), sorted.as<uint64_t>(), (size_t)nkeys,
                                         0, rb + cb, s));
    HIPCHK(ctx, rocprim::unique(tmp.p, tb_uniq, sorted.as<uint64_t>(), keys.as<uint64_t>(), count_d.as<size_t>(),
                                (size_t)nkeys, rocprim::equal_to<uint64_t>(), s));
    size_t nuniq = 0;
    HIPCHK(ctx, hipMemcpyAsync(&nuniq, count_d.p, sizeof(size_t), hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipStreamSynchronize(s));
    uint64_t last = 0;
    if (nuniq > 0) {
      HIPCHK(ctx, hipMemcpyAsync(&last, keys.as<uint64_t>() + nuniq - 1, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
      HIPCHK(ctx, hipStreamSynchronize(s));
    }
    nnz = (int64_t)nuniq - ((nuniq > 0 && last == sentinel) ? 1 : 0);
  }
  if (nnz >= ((int64_t)1 << 31) - 1)
    return set_err(ctx, HFK_EDIM, "HyperFEMKernels: pattern: more than 2^31 - 2 stored entries");
  tmp.reset();
  sorted.reset();
  out.nnz = nnz;
  HIPCHK(ctx, out.colptr.alloc((size_t)(ncols + 1) * sizeof(int64_t)));
  HIPCHK(ctx, out.rowval.alloc((size_t)std::max<int64_t>(nnz, 1) * sizeof(int32_t)));
  HIPCHK(ctx, out.nzpos.alloc((size_t)std::max<int64_t>(ncell_keys, 1) * sizeof(int32_t)));
  hipLaunchKernelGGL(k_colptr, dim3(grid_of(ncols + 1)), dim3(kPT), 0, s, keys.as<uint64_t>(), nnz, ncols, rb,
                     out.colptr.as<int64_t>());
  if (nnz > 0)
    hipLaunchKernelGGL(k_rows, dim3(grid_of(nnz)), dim3(kPT), 0, s, keys.as<uint64_t>(), nnz,
                       ((uint64_t)1 << rb) - 1, out.rowval.as<int32_t>());
  if (ncell_keys > 0)
    hipLaunchKernelGGL(k_nzpos, dim3(grid_of(ncell_keys)), dim3(kPT), 0, s, dofs_d.as<int32_t>(), nd, ncells,
                       out.colptr.as<int64_t>(), out.rowval.as<int32_t>(), out.nzpos.as<int32_t>());
  HIPCHK(ctx, hipGetLastError());
  HIPCHK(ctx, hipStreamSynchronize(s));  // the local buffers are freed on return
  return HFK_OK;
}

int device_csc_to_csr(hfk_ctx* ctx, hipStream_t s, int64_t n, const int64_t* colptr, const int64_t* rowval,
                      DevBuf& rp, DevBuf& col, DevBuf& perm, int* bad) {
  const int64_t nnz = colptr[n] - 1;
  const int cb = bit_width((uint64_t)std::max<int64_t>(n - 1, 1));  // columns 0..n-1
  const int rb = bit_width((uint64_t)std::max<int64_t>(n - 1, 1));  // rows 0..n-1
  DevBuf cp_d, rv_d, keys, sorted, vals, vals_s, tmp, bad_d;
  HIPCHK(ctx, upload(cp_d, colptr, (size_t)n + 1, s));
  HIPCHK(ctx, upload(rv_d, rowval, (size_t)nnz, s));
  HIPCHK(ctx, keys.alloc(std::max<int64_t>(nnz, 1) * sizeof(uint64_t)));
  HIPCHK(ctx, sorted.alloc(std::max<int64_t>(nnz, 1) * sizeof(uint64_t)));
  HIPCHK(ctx, vals.alloc(std::max<int64_t>(nnz, 1) * sizeof(int32_t)));
  HIPCHK(ctx, vals_s.alloc(std::max<int64_t>(nnz, 1) * sizeof(int32_t)));
  HIPCHK(ctx, bad_d.alloc(sizeof(int)));
  HIPCHK(ctx, hipMemsetAsync(bad_d.p, 0, sizeof(int), s));
  hipLaunchKernelGGL(k_csc_keys, dim3(grid_of(n)), dim3(kPT), 0, s, n, cp_d.as<int64_t>(), rv_d.as<int64_t>(), cb,
                     keys.as<uint64_t>(), vals.as<int32_t>(), bad_d.as<int>());
  HIPCHK(ctx, hipGetLastError());
  int flags = 0;
  HIPCHK(ctx, hipMemcpyAsync(&flags, bad_d.p, sizeof(int), hipMemcpyDeviceToHost, s));
  HIPCHK(ctx, hipStreamSynchronize(s));
  *bad = flags;
  if (flags) return HFK_OK;
  rv_d.reset();
  HIPCHK(ctx, rp.alloc((size_t)(n + 1) * sizeof(int64_t)));
  HIPCHK(ctx, col.alloc(std::max<int64_t>(nnz, 1) * sizeof(int32_t)));
  HIPCHK(ctx, perm.alloc(std::max<int64_t>(nnz, 1) * sizeof(int32_t)));
  if (nnz > 0) {
    // rows in the high bits: the sort gives the CSR order, columns ascending within a row (the keys are distinct)
    size_t tb = 0;
    HIPCHK(ctx, rocprim::radix_sort_pairs(nullptr, tb, keys.as<uint64_t>(), sorted.as<uint64_t>(), vals.as<int32_t>(),
                                          vals_s.as<int32_t>(), (size_t)nnz, 0, rb + cb, s));
    HIPCHK(ctx, tmp.alloc(std::max<size_t>(tb, 1)));
    HIPCHK(ctx, rocprim::radix_sort_pairs(tmp.p, tb, keys.as<uint64_t>(), sorted.as<uint64_t>(), vals.as<int32_t>(),
                                          vals_s.as<int32_t>(), (size_t)nnz, 0, rb + cb, s));
    hipLaunchKernelGGL(k_csr_cols, dim3(grid_of(nnz)), dim3(kPT), 0, s, sorted.as<uint64_t>(), nnz,
                       ((uint64_t)1 << cb) - 1, col.as<int32_t>());
    HIPCHK(ctx, hipMemcpyAsync(perm.p, vals_s.p, (size_t)nnz * sizeof(int32_t), hipMemcpyDeviceToDevice, s));
  }
  hipLaunchKernelGGL(k_colptr, dim3(grid_of(n + 1)), dim3(kPT), 0, s, sorted.as<uint64_t>(), nnz, n, cb,
                     rp.as<int64_t>());  // the first position of every row
  HIPCHK(ctx, hipGetLastError());
  HIPCHK(ctx, hipStreamSynchronize(s));  // the local buffers are freed on return
  return HFK_OK;
}

int device_pattern_to_host(hfk_ctx* ctx, hipStream_t s, const DevPattern& pt, int64_t ncols, int64_t ncells, int nd,
                           int64_t* colptr, int64_t* rowval, int32_t* nzpos, const int32_t* perm) {
  DevBuf stage;
  if (colptr) {
    HIPCHK(ctx, stage.alloc((size_t)(ncols + 1) * sizeof(int64_t)));
    hipLaunchKernelGGL(k_colptr_base, dim3(grid_of(ncols + 1)), dim3(kPT), 0, s, pt.colptr.as<int64_t>(), ncols + 1,
                       (int64_t)1, stage.as<int64_t>());
    HIPCHK(ctx, hipGetLastError());
    HIPCHK(ctx, hipMemcpyAsync(colptr, stage.p, (size_t)(ncols + 1) * sizeof(int64_t), hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipStreamSynchronize(s));
  }
  if (rowval && pt.nnz > 0) {
    HIPCHK(ctx, stage.alloc((size_t)pt.nnz * sizeof(int64_t)));
    hipLaunchKernelGGL(k_rows64, dim3(grid_of(pt.nnz)), dim3(kPT), 0, s, pt.rowval.as<int32_t>(), pt.nnz, (int64_t)1,
                       stage.as<int64_t>());
    HIPCHK(ctx, hipGetLastError());
    HIPCHK(ctx, hipMemcpyAsync(rowval, stage.p, (size_t)pt.nnz * sizeof(int64_t), hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipStreamSynchronize(s));
  }
  const int64_t n = ncells * nd * nd;
  if (nzpos && n > 0) {
    const int32_t* src = pt.nzpos.as<int32_t>();
    DevBuf perm_d;
    if (perm) {
      HIPCHK(ctx, upload(perm_d, perm, (size_t)nd, s));
      HIPCHK(ctx, stage.alloc((size_t)n * sizeof(int32_t)));
      hipLaunchKernelGGL(k_nzpos_perm, dim3(grid_of(n)), dim3(kPT), 0, s, src, nd, ncells, perm_d.as<int32_t>(),
                         stage.as<int32_t>());
      HIPCHK(ctx, hipGetLastError());
      src = stage.as<int32_t>();
    }
    HIPCHK(ctx, hipMemcpyAsync(nzpos, src, (size_t)n * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipStreamSynchronize(s));
  }
  return HFK_OK;
}

}  // namespace hfk
