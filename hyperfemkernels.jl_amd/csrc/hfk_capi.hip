// Host side of libhfk: the C ABI of include/hfk.h.  Handles own device memory; host arrays are borrowed.
//
// Set-up work that the reference does once per operator on the host stays on the host here too: the greedy
// cell colouring (src/coloring.jl:4-28), the sparsity pattern with its positions (src/assembly.jl:773-899) and
// the nodal blocks (src/preconditioners.jl:151-171), all integer algorithms whose results (colours,
// colptr/rowval, nzpos, block numbering) are bit-identical to the reference's on the same cell dof ids.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "hfk_op.hpp"

namespace hfk {
bool supported_qp(int Q, int P);
}

using namespace hfk;

namespace hfk {

thread_local std::string t_err;

// Scratch of the allocation-free energy (hfk_energy_noalloc_dev): its partial sums (one per batch, 2 x 256 for a
// traction load, one per Robin face).  One copy per device, held until the process ends.  In host code `scratch`
// names a host-side stand-in, not a device address: kernels only ever receive the address that
// hipGetSymbolAddress returns for the current device (kept per device in g_scratch_ptr, copied into each context).
constexpr int64_t kScratchDoubles = int64_t(1) << 17;  // 1 MiB: up to 131,072 partial sums
static __device__ __align__(256) double scratch[kScratchDoubles];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_m;
double* g_scratch_ptr[kMaxDevices] = {};

// once per device, with that device current: the device address of `scratch`
int scratch_setup(hfk_ctx* ctx) {
  if (ctx->device >= kMaxDevices)
    return set_err(nullptr, HFK_EARG, "HyperFEMKernels: device index beyond the library's device table");
  std::lock_guard<std::mutex> lk(g_scratch_m);
  if (!g_scratch_ptr[ctx->device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return set_err(nullptr, HFK_EHIP, std::string("HyperFEMKernels: hipGetSymbolAddress(scratch): ") +
                                            hipGetErrorString(e));
    g_scratch_ptr[ctx->device] = static_cast<double*>(p);
  }
  ctx->scratch = g_scratch_ptr[ctx->device];
  return HFK_OK;
}

int set_err(hfk_ctx* ctx, int code, const std::string& msg) {
  t_err = msg;
  if (ctx) ctx->err = msg;
  return code;
}

int op_bind(hfk_op* op) {
  if (!op || !op->ctx) return set_err(nullptr, HFK_EARG, "HyperFEMKernels: null operator");
  HIPCHK(op->ctx, hipSetDevice(op->ctx->device));
  return HFK_OK;
}

// prepare!: tangents at the state x_dev; a device copy of the state is kept (the reference's mf.x,
// src/matrixfree.jl:140-143), read by the coarse levels of p-multigrid
int op_run_prepare(hfk_op* op, const double* x_dev, int storage, hipStream_t s) {
  if (storage < 0 || storage > 7) return set_err(op->ctx, HFK_EARG, "HyperFEMKernels: unknown tangent storage");
  const size_t bytes = (size_t)op->nbatches * store_entries(storage) * op->C * op->np * store_bytes(storage);
  HIPCHK(op->ctx, op->tangents.alloc(bytes));
  HIPCHK(op->ctx, op->state.alloc((size_t)op->nfree * sizeof(double)));
  if (x_dev != op->state.as<double>())
    HIPCHK(op->ctx, hipMemcpyAsync(op->state.p, x_dev, (size_t)op->nfree * sizeof(double), hipMemcpyDeviceToDevice, s));
  const int rc = op_prepare_into(op, storage, op->tangents.p, x_dev, s);
  if (rc) return rc;
  op->storage = storage;
  return HFK_OK;
}

int op_prepare_into(hfk_op* op, int storage, void* T, const double* x_dev, hipStream_t s) {
  if (storage < 0 || storage > 7) return set_err(op->ctx, HFK_EARG, "HyperFEMKernels: unknown tangent storage");
  if (store_factorised(storage) && !op->rtc && op->law == 0)
    return set_err(op->ctx, HFK_EARG, "HyperFEMKernels: the factorised tangent storage needs the neo-Hookean or the "
                                      "linear law (St Venant-Kirchhoff has no such form)");
  if (store_affine(storage)) {
    if (!op->affine)
      return set_err(op->ctx, HFK_EARG, "HyperFEMKernels: the affine factorised storage needs parallelepiped cells "
                                        "(an affine map per cell); use the factorised storage");
    HIPCHK(op->ctx, op->cellgeo.alloc((size_t)op->nslots() * 8 * sizeof(double)));
  }
  if (op->rtc) return rtc_prepare(op, storage, T, x_dev, s);
  HIPCHK(op->ctx, launch_prepare(op->dev(), storage, T, x_dev, BatchRange{0, (int)op->nbatches}, s));
  return HFK_OK;
}

// y = J dx (accumulate = false) or y += alpha J dx (accumulate = true)
int op_run_apply_t(hfk_op* op, const void* T, int storage, const double* dx_dev, double* y_dev, hipStream_t s,
                   double alpha, bool accumulate) {
  if (storage < 0 || !T)
    return set_err(op->ctx, HFK_ESTATE, "HyperFEMKernels: mul! before prepare! (no tangents stored)");
  if (!accumulate) HIPCHK(op->ctx, launch_fill(y_dev, op->nfree, 0.0, s));
  const OpDev d = op->dev();
  if (!op->deterministic) {
    HIPCHK(op->ctx, launch_apply(d, storage, T, dx_dev, y_dev, BatchRange{0, 0}, s, accumulate ? alpha : 1.0, &op->amap));
  } else {
    for (const BatchRange& r : op->color_batches)
      HIPCHK(op->ctx, launch_apply(d, storage, T, dx_dev, y_dev, r, s, accumulate ? alpha : 1.0));
  }
  // the boundary group of the Robin faces (its Jacobian is the face mass matrix; increments vanish on the
  // Dirichlet dofs), then the skeleton group at the prepared state
  int rc = robin_apply(op, dx_dev, nullptr, y_dev, accumulate ? alpha : 1.0, s);
  if (rc) return rc;
  return skel_apply(op, dx_dev, y_dev, accumulate ? alpha : 1.0, s);
}

int robin_apply(hfk_op* op, const double* uf, const double* ud, double* y, double alpha, hipStream_t s) {
  if (op->robin_n == 0) return HFK_OK;
  const OpDev d = op->dev();
  for (size_t g = 0; g + 1 < op->robin_start.size(); ++g)
    HIPCHK(op->ctx, launch_face_apply(d, op->robin_slot.as<int64_t>(), op->robin_local.as<int32_t>(),
                                      op->robin_M.as<double>(), op->robin_order.as<int32_t>() + op->robin_start[g],
                                      op->robin_start[g + 1] - op->robin_start[g], uf, ud, y, alpha, s));
  return HFK_OK;
}

int robin_blocks(hfk_op* op, const int32_t* block_of, double* B, hipStream_t s) {
  if (op->robin_n == 0) return HFK_OK;
  const OpDev d = op->dev();
  for (size_t g = 0; g + 1 < op->robin_start.size(); ++g)
    HIPCHK(op->ctx, launch_face_blocks(d, op->robin_slot.as<int64_t>(), op->robin_local.as<int32_t>(),
                                       op->robin_M.as<double>(), op->robin_order.as<int32_t>() + op->robin_start[g],
                                       op->robin_start[g + 1] - op->robin_start[g], block_of, B, s));
  return HFK_OK;
}

int robin_assemble(hfk_op* op, double* nzval, hipStream_t s) {
  if (op->robin_n == 0) return HFK_OK;
  const OpDev d = op->dev();
  for (size_t g = 0; g + 1 < op->robin_start.size(); ++g)
    HIPCHK(op->ctx, launch_face_assemble(d, op->robin_slot.as<int64_t>(), op->robin_local.as<int32_t>(),
                                         op->robin_M.as<double>(), op->robin_order.as<int32_t>() + op->robin_start[g],
                                         op->robin_start[g + 1] - op->robin_start[g], nzval, s));
  return HFK_OK;
}

int op_run_apply(hfk_op* op, const double* dx_dev, double* y_dev, hipStream_t s, double alpha, bool accumulate) {
  return op_run_apply_t(op, op->tangents.p, op->storage, dx_dev, y_dev, s, alpha, accumulate);
}

// sparsity (src/assembly.jl:773-818, full_keys!/cell_keys! :835-893) and cell_positions! (:864-880), with the
// cross blocks of the skeleton (src/skeleton.jl:355-390): built on the device (hfk_pattern.hip), then copied into
// the operator's host arrays in the reference's form (1-based Int64 colptr and rowval, nzpos per cell), which the
// coarse-level set-up, the skeleton and the host API read.
int op_build_pattern(hfk_op* op) {
  if (op->pattern_built) return HFK_OK;
  CellNeighbours nb;
  const bool skel = skel_facets(op) > 0;
  if (skel) skel_neighbours(op, nb);
  DevPattern pt;
  int rc = device_pattern(op->ctx, op->ctx->stream, op->cell_dofs.data(), op->nd, op->ncells, op->nfree, BlockInfo(),
                          skel ? &nb : nullptr, pt);
  if (rc) return rc;
  op->colptr.resize((size_t)op->nfree + 1);
  op->rowval.resize((size_t)pt.nnz);
  op->nzpos.resize((size_t)op->ncells * op->nd * op->nd);
  if ((rc = device_pattern_to_host(op->ctx, op->ctx->stream, pt, op->nfree, op->ncells, op->nd, op->colptr.data(),
                                   op->rowval.data(), op->nzpos.data(), nullptr)))
    return rc;
  op->pattern_built = true;
  return HFK_OK;
}

// src/coloring.jl:4-28: greedy first fit over the free (test) dofs with 128-bit masks; colour = trailing_ones.
// Generic over the cell dof table; returns false when more than 128 colours are needed.
bool greedy_colors(const int32_t* cell_dofs, int nd, int64_t ncells, int64_t nfree, std::vector<int32_t>& color) {
  std::vector<unsigned __int128> masks((size_t)nfree, 0);
  color.assign((size_t)ncells, 0);
  for (int64_t cell = 0; cell < ncells; ++cell) {
    unsigned __int128 used = 0;
    const int32_t* ids = &cell_dofs[(size_t)cell * nd];
    for (int j = 0; j < nd; ++j)
      if (ids[j] > 0) used |= masks[ids[j] - 1];
    int c = 0;
    while (c < 128 && ((used >> c) & 1)) ++c;
    if (c >= 128) return false;
    color[cell] = c + 1;
    const unsigned __int128 bit = (unsigned __int128)1 << c;
    for (int j = 0; j < nd; ++j)
      if (ids[j] > 0) masks[ids[j] - 1] |= bit;
  }
  return true;
}


// field_blocks (src/preconditioners.jl:151-171): a block per node with a free component, first-seen order over
// (cell, table node); block_of in the batch layout, tensor node order
int op_build_blocks(hfk_op* op) {
  NodeBlocks& nb = op->blocks;
  if (nb.built) return HFK_OK;
  const int nn = op->nn, nd = op->nd;
  std::vector<int32_t> block_of_dof((size_t)op->nfree, -1);
  std::vector<int32_t> blk_cell((size_t)op->ncells * nn, -1);
  nb.dofs_h.clear();
  int64_t count = 0;
  for (int64_t cell = 0; cell < op->ncells; ++cell)
    for (int n = 0; n < nn; ++n) {
      int32_t key[3];
      int32_t first = 0;
      for (int c = 0; c < 3; ++c) {
        key[c] = op->cell_dofs[(size_t)cell * nd + op->ldof[c * nn + n]];
        if (key[c] > 0 && first == 0) first = key[c];
      }
      if (first == 0) continue;  // every component fixed: no block
      int32_t blk = block_of_dof[first - 1];
      if (blk < 0) {
        blk = (int32_t)count++;
        for (int c = 0; c < 3; ++c) {
          nb.dofs_h.push_back(key[c] > 0 ? key[c] : 0);
          if (key[c] > 0) block_of_dof[key[c] - 1] = blk;
        }
      }
      blk_cell[(size_t)cell * nn + n] = blk;
    }
  nb.nblocks = count;
  std::vector<int32_t> h((size_t)op->nslots() * nn, -1);
  for (int64_t s = 0; s < op->nslots(); ++s) {
    const int64_t cell = op->batch_cell[s];
    if (cell < 0) continue;
    for (int t = 0; t < nn; ++t) h[(size_t)s * nn + t] = blk_cell[(size_t)cell * nn + op->nodeperm[t]];
  }
  HIPCHK(op->ctx, nb.block_of.alloc(h.size() * sizeof(int32_t)));
  HIPCHK(op->ctx, hipMemcpy(nb.block_of.p, h.data(), h.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  HIPCHK(op->ctx, nb.dofs.alloc(std::max<size_t>(nb.dofs_h.size(), 1) * sizeof(int32_t)));
  if (!nb.dofs_h.empty())
    HIPCHK(op->ctx, hipMemcpy(nb.dofs.p, nb.dofs_h.data(), nb.dofs_h.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  HIPCHK(op->ctx, nb.B.alloc(std::max<int64_t>(count, 1) * 9 * sizeof(double)));
  HIPCHK(op->ctx, nb.Binv.alloc(std::max<int64_t>(count, 1) * 9 * sizeof(double)));
  HIPCHK(op->ctx, nb.bad.alloc(sizeof(int)));
  nb.built = true;
  return HFK_OK;
}

namespace {

// 1D Lagrange basis on Q equispaced nodes of [0,1] (midpoint if Q = 1) at xi (lagrange1d, src/tensorproduct.jl:24-47)
void lagrange1d(int Q, double xi, double* N, double* dN) {
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

// Gauss-Legendre points (ascending) and weights on [0, 1]: Newton iteration on the Legendre three-term recurrence.
void gauss01(int P, double* x, double* w) {
  for (int i = 0; i < P; ++i) {
    double z = std::cos(M_PI * (i + 0.75) / (P + 0.5)), dp = 1;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1, p1 = z;
      for (int j = 2; j <= P; ++j) {
        const double p2 = ((2 * j - 1) * z * p1 - (j - 1) * p0) / j;
        p0 = p1;
        p1 = p2;
      }
      dp = P * (z * p1 - p0) / (z * z - 1);
      const double dz = p1 / dp;
      z -= dz;
      if (std::fabs(dz) < 1e-16) break;
    }
    x[i] = 0.5 * (1 - z);
    w[i] = 1.0 / ((1 - z * z) * dp * dp);
  }
}

int color_cells(hfk_op* op) {
  if (!greedy_colors(op->cell_dofs.data(), op->nd, op->ncells, op->nfree, op->color))
    return set_err(op->ctx, HFK_ESTATE, "HyperFEMKernels: more than 128 colours needed");
  return HFK_OK;
}

int sync(hfk_op* op) {
  HIPCHK(op->ctx, hipStreamSynchronize(op->ctx->stream));
  return HFK_OK;
}

int bind(hfk_op* op) { return op_bind(op); }

int run_prepare(hfk_op* op, const double* x_dev, int storage, hipStream_t s) {
  return op_run_prepare(op, x_dev, storage, s);
}

int run_apply(hfk_op* op, const double* dx_dev, double* y_dev, hipStream_t s) {
  return op_run_apply(op, dx_dev, y_dev, s, 1.0, false);
}

int run_residual(hfk_op* op, const double* x_dev, double* b_dev, hipStream_t s) {
  if (op->rtc) {
    int rc = rtc_residual(op, x_dev, b_dev, s);
    if (rc) return rc;
  } else {
    HIPCHK(op->ctx, launch_fill(b_dev, op->nfree, 0.0, s));
    const OpDev d = op->dev();
    for (const BatchRange& r : op->color_batches) HIPCHK(op->ctx, launch_residual(d, x_dev, b_dev, r, s));
  }
  // the boundary groups after the volume group: minus the traction load on the free rows, plus the springs
  if (op->has_traction)
    HIPCHK(op->ctx, launch_sub_scaled(b_dev, op->face_load.as<double>(), op->load_factor, op->nfree, s));
  int rc = robin_apply(op, x_dev, op->dvals.as<double>(), b_dev, 1.0, s);
  if (rc) return rc;
  return skel_residual(op, x_dev, b_dev, s);  // the skeleton group last
}

// strain energy into out_dev[0]: per-batch sums, then their sum, both in a fixed order
constexpr int kDotBlocks = 256;  // work-groups of the load's dot products (a fixed grid: a fixed order)

// number of partial sums of one energy evaluation
int64_t energy_partials(const hfk_op* op) {
  return op->nbatches + (op->has_traction ? 2 * kDotBlocks : 0) + op->robin_n + skel_facets(op);
}

// the energy with its partial sums in `part` (energy_partials(op) doubles, every one overwritten)
int run_energy_into(hfk_op* op, const double* x_dev, double* out_dev, double* part, hipStream_t s) {
  const int64_t nload = op->has_traction ? 2 * kDotBlocks : 0;
  if (op->rtc) {
    int rc = rtc_energy(op, x_dev, part, s);
    if (rc) return rc;
  } else {
    HIPCHK(op->ctx, launch_energy(op->dev(), x_dev, part, op->nbatches, s));
  }
  if (op->has_traction) {  // minus the work of the load: f . u over the free and the Dirichlet values
    const double* f = op->face_load.as<double>();
    HIPCHK(op->ctx, launch_dot_partials(f, x_dev, op->nfree, -op->load_factor, part + op->nbatches, kDotBlocks, s));
    HIPCHK(op->ctx, launch_dot_partials(f + op->nfree, op->dvals.as<double>(), op->ndir, -op->load_factor,
                                        part + op->nbatches + kDotBlocks, kDotBlocks, s));
  }
  if (op->robin_n > 0)  // plus the energy of the springs, one partial per face
    HIPCHK(op->ctx, launch_face_energy(op->dev(), op->robin_slot.as<int64_t>(), op->robin_local.as<int32_t>(),
                                       op->robin_M.as<double>(), op->robin_n, x_dev, part + op->nbatches + nload, s));
  if (skel_facets(op) > 0) {  // plus the facet energies, one partial per facet
    int rc = skel_energy(op, x_dev, part + op->nbatches + nload + op->robin_n, s);
    if (rc) return rc;
  }
  HIPCHK(op->ctx, launch_reduce_partials(part, energy_partials(op), 0, out_dev, s));
  return HFK_OK;
}

int run_energy(hfk_op* op, const double* x_dev, double* out_dev, hipStream_t s) {
  HIPCHK(op->ctx, op->part.alloc((size_t)std::max<int64_t>(energy_partials(op), 1) * sizeof(double)));
  return run_energy_into(op, x_dev, out_dev, op->part.as<double>(), s);
}

int stage_in(hfk_op* op, DevBuf& buf, const double* h, int64_t n) {
  HIPCHK(op->ctx, buf.alloc((size_t)std::max<int64_t>(n, 1) * sizeof(double)));
  if (n > 0) HIPCHK(op->ctx, hipMemcpyAsync(buf.p, h, (size_t)n * sizeof(double), hipMemcpyHostToDevice, op->ctx->stream));
  return HFK_OK;
}

int build_pattern(hfk_op* op) { return op_build_pattern(op); }

int upload_nzpos(hfk_op* op) {
  if (op->nzpos_d.p) return HFK_OK;
  const size_t per = (size_t)op->nd * op->nd;
  std::vector<int32_t> h((size_t)op->nbatches * op->C * per, 0);
  for (int64_t cell = 0; cell < op->ncells; ++cell)
    std::memcpy(&h[(size_t)op->slot_of_cell[cell] * per], &op->nzpos[(size_t)cell * per], per * sizeof(int32_t));
  HIPCHK(op->ctx, op->nzpos_d.alloc(h.size() * sizeof(int32_t)));
  HIPCHK(op->ctx, hipMemcpy(op->nzpos_d.p, h.data(), h.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  return HFK_OK;
}

}  // namespace

// jacobian!: tangents in FP64 full storage into a buffer of their own (the matrix-free tangents are left
// untouched), zeroed nonzeros, element matrices summed colour by colour
int op_run_jacobian(hfk_op* op, const double* x_dev, double* nz_dev, hipStream_t s) {
  int rc = build_pattern(op);
  if (rc) return rc;
  if ((rc = upload_nzpos(op))) return rc;
  const size_t bytes = (size_t)op->nbatches * 81 * op->C * op->np * sizeof(double);
  HIPCHK(op->ctx, op->D64.alloc(bytes));
  const OpDev d = op->dev();
  if ((rc = op_prepare_into(op, 0, op->D64.p, x_dev, s))) return rc;
  HIPCHK(op->ctx, launch_fill(nz_dev, (int64_t)op->rowval.size(), 0.0, s));
  for (const BatchRange& r : op->color_batches) HIPCHK(op->ctx, launch_assemble(d, op->D64.as<double>(), nz_dev, r, s));
  if ((rc = robin_assemble(op, nz_dev, s))) return rc;
  return skel_assemble(op, x_dev, nz_dev, s);
}

}  // namespace hfk

namespace {
int run_jacobian(hfk_op* op, const double* x_dev, double* nz_dev, hipStream_t s) {
  return op_run_jacobian(op, x_dev, nz_dev, s);
}
}  // namespace
namespace {

// Faces grouped by the colour of their cell and their local face: the faces of one launch group belong to different
// cells of one colour, so they share no dof (two faces of one cell share the nodes of their common edge).  order:
// face indices group by group, start: group bounds, slot: the batch slot of each face's cell.
void face_groups(const hfk_op* op, int64_t nfaces, const int32_t* face_cell, const int32_t* face_local,
                 std::vector<int32_t>& order, std::vector<int64_t>& start, std::vector<int64_t>& slot) {
  const int nc = *std::max_element(op->color.begin(), op->color.end());
  const int ng = 6 * nc;
  auto group = [&](int64_t f) { return 6 * (op->color[face_cell[f]] - 1) + face_local[f]; };
  start.assign((size_t)ng + 1, 0);
  for (int64_t f = 0; f < nfaces; ++f) start[group(f) + 1]++;
  for (int g = 0; g < ng; ++g) start[g + 1] += start[g];
  std::vector<int64_t> fill(start.begin(), start.end() - 1);
  order.assign((size_t)nfaces, 0);
  slot.assign((size_t)nfaces, 0);
  for (int64_t f = 0; f < nfaces; ++f) {
    order[fill[group(f)]++] = (int32_t)f;
    slot[f] = op->slot_of_cell[face_cell[f]];
  }
}

int check_faces(hfk_op* op, int64_t nfaces, const int32_t* face_cell, const int32_t* face_local, const double* v,
                int nv) {
  if (nfaces < 0 || !face_cell || !face_local || !v) return set_err(op->ctx, HFK_EARG, "HyperFEMKernels: invalid faces");
  for (int64_t f = 0; f < nfaces; ++f) {
    if (face_cell[f] < 0 || face_cell[f] >= op->ncells || face_local[f] < 0 || face_local[f] > 5)
      return set_err(op->ctx, HFK_EDIM, "HyperFEMKernels: face out of range (cell, local face 0..5)");
    for (int c = 0; c < nv; ++c)
      if (!std::isfinite(v[nv * f + c])) return set_err(op->ctx, HFK_EARG, "HyperFEMKernels: face data not finite");
  }
  // a face listed twice would land in one launch group, whose faces must not share dofs
  std::vector<int64_t> key((size_t)nfaces);
  for (int64_t f = 0; f < nfaces; ++f) key[f] = (int64_t)face_cell[f] * 6 + face_local[f];
  std::sort(key.begin(), key.end());
  if (std::adjacent_find(key.begin(), key.end()) != key.end())
    return set_err(op->ctx, HFK_EARG, "HyperFEMKernels: a (cell, local face) pair is listed twice");
  return HFK_OK;
}

}  // namespace

extern "C" {

int hfk_version(void) { return 1; }

const char* hfk_last_error(const hfk_ctx* ctx) { return ctx ? ctx->err.c_str() : t_err.c_str(); }

int hfk_open(int device, hfk_ctx** out) {
  if (!out) return set_err(nullptr, HFK_EARG, "HyperFEMKernels: null output pointer");
  *out = nullptr;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0) return set_err(nullptr, HFK_EHIP, "HyperFEMKernels: no HIP device available");
  if (device < 0 || device >= n) return set_err(nullptr, HFK_EARG, "HyperFEMKernels: device index out of range");
  auto* ctx = new hfk_ctx();
  ctx->device = device;
  e = hipSetDevice(device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete ctx;
    return set_err(nullptr, HFK_EHIP, std::string("HyperFEMKernels: ") + hipGetErrorString(e));
  }
  const int rc = scratch_setup(ctx);
  if (rc) {
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return rc;
  }
  *out = ctx;
  return HFK_OK;
}

int hfk_close(hfk_ctx* ctx) {
  if (!ctx) return HFK_OK;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return HFK_OK;
}

int hfk_synchronize(hfk_ctx* ctx) {
  if (!ctx) return set_err(nullptr, HFK_EARG, "HyperFEMKernels: null context");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return HFK_OK;
}

int hfk_op_create(hfk_ctx* ctx, int Q, int P, int64_t ncells, const int32_t* cell_dofs, const int32_t* ldof,
                  const int32_t* nodeperm, const double* coords, int64_t nfree, int64_t ndir, int law, double lam,
                  double mu, hfk_op** out) {
  if (!ctx || !out) return set_err(ctx, HFK_EARG, "HyperFEMKernels: null context or output pointer");
  *out = nullptr;
  if (!supported_qp(Q, P))
    return set_err(ctx, HFK_EARG, "HyperFEMKernels: unsupported (nodes, points) per direction: Q in 2..4, P in Q..4");
  if (ncells <= 0 || !cell_dofs || !ldof || !nodeperm || !coords)
    return set_err(ctx, HFK_EARG, "HyperFEMKernels: empty mesh or null array");
  if (nfree <= 0 || nfree >= (int64_t)1 << 31 || ndir < 0 || ndir >= (int64_t)1 << 31)
    return set_err(ctx, HFK_EDIM, "HyperFEMKernels: free/Dirichlet dof counts out of range for Int32 ids");
  if (law < 0 || law > 2) return set_err(ctx, HFK_EARG, "HyperFEMKernels: unknown law");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  std::unique_ptr<hfk_op> op(new hfk_op());
  op->ctx = ctx;
  op->Q = Q;
  op->P = P;
  op->nn = Q * Q * Q;
  op->nd = 3 * op->nn;
  op->np = P * P * P;
  op->C = cells_per_batch(P);
  op->ncells = ncells;
  op->nfree = nfree;
  op->ndir = ndir;
  op->law = law;
  op->lam = lam;
  op->mu = mu;
  const int nn = op->nn, nd = op->nd;
  op->cell_dofs.assign(cell_dofs, cell_dofs + (size_t)ncells * nd);
  op->ldof.assign(ldof, ldof + nd);
  op->nodeperm.assign(nodeperm, nodeperm + nn);
  // validate the tables and ids
  std::vector<int> seen(nd, 0), seen_n(nn, 0);
  for (int j = 0; j < nd; ++j) {
    if (ldof[j] < 0 || ldof[j] >= nd || seen[ldof[j]]++) return set_err(ctx, HFK_EARG, "HyperFEMKernels: ldof is not a permutation");
  }
  for (int t = 0; t < nn; ++t) {
    if (nodeperm[t] < 0 || nodeperm[t] >= nn || seen_n[nodeperm[t]]++)
      return set_err(ctx, HFK_EARG, "HyperFEMKernels: nodeperm is not a permutation");
  }
  for (size_t i = 0; i < op->cell_dofs.size(); ++i) {
    const int32_t id = op->cell_dofs[i];
    if (id == 0 || id > nfree || -(int64_t)id > ndir)
      return set_err(ctx, HFK_EDIM, "HyperFEMKernels: cell dof id out of range");
  }
  // 1D tables
  double xq[4], wq[4];
  gauss01(P, xq, wq);
  for (int a = 0; a < P; ++a) {
    lagrange1d(Q, xq[a], &op->tab.N[a * Q], &op->tab.dN[a * Q]);
    op->tab.xq[a] = xq[a];
    op->tab.wq[a] = wq[a];
  }
  int rc = color_cells(op.get());
  if (rc) return rc;
  // batches, colour by colour (cells of a colour in increasing order, as the reference's groups)
  const int ncol = ncells ? *std::max_element(op->color.begin(), op->color.end()) : 0;
  std::vector<std::vector<int64_t>> groups(ncol);
  for (int64_t cell = 0; cell < ncells; ++cell) groups[op->color[cell] - 1].push_back(cell);
  const int C = op->C;
  op->slot_of_cell.assign((size_t)ncells, -1);
  for (int c = 0; c < ncol; ++c) {
    const int64_t nb = ((int64_t)groups[c].size() + C - 1) / C;
    op->color_batches.push_back(BatchRange{(int)op->nbatches, (int)nb});
    for (int64_t b = 0; b < nb; ++b)
      for (int i = 0; i < C; ++i) {
        const int64_t k = b * C + i;
        const int64_t cell = k < (int64_t)groups[c].size() ? groups[c][k] : -1;
        if (cell >= 0) op->slot_of_cell[cell] = (int64_t)op->batch_cell.size();
        op->batch_cell.push_back(cell);
      }
    op->nbatches += nb;
  }
  if (op->nbatches >= (int64_t)1 << 30) return set_err(ctx, HFK_EDIM, "HyperFEMKernels: mesh too large");
  // device arrays in batch order, tensor node order
  const size_t nslot = (size_t)op->nbatches * C;
  std::vector<int32_t> hids(nslot * nd, 0);
  std::vector<double> hx(nslot * 24);
  static const double unit[24] = {0, 0, 0, 1, 0, 0, 0, 1, 0, 1, 1, 0, 0, 0, 1, 1, 0, 1, 0, 1, 1, 1, 1, 1};
  for (size_t s = 0; s < nslot; ++s) {
    const int64_t cell = op->batch_cell[s];
    if (cell < 0) {
      std::memcpy(&hx[s * 24], unit, sizeof(unit));
      continue;
    }
    const int32_t* ids = &op->cell_dofs[(size_t)cell * nd];
    for (int c = 0; c < 3; ++c)
      for (int t = 0; t < nn; ++t) hids[s * nd + c * nn + t] = ids[ldof[c * nn + nodeperm[t]]];
    std::memcpy(&hx[s * 24], &coords[(size_t)cell * 24], 24 * sizeof(double));
  }
  op->loc_of.resize(nd);
  for (int c = 0; c < 3; ++c)
    for (int t = 0; t < nn; ++t) op->loc_of[c * nn + t] = ldof[c * nn + nodeperm[t]];
  // affine cells: the bilinear and trilinear terms of every cell map vanish (x = X0 + A xi)
  op->affine = true;
  for (int64_t cell = 0; cell < ncells && op->affine; ++cell) {
    const double* X = &coords[(size_t)cell * 24];
    double scale = 0.0;
    for (int v = 1; v < 8; ++v)
      for (int i = 0; i < 3; ++i) scale = std::max(scale, std::fabs(X[v * 3 + i] - X[i]));
    for (int i = 0; i < 3 && op->affine; ++i) {
      const double x0 = X[i], x1 = X[3 + i], x2 = X[6 + i], x3 = X[9 + i], x4 = X[12 + i], x5 = X[15 + i],
                   x6 = X[18 + i], x7 = X[21 + i];
      const double t[4] = {x3 - x2 - x1 + x0, x5 - x4 - x1 + x0, x6 - x4 - x2 + x0,
                           x7 - x6 - x5 - x3 + x4 + x2 + x1 - x0};
      for (double v : t)
        if (std::fabs(v) > 1e-12 * scale) op->affine = false;
    }
  }
  HIPCHK(ctx, op->ids.alloc(hids.size() * sizeof(int32_t)));
  HIPCHK(ctx, hipMemcpy(op->ids.p, hids.data(), hids.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  HIPCHK(ctx, op->coords.alloc(hx.size() * sizeof(double)));
  HIPCHK(ctx, hipMemcpy(op->coords.p, hx.data(), hx.size() * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(ctx, op->loc_of_d.alloc(nd * sizeof(int32_t)));
  HIPCHK(ctx, hipMemcpy(op->loc_of_d.p, op->loc_of.data(), nd * sizeof(int32_t), hipMemcpyHostToDevice));
  HIPCHK(ctx, op->dvals.alloc((size_t)std::max<int64_t>(ndir, 1) * sizeof(double)));
  HIPCHK(ctx, hipMemset(op->dvals.p, 0, (size_t)std::max<int64_t>(ndir, 1) * sizeof(double)));
  // single-launch map of the atomic product: first batch and batch count per colour
  {
    std::vector<int> fc;
    int maxc = 0;
    for (const BatchRange& r : op->color_batches) {
      fc.push_back(r.first);
      fc.push_back(r.count);
      maxc = std::max(maxc, r.count);
    }
    HIPCHK(ctx, op->colfc.alloc(fc.size() * sizeof(int)));
    HIPCHK(ctx, hipMemcpy(op->colfc.p, fc.data(), fc.size() * sizeof(int), hipMemcpyHostToDevice));
    op->amap.colfc = op->colfc.as<int>();
    op->amap.ncol = (int)op->color_batches.size();
    op->amap.gx = (maxc + 7) / 8;
    op->amap.grid = 8 * op->amap.gx * op->amap.ncol;
  }
  *out = op.release();
  return HFK_OK;
}

int hfk_op_set_deterministic(hfk_op* op, int deterministic) {
  if (!op) return set_err(nullptr, HFK_EARG, "HyperFEMKernels: null operator");
  op->deterministic = deterministic != 0;
  return HFK_OK;
}

int hfk_op_destroy(hfk_op* op) {
  if (!op) return HFK_OK;
  (void)hipSetDevice(op->ctx->device);
  delete op;
  return HFK_OK;
}

int hfk_op_set_refinement(hfk_op* fine, hfk_op* coarse, const int32_t* parent, const int32_t* child) {
  if (!fine || !fine->ctx) return set_err(nullptr, HFK_EARG, "HyperFEMKernels: null operator");
  if (!coarse) {
    fine->refines = nullptr;
    fine->parent.clear();
    fine->child.clear();
    return HFK_OK;
  }
  if (!parent || !child || coarse->ctx != fine->ctx)
    return set_err(fine->ctx, HFK_EARG, "HyperFEMKernels: invalid refinement arguments");
  std::vector<int64_t> key((size_t)fine->ncells);
  for (int64_t f = 0; f < fine->ncells; ++f) {
    if (parent[f] < 0 || parent[f] >= coarse->ncells || child[f] < 0 || child[f] > 7)
      return set_err(fine->ctx, HFK_EDIM, "HyperFEMKernels: refinement out of range (coarse cell, child 0..7)");
    key[f] = (int64_t)parent[f] * 8 + child[f];
  }
  std::sort(key.begin(), key.end());
  if (std::adjacent_find(key.begin(), key.end()) != key.end())
    return set_err(fine->ctx, HFK_EARG, "HyperFEMKernels: two cells claim the same child of a coarse cell");
  fine->refines = coarse;
  fine->parent.assign(parent, parent + fine->ncells);
  fine->child.assign(child, child + fine->ncells);
  return HFK_OK;
}

int hfk_op_set_dirichlet(hfk_op* op, const double* dvals) {
  int rc = bind(op);
  if (rc) return rc;
  if (op->ndir > 0 && !dvals) return set_err(op->ctx, HFK_EARG, "HyperFEMKernels: null Dirichlet values");
  if (op->ndir > 0)
    HIPCHK(op->ctx, hipMemcpy(op->dvals.p, dvals, (size_t)op->ndir * sizeof(double), hipMemcpyHostToDevice));
  return HFK_OK;
}

int hfk_op_set_parameters(hfk_op* op, double lam, double mu) {
  int rc = bind(op);
  if (rc) return rc;
  if (!std::isfinite(lam) || !std::isfinite(mu)) return set_err(op->ctx, HFK_EARG, "HyperFEMKernels: lam and mu must be finite");
  op->lam = lam;
  op->mu = mu;
  op->storage = -1;  // the stored tangents belong to the previous parameters
  return HFK_OK;
}

int hfk_op_set_robin(hfk_op* op, int64_t nfaces, const int32_t* face_cell, const int32_t* face_local,
                     const double* stiffness) {
  int rc = bind(op);
  if (rc) return rc;
  if (nfaces == 0) {
    op->robin_n = 0;
    op->storage = -1;
    return HFK_OK;
  }
  if ((rc = check_faces(op, nfaces, face_cell, face_local, stiffness, 1))) return rc;
  std::vector<int32_t> order;
  std::vector<int64_t> start, slot;
  face_groups(op, nfaces, face_cell, face_local, order, start, slot);
  hipStream_t s = op->ctx->stream;
  DevBuf k_d;
  HIPCHK(op->ctx, k_d.alloc((size_t)nfaces * sizeof(double)));
  HIPCHK(op->ctx, op->robin_slot.alloc((size_t)nfaces * sizeof(int64_t)));
  HIPCHK(op->ctx, op->robin_local.alloc((size_t)nfaces * sizeof(int32_t)));
  HIPCHK(op->ctx, op->robin_order.alloc((size_t)nfaces * sizeof(int32_t)));
  HIPCHK(op->ctx, op->robin_M.alloc((size_t)nfaces * op->Q * op->Q * op->Q * op->Q * sizeof(double)));
  HIPCHK(op->ctx, hipMemcpyAsync(k_d.p, stiffness, (size_t)nfaces * sizeof(double), hipMemcpyHostToDevice, s));
  HIPCHK(op->ctx, hipMemcpyAsync(op->robin_slot.p, slot.data(), (size_t)nfaces * sizeof(int64_t), hipMemcpyHostToDevice, s));
  HIPCHK(op->ctx, hipMemcpyAsync(op->robin_local.p, face_local, (size_t)nfaces * sizeof(int32_t), hipMemcpyHostToDevice, s));
  HIPCHK(op->ctx, hipMemcpyAsync(op->robin_order.p, order.data(), (size_t)nfaces * sizeof(int32_t), hipMemcpyHostToDevice, s));
  HIPCHK(op->ctx, launch_face_mass(op->dev(), op->robin_slot.as<int64_t>(), op->robin_local.as<int32_t>(),
                                   k_d.as<double>(), nfaces, op->robin_M.as<double>(), s));
  HIPCHK(op->ctx, hipStreamSynchronize(s));
  op->robin_start = start;
  op->robin_n = nfaces;
  op->storage = -1;  // the tangents of a matrix-free Jacobian belong to the operator without the springs
  return HFK_OK;
}

int hfk_op_set_traction(hfk_op* op, int64_t nfaces, const int32_t* face_cell, const int32_t* face_local,
                        const double* traction) {
  int rc = bind(op);
  if (rc) return rc;
  if (nfaces == 0) {
    op->has_traction = false;
    return HFK_OK;
  }
  if ((rc = check_faces(op, nfaces, face_cell, face_local, traction, 3))) return rc;
  std::vector<int32_t> order;
  std::vector<int64_t> start, slot;
  face_groups(op, nfaces, face_cell, face_local, order, start, slot);
  const int ng = (int)start.size() - 1;
  hipStream_t s = op->ctx->stream;
  DevBuf slot_d, local_d, trac_d, order_d;
  HIPCHK(op->ctx, slot_d.alloc((size_t)nfaces * sizeof(int64_t)));
  HIPCHK(op->ctx, local_d.alloc((size_t)nfaces * sizeof(int32_t)));
  HIPCHK(op->ctx, trac_d.alloc((size_t)nfaces * 3 * sizeof(double)));
  HIPCHK(op->ctx, order_d.alloc((size_t)nfaces * sizeof(int32_t)));
  HIPCHK(op->ctx, hipMemcpyAsync(slot_d.p, slot.data(), (size_t)nfaces * sizeof(int64_t), hipMemcpyHostToDevice, s));
  HIPCHK(op->ctx, hipMemcpyAsync(local_d.p, face_local, (size_t)nfaces * sizeof(int32_t), hipMemcpyHostToDevice, s));
  HIPCHK(op->ctx, hipMemcpyAsync(trac_d.p, traction, (size_t)nfaces * 3 * sizeof(double), hipMemcpyHostToDevice, s));
  HIPCHK(op->ctx, hipMemcpyAsync(order_d.p, order.data(), (size_t)nfaces * sizeof(int32_t), hipMemcpyHostToDevice, s));
  const int64_t nf = op->nfree + op->ndir;
  HIPCHK(op->ctx, op->face_load.alloc((size_t)std::max<int64_t>(nf, 1) * sizeof(double)));
  HIPCHK(op->ctx, launch_fill(op->face_load.as<double>(), nf, 0.0, s));
  const OpDev d = op->dev();
  for (int g = 0; g < ng; ++g)
    HIPCHK(op->ctx, launch_face_load(d, slot_d.as<int64_t>(), local_d.as<int32_t>(), trac_d.as<double>(),
                                     order_d.as<int32_t>() + start[g], start[g + 1] - start[g], op->nfree,
                                     op->face_load.as<double>(), s));
  HIPCHK(op->ctx, hipStreamSynchronize(s));
  op->has_traction = true;
  return HFK_OK;
}

int hfk_op_set_load_factor(hfk_op* op, double load_factor) {
  int rc = bind(op);
  if (rc) return rc;
  if (!std::isfinite(load_factor)) return set_err(op->ctx, HFK_EARG, "HyperFEMKernels: load factor not finite");
  op->load_factor = load_factor;
  return HFK_OK;
}

int hfk_op_traction_load(hfk_op* op, double* f) {
  int rc = bind(op);
  if (rc) return rc;
  if (!f) return set_err(op->ctx, HFK_EARG, "HyperFEMKernels: null argument");
  const int64_t nf = op->nfree + op->ndir;
  if (!op->has_traction) {
    std::fill(f, f + nf, 0.0);
    return HFK_OK;
  }
  HIPCHK(op->ctx, hipStreamSynchronize(op->ctx->stream));
  HIPCHK(op->ctx, hipMemcpy(f, op->face_load.p, (size_t)nf * sizeof(double), hipMemcpyDeviceToHost));
  return HFK_OK;
}

int hfk_op_set_mu_field(hfk_op* op, const double* mu_points) {
  int rc = bind(op);
  if (rc) return rc;
  if (!mu_points) {
    op->has_mu_field = false;
    return HFK_OK;
  }
  const size_t nslot = (size_t)op->nbatches * op->C;
  std::vector<double> h(nslot * op->np, op->mu);
  for (size_t s = 0; s < nslot; ++s) {
    const int64_t cell = op->batch_cell[s];
    if (cell >= 0) std::memcpy(&h[s * op->np], &mu_points[(size_t)cell * op->np], op->np * sizeof(double));
  }
  HIPCHK(op->ctx, op->mu_pts.alloc(h.size() * sizeof(double)));
  HIPCHK(op->ctx, hipMemcpy(op->mu_pts.p, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
  op->has_mu_field = true;
  return HFK_OK;
}

int hfk_op_info(hfk_op* op, int64_t* info) {
  if (!op || !info) return set_err(op ? op->ctx : nullptr, HFK_EARG, "HyperFEMKernels: null argument");
  info[0] = op->nfree;
  info[1] = op->ndir;
  info[2] = op->ncells;
  info[3] = op->nd;
  info[4] = op->np;
  info[5] = (int64_t)op->color_batches.size();
  info[6] = op->storage < 0 ? 0 : (int64_t)op->nbatches * store_entries(op->storage) * op->C * op->np * store_bytes(op->storage);
  info[7] = op->pattern_built ? (int64_t)op->rowval.size() : -1;
  return HFK_OK;
}

int hfk_op_colors(hfk_op* op, int32_t* color_of_cell) {
  if (!op || !color_of_cell) return set_err(op ? op->ctx : nullptr, HFK_EARG, "HyperFEMKernels: null argument");
  std::memcpy(color_of_cell, op->color.data(), op->color.size() * sizeof(int32_t));
  return HFK_OK;
}

int hfk_residual(hfk_op* op, const double* x, double* b) {
  int rc = bind(op);
  if (rc) return rc;
  if (!x || !b) return set_err(op->ctx, HFK_EARG, "HyperFEMKernels: null vector");
  if ((rc = stage_in(op, op->x_stage, x, op->nfree))) return rc;
  HIPCHK(op->ctx, op->y_stage.alloc((size_t)op->nfree * sizeof(double)));
  if ((rc = run_residual(op, op->x_stage.as<double>(), op->y_stage.as<double>(), op->ctx->stream))) return rc;
  HIPCHK(op->ctx, hipMemcpyAsync(b, op->y_stage.p, (size_t)op->nfree * sizeof(double), hipMemcpyDeviceToHost, op->ctx->stream));
  return sync(op);
}

int hfk_residual_dev(hfk_op* op, const double* x_dev, double* b_dev, void* stream) {
  int rc = bind(op);
  if (rc) return rc;
  if (!x_dev || !b_dev) return set_err(op->ctx, HFK_EARG, "HyperFEMKernels: null vector");
  return run_residual(op, x_dev, b_dev, op->stream(stream));
}

int hfk_energy(hfk_op* op, const double* x, double* energy) {
  int rc = bind(op);
  if (rc) return rc;
  if (!x || !energy) return set_err(op->ctx, HFK_EARG, "HyperFEMKernels: null argument");
  if ((rc = stage_in(op, op->x_stage, x, op->nfree))) return rc;
  HIPCHK(op->ctx, op->y_stage.alloc((size_t)std::max<int64_t>(op->nfree, 1) * sizeof(double)));
  if ((rc = run_energy(op, op->x_stage.as<double>(), op->y_stage.as<double>(), op->ctx->stream))) return rc;
  HIPCHK(op->ctx, hipMemcpyAsync(energy, op->y_stage.p, sizeof(double), hipMemcpyDeviceToHost, op->ctx->stream));
  return sync(op);
}

int hfk_energy_dev(hfk_op* op, const double* x_dev, double* energy_dev, void* stream) {
  int rc = bind(op);
  if (rc) return rc;
  if (!x_dev || !energy_dev) return set_err(op->ctx, HFK_EARG, "HyperFEMKernels: null argument");
  return run_energy(op, x_dev, energy_dev, op->stream(stream));
}

int hfk_energy_noalloc_dev(hfk_op* op, const double* x_dev, double* energy_dev, void* stream) {
  int rc = bind(op);
  if (rc) return rc;
  if (!x_dev || !energy_dev) return set_err(op->ctx, HFK_EARG, "HyperFEMKernels: null argument");
  if (!op->ctx->scratch) return set_err(op->ctx, HFK_ESTATE, "HyperFEMKernels: the context has no scratch");
  const int64_t need = energy_partials(op);
  if (need > kScratchDoubles)
    return set_err(op->ctx, HFK_EDIM, "HyperFEMKernels: hfk_energy_noalloc_dev needs " + std::to_string(need) +
                                          " partial sums; its scratch holds " + std::to_string(kScratchDoubles) +
                                          " (batches + 512 for a traction load + Robin faces)");
  if (op->rtc && !rtc_law_loaded(op))
    return set_err(op->ctx, HFK_ESTATE, "HyperFEMKernels: the compiled law's kernels are not loaded for this "
                                        "operator yet (they are compiled by its first residual/prepare/energy call)");
  return run_energy_into(op, x_dev, energy_dev, op->ctx->scratch, op->stream(stream));
}

int hfk_mf_prepare(hfk_op* op, const double* x, int storage) {
  int rc = bind(op);
  if (rc) return rc;
  if (!x) return set_err(op->ctx, HFK_EARG, "HyperFEMKernels: null state");
  if ((rc = stage_in(op, op->x_stage, x, op->nfree))) return rc;
  if ((rc = run_prepare(op, op->x_stage.as<double>(), storage, op->ctx->stream))) return rc;
  return sync(op);
}

int hfk_mf_prepare_dev(hfk_op* op, const double* x_dev, int storage, void* stream) {
  int rc = bind(op);
  if (rc) return rc;
  if (!x_dev) return set_err(op->ctx, HFK_EARG, "HyperFEMKernels: null state");
  return run_prepare(op, x_dev, storage, op->stream(stream));
}

int hfk_mf_apply(hfk_op* op, const double* dx, double* y) {
  int rc = bind(op);
  if (rc) return rc;
  if (!dx || !y) return set_err(op->ctx, HFK_EARG, "HyperFEMKernels: null vector");
  if ((rc = stage_in(op, op->x_stage, dx, op->nfree))) return rc;
  HIPCHK(op->ctx, op->y_stage.alloc((size_t)op->nfree * sizeof(double)));
  if ((rc = run_apply(op, op->x_stage.as<double>(), op->y_stage.as<double>(), op->ctx->stream))) return rc;
  HIPCHK(op->ctx, hipMemcpyAsync(y, op->y_stage.p, (size_t)op->nfree * sizeof(double), hipMemcpyDeviceToHost, op->ctx->stream));
  return sync(op);
}

int hfk_mf_apply_dev(hfk_op* op, const double* dx_dev, double* y_dev, void* stream) {
  int rc = bind(op);
  if (rc) return rc;
  if (!dx_dev || !y_dev) return set_err(op->ctx, HFK_EARG, "HyperFEMKernels: null vector");
  return run_apply(op, dx_dev, y_dev, op->stream(stream));
}

int hfk_mf_tangents(hfk_op* op, double* D) {
  int rc = bind(op);
  if (rc) return rc;
  if (!D) return set_err(op->ctx, HFK_EARG, "HyperFEMKernels: null output");
  if (op->storage < 0) return set_err(op->ctx, HFK_ESTATE, "HyperFEMKernels: no tangents prepared");
  const size_t nslot = (size_t)op->nbatches * op->C;
  DevBuf tmp;
  HIPCHK(op->ctx, tmp.alloc(nslot * op->np * 81 * sizeof(double)));
  HIPCHK(op->ctx, launch_expand_tangents(op->dev(), op->storage, op->tangents.p, tmp.as<double>(), op->nbatches, op->ctx->stream));
  std::vector<double> h(nslot * op->np * 81);
  HIPCHK(op->ctx, hipMemcpyAsync(h.data(), tmp.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, op->ctx->stream));
  if ((rc = sync(op))) return rc;
  const size_t per = (size_t)op->np * 81;
  for (int64_t cell = 0; cell < op->ncells; ++cell)
    std::memcpy(&D[(size_t)cell * per], &h[(size_t)op->slot_of_cell[cell] * per], per * sizeof(double));
  return HFK_OK;
}

int hfk_pattern_nnz(hfk_op* op, int64_t* nnz) {
  if (!op || !nnz) return set_err(op ? op->ctx : nullptr, HFK_EARG, "HyperFEMKernels: null argument");
  int rc = build_pattern(op);
  if (rc) return rc;
  *nnz = (int64_t)op->rowval.size();
  return HFK_OK;
}

int hfk_pattern(hfk_op* op, int64_t* colptr, int64_t* rowval, int32_t* nzpos) {
  if (!op) return set_err(nullptr, HFK_EARG, "HyperFEMKernels: null operator");
  int rc = build_pattern(op);
  if (rc) return rc;
  if (colptr) std::memcpy(colptr, op->colptr.data(), op->colptr.size() * sizeof(int64_t));
  if (rowval) std::memcpy(rowval, op->rowval.data(), op->rowval.size() * sizeof(int64_t));
  if (nzpos) std::memcpy(nzpos, op->nzpos.data(), op->nzpos.size() * sizeof(int32_t));
  return HFK_OK;
}

int hfk_jacobian(hfk_op* op, const double* x, double* nzval) {
  int rc = bind(op);
  if (rc) return rc;
  if (!x || !nzval) return set_err(op->ctx, HFK_EARG, "HyperFEMKernels: null argument");
  if ((rc = build_pattern(op))) return rc;
  if ((rc = stage_in(op, op->x_stage, x, op->nfree))) return rc;
  const size_t nnz = op->rowval.size();
  DevBuf nz;
  HIPCHK(op->ctx, nz.alloc(std::max<size_t>(nnz, 1) * sizeof(double)));
  if ((rc = run_jacobian(op, op->x_stage.as<double>(), nz.as<double>(), op->ctx->stream))) return rc;
  HIPCHK(op->ctx, hipMemcpyAsync(nzval, nz.p, nnz * sizeof(double), hipMemcpyDeviceToHost, op->ctx->stream));
  return sync(op);
}

int hfk_jacobian_dev(hfk_op* op, const double* x_dev, double* nzval_dev, void* stream) {
  int rc = bind(op);
  if (rc) return rc;
  if (!x_dev || !nzval_dev) return set_err(op->ctx, HFK_EARG, "HyperFEMKernels: null argument");
  return run_jacobian(op, x_dev, nzval_dev, op->stream(stream));
}

}  // extern "C"
