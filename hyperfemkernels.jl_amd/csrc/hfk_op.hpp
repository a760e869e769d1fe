// Host-side handles of libhfk (context, operator) shared by the C-ABI sources.  Not part of the C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "../../include/hfk.h"
#include "hfk_internal.hpp"

struct hfk_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  double* scratch = nullptr;  // device address of the library's static scratch on this device (set by hfk_open)
};

namespace hfk {

int set_err(hfk_ctx* ctx, int code, const std::string& msg);

#define HIPCHK(ctx, expr)                                                                          \
  do {                                                                                             \
    hipError_t e__ = (expr);                                                                       \
    if (e__ != hipSuccess)                                                                         \
      return ::hfk::set_err((ctx), HFK_EHIP, std::string("HyperFEMKernels: HIP error in ") + #expr + ": " + \
                                                 hipGetErrorString(e__));                          \
  } while (0)

// A device allocation owned by a handle.  alloc() keeps the buffer when it is already large enough.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  hipError_t alloc(size_t n) {
    if (p && bytes >= n) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    if (n == 0) return hipSuccess;
    hipError_t e = hipMalloc(&p, n);
    if (e == hipSuccess) bytes = n;
    return e;
  }
  template <class T>
  T* as() const {
    return static_cast<T*>(p);
  }
  void reset() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
};

// Nodal blocks of one operator (FieldBlocks, src/preconditioners.jl:120-171): one block per node with a free
// component, numbered in first-seen order over (cell, table node).
struct NodeBlocks {
  bool built = false;
  int64_t nblocks = 0;
  std::vector<int32_t> dofs_h;  // 3*nblocks, free ids (1-based), 0 = fixed component
  DevBuf block_of;              // [slot][t] tensor order, -1: every component fixed
  DevBuf dofs, B, Binv;
  // the inverses rounded to FP32 when the blocks come from FP32 tangents: the smoother's block steps read them
  DevBuf Binv32;
  bool fp32 = false;
  DevBuf bad;                   // int counter of non-finite inverted blocks
};

// Face integrals of a hexahedral operator with a compiled point form: interior facets (ns = 2, skeleton integrals,
// src/skeleton.jl: a plus and a minus cell per facet) or boundary faces (ns = 1, one cell per face); the faces by
// colour (two faces of one colour share no free dof of their cells), and the face kernels of the form, compiled with
// hiprtc for the operator's (Q, P).
struct SkelData {
  int ns = 2;
  int64_t nf = 0;
  std::vector<int32_t> cells;   // [f][2]
  std::vector<int64_t> start;   // colour c: order[start[c] .. start[c+1])
  DevBuf slot, face, order;     // [f][2] int64 batch slots, [f][3] int32 faces and orientation, int32 facet order
  DevBuf colptr, rowval;        // the CSC pattern (0-based) for the Jacobian's positions
  double p[8] = {};
  hipModule_t mod = nullptr;
  hipFunction_t res = nullptr, jv = nullptr, energy = nullptr, jac = nullptr;
  int split = 1;                // work-groups per facet and side block of the Jacobian kernel
  SkelData() = default;
  SkelData(const SkelData&) = delete;
  SkelData& operator=(const SkelData&) = delete;
  ~SkelData() {
    if (mod) (void)hipModuleUnload(mod);
  }
};

}  // namespace hfk

struct hfk_op {
  hfk_ctx* ctx = nullptr;
  int Q = 0, P = 0, nn = 0, nd = 0, np = 0, C = 0;
  int64_t ncells = 0, nfree = 0, ndir = 0, nbatches = 0;
  int law = 0;
  double lam = 0, mu = 0;
  hfk::Tab tab{};
  std::vector<int32_t> cell_dofs;               // Gridap order, ncells*nd
  std::vector<int32_t> ldof, nodeperm;          // as passed
  // a nested refinement of another operator's mesh (h-multigrid levels): the coarse cell of every cell and the
  // position of the cell in it (cx + 2 cy + 4 cz, each 0 or 1)
  const hfk_op* refines = nullptr;
  std::vector<int32_t> parent, child;
  std::vector<int32_t> color;                   // 1-based colour of each cell
  std::vector<hfk::BatchRange> color_batches;   // batches of each colour
  std::vector<int64_t> batch_cell;              // nbatches*C, -1 padding
  std::vector<int64_t> slot_of_cell;            // cell -> batch*C + position
  std::vector<int32_t> loc_of;                  // tensor index c*nn + t -> table-local dof
  hfk::DevBuf ids, coords, mu_pts, dvals, tangents, loc_of_d, nzpos_d, x_stage, y_stage, D64, state;
  // every cell map affine (parallelepipeds), and the per-cell geometry of the affine factorised storage
  bool affine = false;
  hfk::DevBuf cellgeo;
  hfk::DevBuf part;  // per-batch partial results of the functionals
  // boundary traction (Neumann load): its load vector over the free then the Dirichlet ids, scaled by load_factor in
  // residual! and the energy
  bool has_traction = false;
  double load_factor = 1.0;
  hfk::DevBuf face_load;
  // Robin (spring) faces: face mass matrices (k included), launched in groups of faces of different cells of one
  // colour: robin_order[robin_start[g] .. robin_start[g+1])
  int64_t robin_n = 0;
  std::vector<int64_t> robin_start;
  hfk::DevBuf robin_slot, robin_local, robin_order, robin_M;
  // a second copy of the tangents at the same state in another storage, for a smoother of another precision
  hfk::DevBuf tangents_s;
  bool has_mu_field = false;
  // products: colour by colour with plain adds (the reference's summation order, bitwise reproducible), or one
  // launch with atomic adds (amap)
  bool deterministic = true;
  hfk::DevBuf colfc;
  hfk::ApplyMap amap{};
  int storage = -1;    // storage of the prepared tangents, -1: not prepared
  int storage_s = -1;  // storage of tangents_s
  bool pattern_built = false;
  std::vector<int64_t> colptr, rowval;
  std::vector<int32_t> nzpos;  // Gridap cell order [cell][ii + nd*jj]
  hfk::NodeBlocks blocks;
  // a pointwise law compiled at run time (nullptr: the built-in law)
  hfk_law* rtc = nullptr;
  // skeleton integrals and a boundary form (nullptr: none)
  std::unique_ptr<hfk::SkelData> skel, bnd;

  hfk::OpDev dev() const {
    hfk::OpDev d;
    d.Q = Q;
    d.P = P;
    d.C = C;
    d.nn = nn;
    d.np = np;
    d.nd = nd;
    d.law = law;
    d.lam = lam;
    d.mu = mu;
    d.tab = tab;
    d.ids = ids.as<int32_t>();
    d.coords = coords.as<double>();
    d.mu_pts = has_mu_field ? mu_pts.as<double>() : nullptr;
    d.dvals = dvals.as<double>();
    d.loc_of = loc_of_d.as<int32_t>();
    d.nzpos = nzpos_d.as<int32_t>();
    d.cellgeo = cellgeo.as<double>();
    return d;
  }
  hipStream_t stream(void* s) const { return s ? static_cast<hipStream_t>(s) : ctx->stream; }
  int64_t nslots() const { return nbatches * C; }
};

namespace hfk {
// shared helpers of the C-ABI sources (hfk_capi.hip)
int op_bind(hfk_op* op);
int op_run_prepare(hfk_op* op, const double* x_dev, int storage, hipStream_t s);
int op_run_apply(hfk_op* op, const double* dx_dev, double* y_dev, hipStream_t s, double alpha, bool accumulate);
// the same with explicit tangents (op->tangents_s for a smoother of another precision)
int op_run_apply_t(hfk_op* op, const void* T, int storage, const double* dx_dev, double* y_dev, hipStream_t s,
                   double alpha, bool accumulate);
int op_run_jacobian(hfk_op* op, const double* x_dev, double* nz_dev, hipStream_t s);
int op_build_pattern(hfk_op* op);
// Field blocks of a multi-field cell (block_spec / cell_block_mask, src/assembly.jl:276-323, 825-832): the field of
// each local dof and, per cell, which (test field, trial field) blocks carry a term (mask[cell * nf^2 + fi + nf fj]).
struct BlockInfo {
  const int8_t* field_of = nullptr;  // nd entries
  int nf = 1;
  const uint8_t* mask = nullptr;     // ncells x nf^2; null: every block of every cell
};
// The reference's CSC sparsity and cell positions of a cell dof table (> 0 free ids), built on the device
// (hfk_pattern.hip): the keys of the free pairs of every cell (full_keys!), or of the blocks set on each cell
// (block_keys!); nzpos of a pair outside the cell's blocks is its stored position when another cell stores it, 0
// otherwise (cell_positions!, src/assembly.jl:864-899).  With `nbr`, the rows of every column also hold the free
// dofs of the facet neighbours of the column's cells (the cross blocks of skeleton_keys!, src/skeleton.jl:355-390);
// nzpos stays the cells' own.  cell_dofs, blocks and nbr are host arrays.
struct CellNeighbours;
struct DevPattern {
  int64_t nnz = 0;
  DevBuf colptr;  // int64, 0-based, ncols + 1
  DevBuf rowval;  // int32, 0-based, nnz
  DevBuf nzpos;   // int32 [cell][ii + nd jj]: 1 + stored position, 0 = not stored
};
int device_pattern(hfk_ctx* ctx, hipStream_t s, const int32_t* cell_dofs, int nd, int64_t ncells, int64_t ncols,
                   const BlockInfo& blocks, const CellNeighbours* nbr, DevPattern& out);
// copies in the reference's form: 1-based Int64 colptr and rowval, nzpos (in the local order perm[ii] when perm is
// given: out[cell][perm[ii] + nd perm[jj]]); null outputs are skipped
int device_pattern_to_host(hfk_ctx* ctx, hipStream_t s, const DevPattern& pt, int64_t ncols, int64_t ncells, int nd,
                           int64_t* colptr, int64_t* rowval, int32_t* nzpos, const int32_t* perm);
// CSR of a square CSC pattern given in the reference's 1-based form (host arrays): rp (int64, 0-based, n + 1), the
// columns of every row ascending (int32, 0-based) and perm[k] = the CSC position of CSR entry k; built on the device
// by a radix sort of (row, column) keys.  *bad: 0, or 1 (a row out of range) | 2 (rows not ascending in a column).
int device_csc_to_csr(hfk_ctx* ctx, hipStream_t s, int64_t n, const int64_t* colptr, const int64_t* rowval,
                      DevBuf& rp, DevBuf& col, DevBuf& perm, int* bad);
// the reference's greedy colouring (1-based colours); false when more than 128 colours are needed
bool greedy_colors(const int32_t* cell_dofs, int nd, int64_t ncells, int64_t nfree, std::vector<int32_t>& color);
int op_build_blocks(hfk_op* op);
// tangents of the operator's law at state x into T in `storage` (built-in kernel or the run-time law)
int op_prepare_into(hfk_op* op, int storage, void* T, const double* x_dev, hipStream_t s);
// the run-time law's prepare! and residual! (hfk_rtc.hip)
int rtc_prepare(hfk_op* op, int storage, void* T, const double* x_dev, hipStream_t s);
int rtc_residual(hfk_op* op, const double* x_dev, double* b_dev, hipStream_t s);
// the Robin term: y += alpha M u on the free rows (u: free values uf, Dirichlet values ud or zero); the node-block
// diagonals; the assembled values
int robin_apply(hfk_op* op, const double* uf, const double* ud, double* y, double alpha, hipStream_t s);
int robin_blocks(hfk_op* op, const int32_t* block_of, double* B, hipStream_t s);
int robin_assemble(hfk_op* op, double* nzval, hipStream_t s);
// strain energy per batch with the operator's compiled law into part[nbatches] (HFK_EARG if it has no energy())
int rtc_energy(hfk_op* op, const double* x_dev, double* part, hipStream_t s);
// true when the operator's compiled law already has its kernels loaded for the operator's (Q, P)
bool rtc_law_loaded(hfk_op* op);
// The kernels of an Oseen point form compiled with hiprtc (k_oseen_form_jac / k_oseen_form_res for the form struct
// `name` of `source`, colour-ordered [0] and atomic [1]), loaded on the current device; cached by source text.
struct OseenFormKernels {
  hipModule_t mod = nullptr;
  hipFunction_t jac[2] = {}, res[2] = {};
};
int rtc_oseen_form(hfk_ctx* ctx, const char* source, const char* name, OseenFormKernels* out);
// The face kernels of a point form with ns sides (source null: the built-in gradient-jump penalty for ns = 2, the
// follower pressure for ns = 1) for (Q, P), loaded on the current device into sk (mod, res, jv, energy, jac).
int rtc_skel(hfk_ctx* ctx, int Q, int P, int ns, const char* source, const char* name, SkelData* sk);
// the face terms' launches (hfk_skel.hip), skeleton then boundary form: b += r; y += alpha J dx at the prepared
// state; the CSC values; one energy partial per face into part; skel_facets counts the faces of both (the energy's
// partials); the facet neighbours of every cell (for the pattern)
int skel_residual(hfk_op* op, const double* x_dev, double* b_dev, hipStream_t s);
int skel_apply(hfk_op* op, const double* dx_dev, double* y_dev, double alpha, hipStream_t s);
int skel_assemble(hfk_op* op, const double* x_dev, double* nz_dev, hipStream_t s);
int skel_energy(hfk_op* op, const double* x_dev, double* part, hipStream_t s);
int64_t skel_facets(const hfk_op* op);
struct CellNeighbours {
  std::vector<int64_t> start;  // cell k: cells[start[k] .. start[k+1])
  std::vector<int32_t> cells;
};
void skel_neighbours(const hfk_op* op, CellNeighbours& nb);
}  // namespace hfk
