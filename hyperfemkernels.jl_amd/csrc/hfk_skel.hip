// Face integrals of the hexahedral operator with compiled point forms: skeleton integrals (src/skeleton.jl) on
// interior facets and a boundary form on faces of the cells (the boundary groups of src/assembly.jl:539-578).  The
// face lists and their validation, the colouring, the facet neighbours for the sparsity pattern, and the launches of
// the hiprtc-compiled face kernels (hfk_skel.hpp) in residual!, the matrix-free product, jacobian! and the energy.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "hfk_op.hpp"
#include "hfk_skel.hpp"

namespace hfk {

int64_t skel_facets(const hfk_op* op) { return (op->skel ? op->skel->nf : 0) + (op->bnd ? op->bnd->nf : 0); }

void skel_neighbours(const hfk_op* op, CellNeighbours& nb) {
  nb.start.assign((size_t)op->ncells + 1, 0);
  nb.cells.clear();
  if (!op->skel) return;
  const SkelData& sk = *op->skel;
  for (int64_t f = 0; f < sk.nf; ++f) {
    nb.start[sk.cells[2 * f] + 1]++;
    nb.start[sk.cells[2 * f + 1] + 1]++;
  }
  for (int64_t c = 0; c < op->ncells; ++c) nb.start[c + 1] += nb.start[c];
  std::vector<int64_t> fill(nb.start.begin(), nb.start.end() - 1);
  nb.cells.assign((size_t)nb.start[op->ncells], 0);
  for (int64_t f = 0; f < sk.nf; ++f) {
    const int32_t a = sk.cells[2 * f], b = sk.cells[2 * f + 1];
    nb.cells[fill[a]++] = b;
    nb.cells[fill[b]++] = a;
  }
}

namespace {

// x(xi) of the trilinear map of the 8 vertices X[v][i]
void trilinear_point(const double* X, const double* xi, double* x) {
  x[0] = x[1] = x[2] = 0.0;
  for (int v = 0; v < 8; ++v) {
    double w = 1.0;
    for (int k = 0; k < 3; ++k) w *= ((v >> k) & 1) ? xi[k] : 1.0 - xi[k];
    for (int i = 0; i < 3; ++i) x[i] += w * X[v * 3 + i];
  }
}

void face_point(int lf, double s, double t, double* xi) {
  int d, sd, e1, e2;
  skel_face_dirs(lf, d, sd, e1, e2);
  xi[d] = sd;
  xi[e1] = s;
  xi[e2] = t;
}

void orient_st(int o, double s, double t, double& s2, double& t2) {
  s2 = (o & 1) ? t : s;
  t2 = (o & 1) ? s : t;
  if (o & 2) s2 = 1.0 - s2;
  if (o & 4) t2 = 1.0 - t2;
}

// dof id of component c, tensor node t of a cell (the host's table through ldof and nodeperm)
int32_t tensor_id(const hfk_op* op, int64_t cell, int c, int t) {
  return op->cell_dofs[(size_t)cell * op->nd + op->loc_of[c * op->nn + t]];
}

int face_tnode_host(int Q, int lf, int m1, int m2) {
  int d, sd, e1, e2;
  skel_face_dirs(lf, d, sd, e1, e2);
  int idx[3];
  idx[d] = sd ? Q - 1 : 0;
  idx[e1] = m1;
  idx[e2] = m2;
  return idx[0] + Q * (idx[1] + Q * idx[2]);
}

int launch_skel(hfk_op* op, SkelData& sk, hipFunction_t fn, int per_facet, const double* xf, const double* dx,
                double* out, double alpha, bool jac, hipStream_t s, bool energy = false) {
  OpDev d = op->dev();
  SkelDev sd{};
  sd.slot = sk.slot.as<int64_t>();
  sd.face = sk.face.as<int32_t>();
  sd.colptr = sk.colptr.as<int64_t>();
  sd.rowval = sk.rowval.as<int32_t>();
  std::memcpy(sd.p, sk.p, sizeof(sk.p));
  const int ncol = (int)sk.start.size() - 1;
  for (int c = 0; c < ncol; ++c) {
    const int64_t n = sk.start[c + 1] - sk.start[c];
    if (n == 0) continue;
    sd.order = sk.order.as<int32_t>() + sk.start[c];
    // one energy partial per facet, at the facet's position in the colour order
    double* o = energy ? out + sk.start[c] : out;
    if (jac) {
      for (int sides = 0; sides < sk.ns * sk.ns; ++sides) {  // one launch per (row side, column side) block
        void* args[] = {&d, &sd, &xf, &o, &sides};
        HIPCHK(op->ctx, hipModuleLaunchKernel(fn, (unsigned)(n * per_facet), 1, 1, 256, 1, 1, 0, s, args, nullptr));
      }
    } else {
      void* args[] = {&d, &sd, &xf, &dx, &o, &alpha};
      HIPCHK(op->ctx, hipModuleLaunchKernel(fn, (unsigned)(n * per_facet), 1, 1, 256, 1, 1, 0, s, args, nullptr));
    }
  }
  return HFK_OK;
}

}  // namespace

int skel_residual(hfk_op* op, const double* x_dev, double* b_dev, hipStream_t s) {
  for (SkelData* sk : {op->skel.get(), op->bnd.get()})
    if (sk) {
      int rc = launch_skel(op, *sk, sk->res, 1, x_dev, nullptr, b_dev, 1.0, false, s);
      if (rc) return rc;
    }
  return HFK_OK;
}

int skel_apply(hfk_op* op, const double* dx_dev, double* y_dev, double alpha, hipStream_t s) {
  if (skel_facets(op) == 0) return HFK_OK;
  if (!op->state.p) return set_err(op->ctx, HFK_ESTATE, "HyperFEMKernels: mul! before prepare! (no state)");
  for (SkelData* sk : {op->skel.get(), op->bnd.get()})
    if (sk) {
      int rc = launch_skel(op, *sk, sk->jv, 1, op->state.as<double>(), dx_dev, y_dev, alpha, false, s);
      if (rc) return rc;
    }
  return HFK_OK;
}

int skel_energy(hfk_op* op, const double* x_dev, double* part, hipStream_t s) {
  for (SkelData* sk : {op->skel.get(), op->bnd.get()})
    if (sk) {
      int rc = launch_skel(op, *sk, sk->energy, 1, x_dev, nullptr, part, 1.0, false, s, true);
      if (rc) return rc;
      part += sk->nf;
    }
  return HFK_OK;
}

namespace {
int face_assemble(hfk_op* op, SkelData& sk, const double* x_dev, double* nz_dev, hipStream_t s) {
  if (!sk.colptr.p) {  // the pattern (built by the caller) in 0-based form on the device
    std::vector<int64_t> cp(op->colptr.size());
    std::vector<int32_t> rv(op->rowval.size());
    for (size_t i = 0; i < cp.size(); ++i) cp[i] = op->colptr[i] - 1;
    for (size_t i = 0; i < rv.size(); ++i) rv[i] = (int32_t)(op->rowval[i] - 1);
    HIPCHK(op->ctx, sk.colptr.alloc(cp.size() * sizeof(int64_t)));
    HIPCHK(op->ctx, sk.rowval.alloc(std::max<size_t>(rv.size(), 1) * sizeof(int32_t)));
    HIPCHK(op->ctx, hipMemcpy(sk.colptr.p, cp.data(), cp.size() * sizeof(int64_t), hipMemcpyHostToDevice));
    HIPCHK(op->ctx, hipMemcpy(sk.rowval.p, rv.data(), rv.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  }
  return launch_skel(op, sk, sk.jac, sk.split, x_dev, nullptr, nz_dev, 1.0, true, s);
}
}  // namespace

int skel_assemble(hfk_op* op, const double* x_dev, double* nz_dev, hipStream_t s) {
  for (SkelData* sk : {op->skel.get(), op->bnd.get()})
    if (sk) {
      int rc = face_assemble(op, *sk, x_dev, nz_dev, s);
      if (rc) return rc;
    }
  return HFK_OK;
}

// greedy first fit over the free dofs of the faces' cells (color_facets, src/skeleton.jl:39-69), 128-bit masks; the
// faces in colour order (order, start); false when more than 128 colours are needed
bool color_faces(const hfk_op* op, const std::vector<int32_t>& cells, int ns, std::vector<int32_t>& order,
                 std::vector<int64_t>& start) {
  const int64_t nf = (int64_t)cells.size() / ns;
  std::vector<unsigned __int128> masks((size_t)op->nfree, 0);
  std::vector<int> color((size_t)nf, 0);
  int ncol = 0;
  const int nd = op->nd;
  for (int64_t f = 0; f < nf; ++f) {
    unsigned __int128 busy = 0;
    for (int side = 0; side < ns; ++side)
      for (int j = 0; j < nd; ++j) {
        const int32_t id = op->cell_dofs[(size_t)cells[ns * f + side] * nd + j];
        if (id > 0) busy |= masks[id - 1];
      }
    int c = 0;
    while (c < 128 && ((busy >> c) & 1)) ++c;
    if (c >= 128) return false;
    color[f] = c;
    ncol = std::max(ncol, c + 1);
    const unsigned __int128 bit = (unsigned __int128)1 << c;
    for (int side = 0; side < ns; ++side)
      for (int j = 0; j < nd; ++j) {
        const int32_t id = op->cell_dofs[(size_t)cells[ns * f + side] * nd + j];
        if (id > 0) masks[id - 1] |= bit;
      }
  }
  start.assign((size_t)ncol + 1, 0);
  for (int64_t f = 0; f < nf; ++f) start[color[f] + 1]++;
  for (int c = 0; c < ncol; ++c) start[c + 1] += start[c];
  std::vector<int64_t> fillp(start.begin(), start.end() - 1);
  order.assign((size_t)nf, 0);
  for (int64_t f = 0; f < nf; ++f) order[fillp[color[f]]++] = (int32_t)f;
  return true;
}

// upload the face arrays and compile the form: sk->cells and sk->nf set by the caller
int face_setup(hfk_op* op, SkelData* sk, const int32_t* plus_face, const int32_t* minus_face, const int32_t* orient,
               const char* source, const char* name, const double* params, int nparams, const char* what) {
  const int64_t nf = sk->nf;
  if (params) std::memcpy(sk->p, params, (size_t)nparams * sizeof(double));
  std::vector<int32_t> order;
  if (!color_faces(op, sk->cells, sk->ns, order, sk->start))
    return set_err(op->ctx, HFK_EDIM, std::string("HyperFEMKernels: more than 128 colours needed for the ") + what);
  std::vector<int64_t> slot((size_t)2 * nf);
  std::vector<int32_t> face((size_t)3 * nf);
  for (int64_t f = 0; f < nf; ++f) {
    slot[2 * f] = op->slot_of_cell[sk->cells[sk->ns * f]];
    slot[2 * f + 1] = sk->ns == 2 ? op->slot_of_cell[sk->cells[2 * f + 1]] : slot[2 * f];
    face[3 * f] = plus_face[f];
    face[3 * f + 1] = minus_face ? minus_face[f] : plus_face[f];
    face[3 * f + 2] = orient ? orient[f] : 0;
  }
  HIPCHK(op->ctx, sk->slot.alloc(slot.size() * sizeof(int64_t)));
  HIPCHK(op->ctx, sk->face.alloc(face.size() * sizeof(int32_t)));
  HIPCHK(op->ctx, sk->order.alloc(order.size() * sizeof(int32_t)));
  HIPCHK(op->ctx, hipMemcpy(sk->slot.p, slot.data(), slot.size() * sizeof(int64_t), hipMemcpyHostToDevice));
  HIPCHK(op->ctx, hipMemcpy(sk->face.p, face.data(), face.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  HIPCHK(op->ctx, hipMemcpy(sk->order.p, order.data(), order.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  int rc = rtc_skel(op->ctx, op->Q, op->P, sk->ns, source, name, sk);
  if (rc) return rc;
  sk->split = skel_jac_split(op->Q);
  return HFK_OK;
}

}  // namespace hfk

using namespace hfk;

extern "C" {

int hfk_op_set_skeleton(hfk_op* op, int64_t nfacets, const int32_t* plus_cell, const int32_t* plus_face,
                        const int32_t* minus_cell, const int32_t* minus_face, const int32_t* orient,
                        const char* source, const char* name, const double* params, int nparams) {
  if (!op) return set_err(nullptr, HFK_EARG, "HyperFEMKernels: null operator");
  HIPCHK(op->ctx, hipSetDevice(op->ctx->device));
  if (nfacets < 0 || nparams < 0 || nparams > 8 || (nparams > 0 && !params))
    return set_err(op->ctx, HFK_EARG, "HyperFEMKernels: skeleton: invalid facet count or parameters (at most 8)");
  // the pattern gains or loses the cross blocks: it, its positions and their device copies are rebuilt at the next use
  op->pattern_built = false;
  op->nzpos_d.reset();
  op->storage = -1;  // the tangents of a matrix-free Jacobian belong to the operator without the facets
  if (nfacets == 0) {
    op->skel.reset();
    return HFK_OK;
  }
  if (!plus_cell || !plus_face || !minus_cell || !minus_face || !orient)
    return set_err(op->ctx, HFK_EARG, "HyperFEMKernels: skeleton: null facet array");
  const int Q = op->Q;
  // the cells' vertices (batch order on the device)
  std::vector<double> coords_h((size_t)op->nslots() * 24);
  HIPCHK(op->ctx, hipMemcpy(coords_h.data(), op->coords.p, coords_h.size() * sizeof(double), hipMemcpyDeviceToHost));
  std::vector<int64_t> used((size_t)2 * nfacets);
  for (int64_t f = 0; f < nfacets; ++f) {
    const int32_t cp = plus_cell[f], cm = minus_cell[f];
    if (cp < 0 || cp >= op->ncells || cm < 0 || cm >= op->ncells || cp == cm || plus_face[f] < 0 || plus_face[f] > 5 ||
        minus_face[f] < 0 || minus_face[f] > 5 || orient[f] < 0 || orient[f] > 7)
      return set_err(op->ctx, HFK_EDIM, "HyperFEMKernels: skeleton: facet " + std::to_string(f) +
                                            " out of range (two different cells, local faces 0..5, orientation 0..7)");
    used[2 * f] = (int64_t)cp * 6 + plus_face[f];
    used[2 * f + 1] = (int64_t)cm * 6 + minus_face[f];
    // the two faces are one: the corners coincide, and so do the dof ids of every face node, under the orientation
    const double* Xp = &coords_h[(size_t)op->slot_of_cell[cp] * 24];
    const double* Xm = &coords_h[(size_t)op->slot_of_cell[cm] * 24];
    double scale = 0.0;
    for (int k = 0; k < 24; ++k) scale = std::max(scale, std::fabs(Xp[k]));
    for (int corner = 0; corner < 4; ++corner) {
      const double s = corner & 1, t = corner >> 1;
      double s2, t2, xip[3], xim[3], xp[3], xm[3];
      orient_st(orient[f], s, t, s2, t2);
      face_point(plus_face[f], s, t, xip);
      face_point(minus_face[f], s2, t2, xim);
      trilinear_point(Xp, xip, xp);
      trilinear_point(Xm, xim, xm);
      for (int i = 0; i < 3; ++i)
        if (std::fabs(xp[i] - xm[i]) > 1e-10 * std::max(scale, 1.0))
          return set_err(op->ctx, HFK_EARG, "HyperFEMKernels: skeleton: the faces of facet " + std::to_string(f) +
                                                " do not coincide under its orientation");
    }
    for (int m2 = 0; m2 < Q; ++m2)
      for (int m1 = 0; m1 < Q; ++m1) {
        // the orientation on node indices (equispaced nodes: 1 - s becomes Q - 1 - m)
        int n1 = (orient[f] & 1) ? m2 : m1, n2 = (orient[f] & 1) ? m1 : m2;
        if (orient[f] & 2) n1 = Q - 1 - n1;
        if (orient[f] & 4) n2 = Q - 1 - n2;
        const int tp = face_tnode_host(Q, plus_face[f], m1, m2), tm = face_tnode_host(Q, minus_face[f], n1, n2);
        for (int c = 0; c < 3; ++c)
          if (tensor_id(op, cp, c, tp) != tensor_id(op, cm, c, tm))
            return set_err(op->ctx, HFK_EARG, "HyperFEMKernels: skeleton: the face nodes of facet " + std::to_string(f) +
                                                  " carry different dofs on its two cells");
      }
  }
  std::sort(used.begin(), used.end());
  if (std::adjacent_find(used.begin(), used.end()) != used.end())
    return set_err(op->ctx, HFK_EARG, "HyperFEMKernels: skeleton: a (cell, local face) is used by two facets");
  std::unique_ptr<SkelData> sk(new SkelData());
  sk->ns = 2;
  sk->nf = nfacets;
  sk->cells.resize((size_t)2 * nfacets);
  for (int64_t f = 0; f < nfacets; ++f) {
    sk->cells[2 * f] = plus_cell[f];
    sk->cells[2 * f + 1] = minus_cell[f];
  }
  int rc = face_setup(op, sk.get(), plus_face, minus_face, orient, source, name, params, nparams, "skeleton facets");
  if (rc) return rc;
  op->skel = std::move(sk);
  return HFK_OK;
}

int hfk_op_set_boundary_form(hfk_op* op, int64_t nfaces, const int32_t* face_cell, const int32_t* face_local,
                             const char* source, const char* name, const double* params, int nparams) {
  if (!op) return set_err(nullptr, HFK_EARG, "HyperFEMKernels: null operator");
  HIPCHK(op->ctx, hipSetDevice(op->ctx->device));
  if (nfaces < 0 || nparams < 0 || nparams > 8 || (nparams > 0 && !params))
    return set_err(op->ctx, HFK_EARG, "HyperFEMKernels: boundary form: invalid face count or parameters (at most 8)");
  op->storage = -1;  // the tangents of a matrix-free Jacobian belong to the operator without the boundary form
  if (nfaces == 0) {
    op->bnd.reset();
    return HFK_OK;
  }
  if (!face_cell || !face_local) return set_err(op->ctx, HFK_EARG, "HyperFEMKernels: boundary form: null face array");
  std::vector<int64_t> used((size_t)nfaces);
  for (int64_t f = 0; f < nfaces; ++f) {
    if (face_cell[f] < 0 || face_cell[f] >= op->ncells || face_local[f] < 0 || face_local[f] > 5)
      return set_err(op->ctx, HFK_EDIM, "HyperFEMKernels: boundary form: face out of range (cell, local face 0..5)");
    used[f] = (int64_t)face_cell[f] * 6 + face_local[f];
  }
  std::sort(used.begin(), used.end());
  if (std::adjacent_find(used.begin(), used.end()) != used.end())
    return set_err(op->ctx, HFK_EARG, "HyperFEMKernels: boundary form: a (cell, local face) pair is listed twice");
  std::unique_ptr<SkelData> sk(new SkelData());
  sk->ns = 1;
  sk->nf = nfaces;
  sk->cells.assign(face_cell, face_cell + nfaces);
  int rc = face_setup(op, sk.get(), face_local, nullptr, nullptr, source, name, params, nparams, "boundary faces");
  if (rc) return rc;
  op->bnd = std::move(sk);
  return HFK_OK;
}

int hfk_op_skeleton_colors(hfk_op* op, int32_t* color) {
  if (!op || !color) return set_err(op ? op->ctx : nullptr, HFK_EARG, "HyperFEMKernels: null argument");
  if (skel_facets(op) == 0) return set_err(op->ctx, HFK_ESTATE, "HyperFEMKernels: the operator has no skeleton");
  std::vector<int32_t> order((size_t)op->skel->nf);
  HIPCHK(op->ctx, hipMemcpy(order.data(), op->skel->order.p, order.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
  const int ncol = (int)op->skel->start.size() - 1;
  for (int c = 0; c < ncol; ++c)
    for (int64_t k = op->skel->start[c]; k < op->skel->start[c + 1]; ++k) color[order[k]] = c + 1;
  return HFK_OK;
}

}  // extern "C"
