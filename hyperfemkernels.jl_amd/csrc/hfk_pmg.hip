// Host orchestration of the solver layer (C ABI): BlockJacobi, the Chebyshev smoother with its Arnoldi
// interval, the p-multigrid transfers, the V-cycle with an assembled coarse level, and CG / flexible CG.
// Every vector lives on the device; the host only reads back the scalars a decision needs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "hfk_krylov.hpp"
#include "hfk_net.hpp"
#include "hfk_op.hpp"
#include "hfk_solver.hpp"

using namespace hfk;

namespace {

// ---------------------------------------------------------------------------------------------------------
// Eigenvalues (real parts) of a small upper Hessenberg matrix H (n x n, column-major h[i + n j]): shifted QR
// iterations with Givens rotations on the active window, Wilkinson shift, deflation of 1x1 and 2x2 blocks
// (a 2x2 block contributes the real part of its eigenvalue pair).  n <= 64.
void hessenberg_eig_real(std::vector<double> h, int n, std::vector<double>& out) {
  out.clear();
  auto H = [&](int i, int j) -> double& { return h[i + n * j]; };
  int hi = n - 1, its = 0;
  while (hi >= 0) {
    if (hi == 0) {
      out.push_back(H(0, 0));
      hi = -1;
      break;
    }
    int l = hi;
    while (l > 0) {
      const double s = std::fabs(H(l - 1, l - 1)) + std::fabs(H(l, l));
      if (std::fabs(H(l, l - 1)) <= 1e-15 * (s == 0 ? 1.0 : s)) {
        H(l, l - 1) = 0.0;
        break;
      }
      --l;
    }
    if (l == hi) {
      out.push_back(H(hi, hi));
      --hi;
      its = 0;
      continue;
    }
    if (l == hi - 1 || its > 200) {
      const double a = H(hi - 1, hi - 1), b = H(hi - 1, hi), c = H(hi, hi - 1), d = H(hi, hi);
      const double tr = 0.5 * (a + d), disc = 0.25 * (a - d) * (a - d) + b * c;
      if (disc >= 0) {
        out.push_back(tr + std::sqrt(disc));
        out.push_back(tr - std::sqrt(disc));
      } else {
        out.push_back(tr);
        out.push_back(tr);
      }
      hi -= 2;
      its = 0;
      continue;
    }
    // Wilkinson shift from the trailing 2x2 of the window (its real part when the pair is complex)
    const double a = H(hi - 1, hi - 1), b = H(hi - 1, hi), c = H(hi, hi - 1), d = H(hi, hi);
    const double tr = 0.5 * (a + d), disc = 0.25 * (a - d) * (a - d) + b * c;
    double mu = tr;
    if (disc >= 0) {
      const double e1 = tr + std::sqrt(disc), e2 = tr - std::sqrt(disc);
      mu = std::fabs(e1 - d) < std::fabs(e2 - d) ? e1 : e2;
    }
    if (its == 20 || its == 60) mu += 0.75 * std::fabs(H(hi, hi - 1));  // exceptional shift
    // QR step on the window [l, hi]: H - mu I = QR, H <- RQ + mu I (rotations applied to the window only)
    for (int i = l; i <= hi; ++i) H(i, i) -= mu;
    std::vector<double> cs(hi - l), sn(hi - l);
    for (int k = l; k < hi; ++k) {
      const double x = H(k, k), y = H(k + 1, k);
      const double r = std::hypot(x, y);
      const double cc = r == 0 ? 1.0 : x / r, ss = r == 0 ? 0.0 : y / r;
      cs[k - l] = cc;
      sn[k - l] = ss;
      for (int j = k; j <= hi; ++j) {
        const double t1 = H(k, j), t2 = H(k + 1, j);
        H(k, j) = cc * t1 + ss * t2;
        H(k + 1, j) = -ss * t1 + cc * t2;
      }
    }
    for (int k = l; k < hi; ++k) {
      const double cc = cs[k - l], ss = sn[k - l];
      for (int i = l; i <= std::min(k + 2, hi); ++i) {
        const double t1 = H(i, k), t2 = H(i, k + 1);
        H(i, k) = cc * t1 + ss * t2;
        H(i, k + 1) = -ss * t1 + cc * t2;
      }
    }
    for (int i = l; i <= hi; ++i) H(i, i) += mu;
    ++its;
  }
}

// chebyshev_start (src/smoothers.jl:68-71): v_i = sin(1 + 0.7 i) + 0.3 cos(2.3 i), i = 1..n, normalised
std::vector<double> chebyshev_start(int64_t n) {
  std::vector<double> v((size_t)n);
  double s = 0;
  for (int64_t i = 0; i < n; ++i) {
    const double k = (double)(i + 1);
    v[i] = std::sin(1.0 + 0.7 * k) + 0.3 * std::cos(2.3 * k);
    s += v[i] * v[i];
  }
  s = std::sqrt(s);
  for (auto& x : v) x /= s;
  return v;
}

// ---------------------------------------------------------------------------------------------------------
// BlockJacobi on the stored tangents of op (block_diagonal!, src/preconditioners.jl:186-218)
int blocks_update(hfk_op* op, const void* T, int storage, hipStream_t s, bool* ok) {
  if (skel_facets(op) > 0)  // src/preconditioners.jl:141 (and the p-multigrid's smoother built on it)
    return set_err(op->ctx, HFK_ESTATE,
                   "HyperFEMKernels: BlockJacobi is not available with skeleton integrals or a compiled boundary form");
  int rc = op_build_blocks(op);
  if (rc) return rc;
  if (storage < 0 || !T) return set_err(op->ctx, HFK_ESTATE, "HyperFEMKernels: BlockJacobi before prepare!");
  NodeBlocks& nb = op->blocks;
  HIPCHK(op->ctx, launch_fill(nb.B.as<double>(), nb.nblocks * 9, 0.0, s));
  const OpDev d = op->dev();
  for (const BatchRange& r : op->color_batches)
    HIPCHK(op->ctx, launch_blocks(d, storage, T, nb.block_of.as<int32_t>(), nb.B.as<double>(), r, s));
  if ((rc = robin_blocks(op, nb.block_of.as<int32_t>(), nb.B.as<double>(), s))) return rc;  // the Robin faces
  HIPCHK(op->ctx, hipMemsetAsync(nb.bad.p, 0, sizeof(int), s));
  HIPCHK(op->ctx, blocks_invert(nb.nblocks, nb.dofs.as<int32_t>(), nb.B.as<double>(), nb.Binv.as<double>(), nb.bad.as<int>(), s));
  nb.fp32 = store_fp32(storage);  // a smoother on FP32 tangents reads the inverses in FP32
  if (nb.fp32) {
    HIPCHK(op->ctx, nb.Binv32.alloc((size_t)std::max<int64_t>(nb.nblocks, 1) * 9 * sizeof(float)));
    HIPCHK(op->ctx, vec_to_fp32(nb.nblocks * 9, nb.Binv.as<double>(), nb.Binv32.as<float>(), s));
  }
  int bad = 0;
  HIPCHK(op->ctx, hipMemcpyAsync(&bad, nb.bad.p, sizeof(int), hipMemcpyDeviceToHost, s));
  HIPCHK(op->ctx, hipStreamSynchronize(s));
  *ok = bad == 0;
  return HFK_OK;
}

// one block step of a smoother: z = M^-1 r; d = cd d + cz z; x += d (the FP32 inverses for FP32 tangents)
hipError_t smoother_block_step(const NodeBlocks& nb, const double* r, double* d, double* x, double cd, double cz,
                               hipStream_t s) {
  if (nb.fp32) return blocks_step(nb.nblocks, nb.dofs.as<int32_t>(), nb.Binv32.as<float>(), r, d, x, cd, cz, s);
  return blocks_step(nb.nblocks, nb.dofs.as<int32_t>(), nb.Binv.as<double>(), r, d, x, cd, cz, s);
}

int blocks_apply(hfk_op* op, const double* r, double* y, hipStream_t s) {
  NodeBlocks& nb = op->blocks;
  if (!nb.built) return set_err(op->ctx, HFK_ESTATE, "HyperFEMKernels: BlockJacobi applied before its update");
  HIPCHK(op->ctx, blocks_step(nb.nblocks, nb.dofs.as<int32_t>(), nb.Binv.as<double>(), r, y, nullptr, 0.0, 1.0, s));
  return HFK_OK;
}

}  // namespace

// ---------------------------------------------------------------------------------------------------------
struct hfk_cheb {
  hfk_op* op = nullptr;
  int degree = 2, eig_iters = 10;
  double range = 20, inflate = 1.1;
  double lmin = 0, lmax = 0, a = 0, b = 0;
  bool ok = false, blocks_ok = false;
  bool secondary = false;  // apply the operator's second tangents (another precision) instead of its own
  // an assembled operator instead of the stored tangents (CSC of A, 0-based; A symmetric): y = alpha A x + beta y
  const int32_t* csr_colptr = nullptr;
  const int32_t* csr_rowval = nullptr;
  const double* csr_val = nullptr;
  // the same matrix in node-block ELL form when available (values FP64, or FP32 with bell_fp32)
  int64_t bell_nb = 0;
  int bell_W = 0;
  const int32_t* bell_nbr = nullptr;
  const int32_t* bell_dofs = nullptr;
  const void* bell_val = nullptr;
  bool bell_fp32 = false;
  // with an assembled operator: the nodal blocks still come from the level's stored tangents (else from the
  // assembled matrix, set up by its owner)
  bool csr_tangent_blocks = false;
  int nbuild = 0;
  DevBuf r, d, tmp, start, V, hess, hpart;
  Dots dots;
  const void* T() const { return secondary ? op->tangents_s.p : op->tangents.p; }
  int storage() const { return secondary ? op->storage_s : op->storage; }
};

namespace {

// y = alpha A x (+ y when accumulate) with the smoother's operator
int cheb_op_apply(hfk_cheb* ch, const double* x, double* y, double alpha, bool accumulate, hipStream_t s) {
  hfk_op* op = ch->op;
  if (ch->bell_val) {
    HIPCHK(op->ctx, bell_spmv(ch->bell_nb, ch->bell_W, ch->bell_nbr, ch->bell_val, ch->bell_fp32, ch->bell_dofs, x, y, alpha,
                              accumulate ? 1.0 : 0.0, s));
    return HFK_OK;
  }
  if (ch->csr_val) {
    HIPCHK(op->ctx, csc_t_spmv(op->nfree, ch->csr_colptr, ch->csr_rowval, ch->csr_val, x, y, alpha, accumulate ? 1.0 : 0.0, s));
    return HFK_OK;
  }
  return op_run_apply_t(op, ch->T(), ch->storage(), x, y, s, alpha, accumulate);
}

// eigenvalue_bounds (src/preconditioners.jl:374-389): k Arnoldi steps (modified Gram-Schmidt) on J M^-1 from
// `start`; extreme real parts of the Ritz values of the (truncated at a near breakdown) Hessenberg matrix
int arnoldi_bounds(hfk_op* op, hfk_cheb* ch, int k, hipStream_t s, double* lmin, double* lmax) {
  const int64_t n = op->nfree;
  k = (int)std::min<int64_t>(k, n);
  HIPCHK(op->ctx, ch->V.alloc((size_t)(k + 1) * n * sizeof(double)));
  double* V = ch->V.as<double>();
  HIPCHK(op->ctx, vec_copy(n, ch->start.as<double>(), V, s));
  // The Hessenberg matrix stays on the device until the last step: each step orthogonalises against the basis with
  // classical Gram-Schmidt applied twice (two batched projections, as GMRES does), so no step waits for the host;
  // the matrix is read back once.  A zero or non-finite norm leaves the later columns non-finite: the host stops
  // at the first such step, as a step-by-step process would.
  const int ldh = k + 1;
  HIPCHK(op->ctx, ch->hess.alloc((size_t)(ldh * k + 2 * ldh + 4) * sizeof(double)));
  HIPCHK(op->ctx, ch->hpart.alloc(std::max(multi_dot_partial_doubles(ldh), dot_partial_doubles()) * sizeof(double)));
  double* Hd = ch->hess.as<double>();
  double* c1 = Hd + (size_t)ldh * k;
  double* c2 = c1 + ldh;
  double* ww = c2 + ldh;
  double* part = ch->hpart.as<double>();
  HIPCHK(op->ctx, launch_fill(Hd, (int64_t)ldh * k, 0.0, s));
  int rc;
  for (int j = 0; j < k; ++j) {
    double* vj = V + (size_t)j * n;
    double* w = V + (size_t)(j + 1) * n;
    if ((rc = blocks_apply(op, vj, ch->tmp.as<double>(), s))) return rc;
    if ((rc = cheb_op_apply(ch, ch->tmp.as<double>(), w, 1.0, false, s))) return rc;
    HIPCHK(op->ctx, multi_dot(n, V, n, j + 1, w, part, c1, s));
    HIPCHK(op->ctx, multi_axpy(n, V, n, j + 1, c1, -1.0, w, s));
    HIPCHK(op->ctx, multi_dot(n, V, n, j + 1, w, part, c2, s));
    HIPCHK(op->ctx, multi_axpy(n, V, n, j + 1, c2, -1.0, w, s));
    HIPCHK(op->ctx, vec_dots(n, w, w, nullptr, nullptr, nullptr, nullptr, part, ww, s));
    HIPCHK(op->ctx, arnoldi_column(n, j, ldh, c1, c2, ww, Hd, w, s));
  }
  std::vector<double> Hm((size_t)ldh * k, 0.0);  // (k+1) x k column-major
  HIPCHK(op->ctx, hipMemcpyAsync(Hm.data(), Hd, Hm.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(op->ctx, hipStreamSynchronize(s));
  int steps = k;
  for (int j = 0; j < k; ++j) {
    const double hn = Hm[j + 1 + ldh * j];
    if (!(hn > 0) || !std::isfinite(hn)) {
      steps = j + 1;
      break;
    }
  }
  // truncate at the first (near) breakdown, as the reference
  int jj = steps;
  for (int i = 1; i < steps; ++i) {
    double mx = 1.0;
    for (int c = 0; c < i; ++c)
      for (int r = 0; r < i; ++r) mx = std::max(mx, std::fabs(Hm[r + (k + 1) * c]));
    if (std::fabs(Hm[i + (k + 1) * (i - 1)]) <= 1e-8 * mx) {
      jj = i;
      break;
    }
  }
  std::vector<double> h((size_t)jj * jj);
  bool fin = true;
  for (int c = 0; c < jj; ++c)
    for (int r = 0; r < jj; ++r) {
      h[r + jj * c] = Hm[r + (k + 1) * c];
      fin = fin && std::isfinite(h[r + jj * c]);
    }
  if (!fin) {
    *lmin = *lmax = NAN;
    return HFK_OK;
  }
  std::vector<double> ev;
  hessenberg_eig_real(h, jj, ev);
  *lmin = *std::min_element(ev.begin(), ev.end());
  *lmax = *std::max_element(ev.begin(), ev.end());
  return HFK_OK;
}

int cheb_update(hfk_cheb* ch, hipStream_t s) {
  hfk_op* op = ch->op;
  bool bok = ch->blocks_ok;
  int rc;
  // an assembled operator brings its nodal blocks (and their check) from its own set-up
  if ((!ch->csr_val || ch->csr_tangent_blocks) && (rc = blocks_update(op, ch->T(), ch->storage(), s, &bok))) return rc;
  ch->blocks_ok = bok;
  if (!bok) {
    ch->ok = false;
    return HFK_OK;
  }
  double lmin, lmax;
  if ((rc = arnoldi_bounds(op, ch, ch->eig_iters, s, &lmin, &lmax))) return rc;
  ch->lmin = lmin;
  ch->lmax = lmax;
  // smoother_interval! (src/smoothers.jl:92-107): usable only for a finite positive spectrum estimate
  ch->ok = std::isfinite(lmax) && lmax > 0 && lmin > 0;
  if (!ch->ok) return HFK_OK;
  ch->b = ch->inflate * lmax;
  ch->a = ch->b / ch->range;
  ch->nbuild++;
  return HFK_OK;
}

// smooth! (src/smoothers.jl:116-146), the block application fused with the direction update and x += d
int cheb_smooth(hfk_cheb* ch, double* x, const double* rhs, bool zero_guess, double* residual, hipStream_t s) {
  hfk_op* op = ch->op;
  if (!ch->ok)
    return set_err(op->ctx, HFK_ESTATE,
                   "HyperFEMKernels: Chebyshev: the smoother has no usable interval (non-finite or indefinite tangent)");
  const int64_t n = op->nfree;
  double* r = ch->r.as<double>();
  double* d = ch->d.as<double>();
  const NodeBlocks& nb = op->blocks;
  const double th = 0.5 * (ch->b + ch->a), de = 0.5 * (ch->b - ch->a), sg = th / de;
  int rc;
  HIPCHK(op->ctx, vec_copy(n, rhs, r, s));
  if (zero_guess) {
    HIPCHK(op->ctx, launch_fill(x, n, 0.0, s));
  } else {
    if ((rc = cheb_op_apply(ch, x, r, -1.0, true, s))) return rc;  // r = rhs - A x
  }
  // z = M^-1 r; d = z / theta; x += d
  HIPCHK(op->ctx, smoother_block_step(nb, r, d, x, 0.0, 1.0 / th, s));
  double rho = 1.0 / sg;
  const int k = ch->degree;
  for (int i = 1; i <= k; ++i) {
    if (i == k && !residual) break;
    if ((rc = cheb_op_apply(ch, d, r, -1.0, true, s))) return rc;  // r -= A d
    if (i == k) break;
    const double rn = 1.0 / (2.0 * sg - rho);
    HIPCHK(op->ctx, smoother_block_step(nb, r, d, x, rn * rho, 2.0 * rn / de, s));
    rho = rn;
  }
  if (residual) HIPCHK(op->ctx, vec_copy(n, r, residual, s));
  return HFK_OK;
}

int cheb_init(hfk_op* op, int degree, double range, double inflate, int eig_iters, hfk_cheb* ch) {
  ch->op = op;
  ch->degree = degree;
  ch->range = range;
  ch->inflate = inflate;
  ch->eig_iters = eig_iters;
  const int64_t n = op->nfree;
  int rc;
  if ((rc = vec_alloc(op->ctx, ch->r, n)) || (rc = vec_alloc(op->ctx, ch->d, n)) || (rc = vec_alloc(op->ctx, ch->tmp, n)) ||
      (rc = vec_alloc(op->ctx, ch->start, n)) || (rc = ch->dots.init(op->ctx)))
    return rc;
  std::vector<double> v = chebyshev_start(n);
  HIPCHK(op->ctx, hipMemcpy(ch->start.p, v.data(), n * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(op->ctx, launch_fill(ch->d.as<double>(), n, 0.0, op->ctx->stream));
  return HFK_OK;
}

int check_cheb_args(hfk_ctx* ctx, int degree, double range, double inflate, int eig_iters) {
  if (degree < 1) return set_err(ctx, HFK_EARG, "HyperFEMKernels: Chebyshev: degree must be >= 1");
  if (!(range > 1)) return set_err(ctx, HFK_EARG, "HyperFEMKernels: Chebyshev: range must be > 1");
  if (!(inflate >= 1)) return set_err(ctx, HFK_EARG, "HyperFEMKernels: Chebyshev: inflate must be >= 1");
  if (eig_iters < 2) return set_err(ctx, HFK_EARG, "HyperFEMKernels: Chebyshev: eig_iters must be >= 2");
  return HFK_OK;
}

}  // namespace

// ---------------------------------------------------------------------------------------------------------
// Prolongation between nested Q_k spaces on one mesh (src/transfers.jl:16-154)
struct hfk_transfer {
  hfk_op* coarse = nullptr;
  hfk_op* fine = nullptr;
  int NF = 0, NC = 0;
  DevBuf E, Einj, cid, invmult;
  // between nested meshes (the fine operator refines the coarse one): P (fine free x coarse free) and R = P' as CSR,
  // and the fine source of every coarse free and Dirichlet value for the injection
  bool nested = false;
  DevBuf prp, pcol, pval, rrp, rcol, rval, inj, injd;
  // a transfer on one mesh to Q1 (at most 8 coarse dofs per fine row) keeps R as CSR and prolongates from the cell
  // owning each fine dof: each fine value written once and each coarse sum formed by one group of lanes, instead
  // of per cell with a colour per launch
  bool csr = false;
  int p_lanes = 1, r_lanes = 1;  // threads per row of the two CSR products
  DevBuf owner;                  // same mesh, coarse Q1: the cell slot (slot * 3 + c) * NF + tf of each fine free dof
};

namespace {

// P (CSR over the fine free dofs, its rows given with sorted columns; kept on the device when with_p) and R = P'
int upload_transfer_csr(hfk_ctx* ctx, hfk_transfer* T, int64_t nf, int64_t nc, const std::vector<int64_t>& prp,
                        const std::vector<int32_t>& pcol, const std::vector<double>& pval, bool with_p) {
  std::vector<int64_t> rrp((size_t)nc + 1, 0);
  for (int64_t k = 0; k < prp[nf]; ++k) rrp[pcol[k] + 1]++;
  for (int64_t j = 0; j < nc; ++j) rrp[j + 1] += rrp[j];
  std::vector<int32_t> rcol((size_t)prp[nf]);
  std::vector<double> rval((size_t)prp[nf]);
  std::vector<int64_t> fill(rrp.begin(), rrp.end() - 1);
  for (int64_t i = 0; i < nf; ++i)
    for (int64_t k = prp[i]; k < prp[i + 1]; ++k) {
      const int64_t q = fill[pcol[k]]++;
      rcol[q] = (int32_t)i;
      rval[q] = pval[k];
    }
  auto up = [&](DevBuf& b, const void* h, size_t bytes) -> hipError_t {
    hipError_t e = b.alloc(std::max<size_t>(bytes, 8));
    if (e == hipSuccess && bytes) e = hipMemcpy(b.p, h, bytes, hipMemcpyHostToDevice);
    return e;
  };
  if (with_p) {
    HIPCHK(ctx, up(T->prp, prp.data(), prp.size() * sizeof(int64_t)));
    HIPCHK(ctx, up(T->pcol, pcol.data(), pcol.size() * sizeof(int32_t)));
    HIPCHK(ctx, up(T->pval, pval.data(), pval.size() * sizeof(double)));
  }
  HIPCHK(ctx, up(T->rrp, rrp.data(), rrp.size() * sizeof(int64_t)));
  HIPCHK(ctx, up(T->rcol, rcol.data(), rcol.size() * sizeof(int32_t)));
  HIPCHK(ctx, up(T->rval, rval.data(), rval.size() * sizeof(double)));
  T->p_lanes = csr_lanes(nf, prp[nf]);
  T->r_lanes = csr_lanes(nc, prp[nf]);
  return HFK_OK;
}

// A transfer between nested meshes (h-multigrid): the fine operator's cells each lie in one coarse cell (parent) as
// one of its 8 children.  P[f, c] = coarse shape function of dof c at the node of fine dof f (trilinear for Q1,
// snapped to 0/1 like the p-transfers), R = P', injection: each coarse node is a node of the fine mesh.  Built on
// the host from the two operators' cell dof tables.
int transfer_build_nested(hfk_op* coarse, hfk_op* fine, hfk_transfer* T) {
  hfk_ctx* ctx = fine->ctx;
  if (fine->refines != coarse)
    return set_err(ctx, HFK_EDIM, "HyperFEMKernels: the two levels live on different meshes and the fine one is not "
                                  "declared a refinement of the coarse one (hfk_op_set_refinement)");
  const int Qf = fine->Q, Qc = coarse->Q, NF = fine->nn, NC = coarse->nn;
  T->coarse = coarse;
  T->fine = fine;
  T->NF = NF;
  T->NC = NC;
  T->nested = true;
  auto lag = [](int Q, double xi, int i) {
    const double h = 1.0 / (Q - 1);
    double v = 1;
    for (int j = 0; j < Q; ++j)
      if (j != i) v *= (xi - j * h) / (i * h - j * h);
    return v;
  };
  auto snap = [](double e) { return std::fabs(e) < 1e-12 ? 0.0 : (std::fabs(e - 1) < 1e-12 ? 1.0 : e); };
  const int64_t nf = fine->nfree, nc = coarse->nfree;
  std::vector<std::vector<std::pair<int32_t, double>>> rows((size_t)nf);
  std::vector<int32_t> inj((size_t)nc, 0), injd((size_t)coarse->ndir, 0);
  for (int64_t f = 0; f < fine->ncells; ++f) {
    const int64_t pc = fine->parent[f];
    const int ch = fine->child[f];
    const int cpos[3] = {ch & 1, (ch >> 1) & 1, (ch >> 2) & 1};
    const int32_t* fd = &fine->cell_dofs[(size_t)f * fine->nd];
    const int32_t* cd = &coarse->cell_dofs[(size_t)pc * coarse->nd];
    for (int tf = 0; tf < NF; ++tf) {
      const int fi[3] = {tf % Qf, (tf / Qf) % Qf, tf / (Qf * Qf)};
      double xi[3];
      for (int d = 0; d < 3; ++d) xi[d] = 0.5 * (cpos[d] + fi[d] / double(Qf - 1));
      for (int c = 0; c < 3; ++c) {
        const int32_t fid = fd[fine->ldof[c * NF + fine->nodeperm[tf]]];
        for (int tc = 0; tc < NC; ++tc) {
          const int ci[3] = {tc % Qc, (tc / Qc) % Qc, tc / (Qc * Qc)};
          double e = 1;
          for (int d = 0; d < 3; ++d) e *= lag(Qc, xi[d], ci[d]);
          e = snap(e);
          if (e == 0.0) continue;
          const int32_t cid = cd[coarse->ldof[c * NC + coarse->nodeperm[tc]]];
          if (cid > 0 && fid <= 0)
            return set_err(ctx, HFK_ESTATE, "HyperFEMKernels: the two levels have different Dirichlet boundaries (a "
                                            "fine Dirichlet dof is interpolated from a free coarse dof)");
          if (cid <= 0 || fid <= 0) continue;
          auto& row = rows[fid - 1];
          bool seen = false;
          for (auto& pr : row) seen = seen || pr.first == cid - 1;
          if (!seen) row.emplace_back(cid - 1, e);
        }
      }
    }
    // injection: the coarse nodes of the parent that are nodes of this child
    for (int tc = 0; tc < NC; ++tc) {
      const int ci[3] = {tc % Qc, (tc / Qc) % Qc, tc / (Qc * Qc)};
      int fi[3];
      bool inside = true;
      for (int d = 0; d < 3; ++d) {
        const double t = (2.0 * ci[d] / double(Qc - 1) - cpos[d]) * (Qf - 1);
        fi[d] = (int)std::lround(t);
        inside = inside && t > -1e-9 && t < Qf - 1 + 1e-9 && std::fabs(t - fi[d]) < 1e-9;
      }
      if (!inside) continue;
      const int tf = fi[0] + Qf * (fi[1] + Qf * fi[2]);
      for (int c = 0; c < 3; ++c) {
        const int32_t cid = cd[coarse->ldof[c * NC + coarse->nodeperm[tc]]];
        const int32_t fid = fd[fine->ldof[c * NF + fine->nodeperm[tf]]];
        if ((cid > 0) != (fid > 0))
          return set_err(ctx, HFK_ESTATE, "HyperFEMKernels: the two levels have different Dirichlet boundaries (a node "
                                          "is free on one level and Dirichlet on the other)");
        if (cid > 0) inj[cid - 1] = fid;
        else injd[-cid - 1] = fid;
      }
    }
  }
  for (int32_t v : inj)
    if (v == 0) return set_err(ctx, HFK_ESTATE, "HyperFEMKernels: a coarse node is not a node of the fine mesh");
  for (int32_t v : injd)
    if (v == 0) return set_err(ctx, HFK_ESTATE, "HyperFEMKernels: a coarse Dirichlet node is not a node of the fine mesh");
  // P (CSR over the fine free dofs, columns sorted) and R = P'
  std::vector<int64_t> prp((size_t)nf + 1, 0);
  for (int64_t i = 0; i < nf; ++i) {
    std::sort(rows[i].begin(), rows[i].end());
    prp[i + 1] = prp[i] + (int64_t)rows[i].size();
  }
  std::vector<int32_t> pcol((size_t)prp[nf]);
  std::vector<double> pval((size_t)prp[nf]);
  for (int64_t i = 0; i < nf; ++i)
    for (size_t k = 0; k < rows[i].size(); ++k) {
      pcol[prp[i] + k] = rows[i][k].first;
      pval[prp[i] + k] = rows[i][k].second;
    }
  if (int rc = upload_transfer_csr(ctx, T, nf, nc, prp, pcol, pval, true)) return rc;
  auto up = [&](DevBuf& b, const void* h, size_t bytes) -> hipError_t {
    hipError_t e = b.alloc(std::max<size_t>(bytes, 8));
    if (e == hipSuccess && bytes) e = hipMemcpy(b.p, h, bytes, hipMemcpyHostToDevice);
    return e;
  };
  HIPCHK(ctx, up(T->inj, inj.data(), inj.size() * sizeof(int32_t)));
  HIPCHK(ctx, up(T->injd, injd.data(), injd.size() * sizeof(int32_t)));
  return HFK_OK;
}

int transfer_build(hfk_op* coarse, hfk_op* fine, hfk_transfer* T) {
  hfk_ctx* ctx = fine->ctx;
  if (coarse->ncells != fine->ncells || fine->refines == coarse) return transfer_build_nested(coarse, fine, T);
  if (coarse->Q >= fine->Q) return set_err(ctx, HFK_EARG, "HyperFEMKernels: the coarse space has more nodes per cell than the fine one");
  const int Qf = fine->Q, Qc = coarse->Q, NF = fine->nn, NC = coarse->nn;
  T->coarse = coarse;
  T->fine = fine;
  T->NF = NF;
  T->NC = NC;
  // E[tf, tc] = coarse shape function tc at fine node tf; Einj[tc, tf] = fine shape function tf at coarse node tc
  auto lag = [](int Q, double xi, int i) {
    const double h = 1.0 / (Q - 1);
    double v = 1;
    for (int j = 0; j < Q; ++j)
      if (j != i) v *= (xi - j * h) / (i * h - j * h);
    return v;
  };
  auto snap = [](double e) { return std::fabs(e) < 1e-7 ? 0.0 : (std::fabs(e - 1) < 1e-7 ? 1.0 : e); };
  std::vector<double> E((size_t)NF * NC), Ei((size_t)NC * NF);
  for (int tf = 0; tf < NF; ++tf)
    for (int tc = 0; tc < NC; ++tc) {
      const int fi[3] = {tf % Qf, (tf / Qf) % Qf, tf / (Qf * Qf)};
      const int ci[3] = {tc % Qc, (tc / Qc) % Qc, tc / (Qc * Qc)};
      double e = 1, ei = 1;
      for (int d = 0; d < 3; ++d) {
        e *= lag(Qc, fi[d] / double(Qf - 1), ci[d]);
        ei *= lag(Qf, ci[d] / double(Qc - 1), fi[d]);
      }
      E[(size_t)tf * NC + tc] = snap(e);
      Ei[(size_t)tc * NF + tf] = snap(ei);
    }
  // coarse ids in the fine operator's batch layout, Dirichlet consistency (src/transfers.jl:112-127)
  const int64_t nslot = fine->nslots();
  std::vector<int32_t> cid((size_t)nslot * 3 * NC, 0);
  for (int64_t sl = 0; sl < nslot; ++sl) {
    const int64_t cell = fine->batch_cell[sl];
    if (cell < 0) continue;
    const int32_t* cd = &coarse->cell_dofs[(size_t)cell * coarse->nd];
    const int32_t* fd = &fine->cell_dofs[(size_t)cell * fine->nd];
    for (int c = 0; c < 3; ++c)
      for (int tc = 0; tc < NC; ++tc) cid[((size_t)sl * 3 + c) * NC + tc] = cd[coarse->ldof[c * NC + coarse->nodeperm[tc]]];
    for (int c = 0; c < 3; ++c) {
      for (int tf = 0; tf < NF; ++tf) {
        if (fd[fine->ldof[c * NF + fine->nodeperm[tf]]] > 0) continue;
        for (int tc = 0; tc < NC; ++tc)
          if (E[(size_t)tf * NC + tc] != 0 && cid[((size_t)sl * 3 + c) * NC + tc] > 0)
            return set_err(ctx, HFK_ESTATE,
                           "HyperFEMKernels: the two levels have different Dirichlet boundaries (a fine Dirichlet dof is "
                           "interpolated from a free coarse dof; use the same closed Dirichlet set on every level)");
      }
      for (int tc = 0; tc < NC; ++tc) {
        if (cid[((size_t)sl * 3 + c) * NC + tc] > 0) continue;
        for (int tf = 0; tf < NF; ++tf)
          if (Ei[(size_t)tc * NF + tf] != 0 && fd[fine->ldof[c * NF + fine->nodeperm[tf]]] > 0)
            return set_err(ctx, HFK_ESTATE,
                           "HyperFEMKernels: the two levels have different Dirichlet boundaries (a coarse Dirichlet dof "
                           "is interpolated from a free fine dof)");
      }
    }
  }
  std::vector<double> inv((size_t)fine->nfree, 0.0);
  for (int64_t cell = 0; cell < fine->ncells; ++cell)
    for (int j = 0; j < fine->nd; ++j) {
      const int32_t id = fine->cell_dofs[(size_t)cell * fine->nd + j];
      if (id > 0) inv[id - 1] += 1.0;
    }
  for (auto& v : inv) v = v > 0 ? 1.0 / v : 0.0;
  HIPCHK(ctx, T->E.alloc(E.size() * sizeof(double)));
  HIPCHK(ctx, hipMemcpy(T->E.p, E.data(), E.size() * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(ctx, T->Einj.alloc(Ei.size() * sizeof(double)));
  HIPCHK(ctx, hipMemcpy(T->Einj.p, Ei.data(), Ei.size() * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(ctx, T->cid.alloc(cid.size() * sizeof(int32_t)));
  HIPCHK(ctx, hipMemcpy(T->cid.p, cid.data(), cid.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  HIPCHK(ctx, T->invmult.alloc(inv.size() * sizeof(double)));
  HIPCHK(ctx, hipMemcpy(T->invmult.p, inv.data(), inv.size() * sizeof(double), hipMemcpyHostToDevice));
  if (NC == 8) {
    // the row of a fine free dof from the first cell holding it: every cell holding the fine node holds the coarse
    // nodes whose shape functions are nonzero there (they lie on the smallest face or edge through the node), with
    // the same E, so the rows agree across cells; each fine dof summed once makes R = P' the sum over cells of
    // E yf / mult
    const int64_t nf = fine->nfree, nc = coarse->nfree;
    std::vector<int64_t> owner((size_t)nf, -1);  // (slot * 3 + c) * NF + tf
    for (int64_t sl = 0; sl < nslot; ++sl) {
      const int64_t cell = fine->batch_cell[sl];
      if (cell < 0) continue;
      const int32_t* fd = &fine->cell_dofs[(size_t)cell * fine->nd];
      for (int c = 0; c < 3; ++c)
        for (int tf = 0; tf < NF; ++tf) {
          const int32_t id = fd[fine->ldof[c * NF + fine->nodeperm[tf]]];
          if (id > 0 && owner[id - 1] < 0) owner[id - 1] = ((int64_t)sl * 3 + c) * NF + tf;
        }
    }
    std::vector<int64_t> prp((size_t)nf + 1, 0);
    std::vector<int32_t> pcol;
    std::vector<double> pval;
    pcol.reserve((size_t)nf * 4);
    pval.reserve((size_t)nf * 4);
    for (int64_t i = 0; i < nf; ++i) {
      std::pair<int32_t, double> row[8];
      int m = 0;
      if (owner[i] >= 0) {
        const int64_t sc = owner[i] / NF;
        const int tf = (int)(owner[i] - sc * NF);
        for (int tc = 0; tc < NC; ++tc) {
          const double e = E[(size_t)tf * NC + tc];
          const int32_t j = cid[(size_t)sc * NC + tc];
          if (e != 0.0 && j > 0) row[m++] = {j - 1, e};
        }
      }
      std::sort(row, row + m);
      for (int k = 0; k < m; ++k) {
        pcol.push_back(row[k].first);
        pval.push_back(row[k].second);
      }
      prp[i + 1] = (int64_t)pcol.size();
    }
    if (int rc = upload_transfer_csr(ctx, T, nf, nc, prp, pcol, pval, false)) return rc;
    HIPCHK(ctx, T->owner.alloc(std::max<size_t>(owner.size() * sizeof(int64_t), 8)));
    if (nf) HIPCHK(ctx, hipMemcpy(T->owner.p, owner.data(), owner.size() * sizeof(int64_t), hipMemcpyHostToDevice));
    T->csr = true;
  }
  return HFK_OK;
}

// prolongate! (src/transfers.jl:238-245): yf = P xc, Dirichlet dofs read as zero and never written
int do_prolongate(hfk_transfer* T, const double* xc, double* yf, hipStream_t s) {
  hfk_op* f = T->fine;
  if (T->csr) {  // one thread per fine dof from its owning cell: each value written once, no matrix read
    HIPCHK(f->ctx, transfer_prolongate_owner(f->nfree, T->NF, T->owner.as<int64_t>(), T->cid.as<int32_t>(),
                                             T->E.as<double>(), xc, yf, s));
    return HFK_OK;
  }
  if (T->nested) {
    HIPCHK(f->ctx, csr_spmv_lanes(f->nfree, T->prp.as<int64_t>(), T->pcol.as<int32_t>(), T->pval.as<double>(), xc, yf,
                                  T->p_lanes, s));
    return HFK_OK;
  }
  HIPCHK(f->ctx, launch_fill(yf, f->nfree, 0.0, s));
  HIPCHK(f->ctx, transfer_prolongate(f->nslots(), T->NF, T->NC, f->ids.as<int32_t>(), T->cid.as<int32_t>(),
                                     T->E.as<double>(), xc, yf, s));
  return HFK_OK;
}

// restrict! (src/transfers.jl:252-259): yc = P' yf, colour by colour of the fine cells
int do_restrict(hfk_transfer* T, const double* yf, double* yc, hipStream_t s) {
  hfk_op* f = T->fine;
  if (T->nested || T->csr) {
    HIPCHK(f->ctx, csr_spmv_lanes(T->coarse->nfree, T->rrp.as<int64_t>(), T->rcol.as<int32_t>(), T->rval.as<double>(),
                                  yf, yc, T->r_lanes, s));
    return HFK_OK;
  }
  HIPCHK(f->ctx, launch_fill(yc, T->coarse->nfree, 0.0, s));
  for (const BatchRange& r : f->color_batches)
    HIPCHK(f->ctx, transfer_restrict((int64_t)r.first * f->C, (int64_t)r.count * f->C, T->NF, T->NC, f->ids.as<int32_t>(),
                                     T->cid.as<int32_t>(), T->E.as<double>(), T->invmult.as<double>(), yf, yc, s));
  return HFK_OK;
}

// inject! (src/transfers.jl:285-313)
int do_inject(hfk_transfer* T, const double* xf, const double* df, double* xc, double* dc, hipStream_t s) {
  hfk_op* f = T->fine;
  if (T->nested) {
    HIPCHK(f->ctx, inject_map(T->coarse->nfree, T->inj.as<int32_t>(), xf, df, xc, s));
    if (dc) HIPCHK(f->ctx, inject_map(T->coarse->ndir, T->injd.as<int32_t>(), xf, df, dc, s));
    return HFK_OK;
  }
  HIPCHK(f->ctx, transfer_inject(f->nslots(), T->NF, T->NC, f->ids.as<int32_t>(), T->cid.as<int32_t>(),
                                 T->Einj.as<double>(), xf, df, xc, dc, s));
  return HFK_OK;
}

}  // namespace

// ---------------------------------------------------------------------------------------------------------
// p-multigrid (src/pmultigrid.jl): levels fine first; the last is assembled
struct hfk_pmg {
  hfk_ctx* ctx = nullptr;
  std::vector<hfk_op*> ops;
  std::vector<std::unique_ptr<hfk_transfer>> P;  // P[l]: level l+1 -> level l
  std::vector<std::unique_ptr<hfk_cheb>> cheb;   // smoothed levels 0..L-2
  std::vector<std::unique_ptr<DevBuf>> r, e, bc, xc;
  int coarse_kind = 0, coarse_maxiter = 500;  // 0: block-Jacobi PCG to coarse_rtol; 1: direct (dense Cholesky)
  DevBuf cinv;                                 // kind 1: the inverse of the coarse matrix, n x n
  hfk_net* net = nullptr;  // learned coarse correction (not owned)
  int net_mode = 0;
  // the plain cycle as the fallback of a solve with the network: set aside for the rest of a Krylov solve whose
  // residual has not fallen by fb_factor over the last fb_window iterations (fb_window 0: never)
  int fb_window = 16;
  double fb_factor = 0.5;
  bool net_suspended = false;
  int64_t fallbacks = 0;
  double coarse_rtol = 1e-2;
  // coarse CSR (CSC of A, 0-based int32)
  DevBuf colptr, rowval, nzval;
  int64_t nnz = 0;
  DevBuf cr, cz, cp, cq, ct, cu, cv, cs;  // inner CG work vectors and device scalars
  DevBuf pcg_part;                        // partial sums of the fused coarse PCG
  DevBuf cdscale;                         // D^-1/2 of the coarse matrix (network scaling)
  // the coarse matrix in node-block ELL form (3x3 blocks of the nodal dofs), filled from the CSC at every setup
  struct BlockEll {
    bool built = false, has32 = false;
    int64_t nb = 0;
    int W = 0;
    DevBuf nbr, pos, v64, v32;
  } bell;
  bool net_scaled = true;                 // the network acts on D^-1/2 A D^-1/2
  std::unique_ptr<hfk_cheb> ccheb;      // Chebyshev on the assembled coarse matrix (network cycle, modes 3 and 4)
  bool ccheb_stale = true;
  int ccheb_degree = 0;  // its degree (0: the matrix-free levels' degree)
  Dots dots;
  bool ok = false, indefinite = false, built = false, coarse_ok = false;
  // the coarse PCG's products on FP32 copies of the coarse values (FP64 vectors and accumulation)
  bool coarse_fp32 = false;
  // coarse re-assembly every coarse_every updates (PLevels coarse_every, src/pmultigrid.jl:200-216)
  int coarse_every = 1, since = 0;
  bool coarse_built = false;
  int64_t coarse_builds = 0;
  int nupdate = 0, ncycle = 0, last_inner = 0, total_inner = 0;
  int storage = 0;           // storage of the matrix-free coarse levels' tangents
  int smoother_storage = -1;  // -1: the working operator's own tangents; else a copy in this storage
  // Intermediate levels with at most kAssembledLevel free dofs: their eight colour launches of a few dozen
  // work-groups are bound by launch latency, so each update also assembles the level (CSC, FP64) and its smoother
  // applies it with one CSC product per Jv; the nodal blocks still come from the stored tangents.
  struct LevelCsc {
    bool built = false;
    DevBuf colptr, rowval, nzval;  // 0-based int32 pattern, FP64 values
  };
  std::vector<std::unique_ptr<LevelCsc>> lcsc;  // per level; null: matrix-free
};

namespace {

// The node-block ELL structure of the coarse matrix from its CSC pattern and nodal blocks (host, once): the
// neighbour blocks of every block row (sorted, padded to the widest row with the row's own block) and, for every
// CSC entry, its position in the ELL values.  Left unbuilt (the scalar CSC product is used) when the positions
// would not fit in 32 bits.
int bell_build(hfk_pmg* g) {
  hfk_op* c = g->ops.back();
  const NodeBlocks& nbk = c->blocks;
  const int64_t n = c->nfree, nb = nbk.nblocks;
  if (nb <= 0) return HFK_OK;
  std::vector<int32_t> blk((size_t)n, -1), comp((size_t)n, 0);
  for (int64_t b = 0; b < nb; ++b)
    for (int k = 0; k < 3; ++k) {
      const int32_t d = nbk.dofs_h[3 * b + k];
      if (d > 0) {
        blk[d - 1] = (int32_t)b;
        comp[d - 1] = k;
      }
    }
  for (int64_t d = 0; d < n; ++d)
    if (blk[d] < 0) return HFK_OK;  // a free dof outside the nodal blocks: keep the scalar product
  // neighbour lists: block j (of column c) is a neighbour of block i (of row r)
  constexpr int kMaxW = 64;
  std::vector<int32_t> lst((size_t)nb * kMaxW), cnt((size_t)nb, 0), last((size_t)nb, -1);
  for (int64_t col = 0; col < n; ++col) {
    const int32_t j = blk[col];
    for (int64_t p = c->colptr[col] - 1; p < c->colptr[col + 1] - 1; ++p) {
      const int32_t i = blk[c->rowval[p] - 1];
      if (last[i] == j) continue;
      bool found = false;
      for (int k = 0; k < cnt[i]; ++k) found = found || lst[(size_t)i * kMaxW + k] == j;
      if (found) continue;
      if (cnt[i] >= kMaxW) return HFK_OK;  // unusually wide block row: keep the scalar product
      lst[(size_t)i * kMaxW + cnt[i]++] = j;
      last[i] = j;
    }
  }
  int W = 1;
  for (int64_t i = 0; i < nb; ++i) {
    std::sort(lst.begin() + (size_t)i * kMaxW, lst.begin() + (size_t)i * kMaxW + cnt[i]);
    W = std::max(W, cnt[i]);
  }
  if ((int64_t)9 * W * nb >= ((int64_t)1 << 31)) return HFK_OK;
  std::vector<int32_t> nbr((size_t)W * nb);
  for (int64_t i = 0; i < nb; ++i)
    for (int k = 0; k < W; ++k) nbr[(size_t)k * nb + i] = k < cnt[i] ? lst[(size_t)i * kMaxW + k] : (int32_t)i;
  std::vector<int32_t> pos((size_t)g->nnz);
  for (int64_t col = 0; col < n; ++col) {
    const int32_t j = blk[col];
    for (int64_t p = c->colptr[col] - 1; p < c->colptr[col + 1] - 1; ++p) {
      const int64_t r = c->rowval[p] - 1;
      const int32_t i = blk[r];
      const int32_t* L = lst.data() + (size_t)i * kMaxW;
      const int k = (int)(std::lower_bound(L, L + cnt[i], j) - L);
      pos[p] = (int32_t)(((int64_t)9 * k + comp[r] + 3 * comp[col]) * nb + i);
    }
  }
  auto& B = g->bell;
  HIPCHK(g->ctx, B.nbr.alloc(nbr.size() * sizeof(int32_t)));
  HIPCHK(g->ctx, hipMemcpy(B.nbr.p, nbr.data(), nbr.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  HIPCHK(g->ctx, B.pos.alloc(std::max<size_t>(pos.size(), 1) * sizeof(int32_t)));
  if (!pos.empty()) HIPCHK(g->ctx, hipMemcpy(B.pos.p, pos.data(), pos.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  const size_t nv = (size_t)9 * W * nb;
  HIPCHK(g->ctx, B.v64.alloc(nv * sizeof(double)));
  B.has32 = (g->smoother_storage >= 0 && store_fp32(g->smoother_storage)) || g->coarse_fp32;
  if (B.has32) HIPCHK(g->ctx, B.v32.alloc(nv * sizeof(float)));
  B.nb = nb;
  B.W = W;
  B.built = true;
  return HFK_OK;
}

// y = alpha A x + beta y with the assembled coarse matrix
hipError_t coarse_spmv(hfk_pmg* g, const double* x, double* y, double alpha, double beta, hipStream_t s) {
  if (g->bell.built)
    return bell_spmv(g->bell.nb, g->bell.W, g->bell.nbr.as<int32_t>(), g->bell.v64.p, false,
                     g->ops.back()->blocks.dofs.as<int32_t>(), x, y, alpha, beta, s);
  return csc_t_spmv(g->ops.back()->nfree, g->colptr.as<int32_t>(), g->rowval.as<int32_t>(), g->nzval.as<double>(), x, y,
                    alpha, beta, s);
}

// The direct coarse solve (the role of the reference's LUSolver on the coarsest level, src/pmultigrid.jl:229-239):
// the symmetric part of the assembled matrix of a small coarsest level (at most kMaxDirect free dofs, the bottom of
// an h-hierarchy) is inverted on the device once per set-up (Gauss-Jordan, symmetric positive definite: no pivoting),
// so that every cycle applies it with one dense product.  *ok = false when a pivot is not positive or not finite.
constexpr int64_t kMaxDirect = 4000;

int coarse_direct_setup(hfk_pmg* g, hipStream_t s, bool* ok) {
  const int64_t n = g->ops.back()->nfree;
  DevBuf W, rc, bad;
  HIPCHK(g->ctx, W.alloc((size_t)n * n * sizeof(double)));
  HIPCHK(g->ctx, rc.alloc((size_t)2 * n * sizeof(double)));
  HIPCHK(g->ctx, bad.alloc(sizeof(int)));
  HIPCHK(g->ctx, g->cinv.alloc((size_t)n * n * sizeof(double)));
  HIPCHK(g->ctx, hipMemsetAsync(bad.p, 0, sizeof(int), s));
  HIPCHK(g->ctx, dense_spd_inverse(n, g->colptr.as<int32_t>(), g->rowval.as<int32_t>(), g->nzval.as<double>(),
                                   W.as<double>(), rc.as<double>(), rc.as<double>() + n, g->cinv.as<double>(),
                                   bad.as<int>(), s));
  int nbad = 0;
  HIPCHK(g->ctx, hipMemcpyAsync(&nbad, bad.p, sizeof(int), hipMemcpyDeviceToHost, s));
  HIPCHK(g->ctx, hipStreamSynchronize(s));  // also: W and rc are freed on return
  *ok = nbad == 0;
  return HFK_OK;
}

constexpr int64_t kAssembledLevel = 200000;

// the assembled matrix of intermediate level l at its injected state, handed to the level's smoother
int level_csc_update(hfk_pmg* g, size_t l, hipStream_t s) {
  hfk_op* op = g->ops[l];
  auto& L = *g->lcsc[l];
  int rc;
  if (!L.built) {
    if ((rc = op_build_pattern(op))) return rc;
    if ((int64_t)op->rowval.size() >= ((int64_t)1 << 31))
      return set_err(g->ctx, HFK_EDIM, "HyperFEMKernels: level matrix too large");
    std::vector<int32_t> cp(op->colptr.size()), rv(op->rowval.size());
    for (size_t i = 0; i < cp.size(); ++i) cp[i] = (int32_t)(op->colptr[i] - 1);
    for (size_t i = 0; i < rv.size(); ++i) rv[i] = (int32_t)(op->rowval[i] - 1);
    HIPCHK(g->ctx, L.colptr.alloc(cp.size() * sizeof(int32_t)));
    HIPCHK(g->ctx, hipMemcpy(L.colptr.p, cp.data(), cp.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    HIPCHK(g->ctx, L.rowval.alloc(std::max<size_t>(rv.size(), 1) * sizeof(int32_t)));
    HIPCHK(g->ctx, hipMemcpy(L.rowval.p, rv.data(), rv.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    HIPCHK(g->ctx, L.nzval.alloc(std::max<size_t>(rv.size(), 1) * sizeof(double)));
    L.built = true;
  }
  if ((rc = op_run_jacobian(op, op->state.as<double>(), L.nzval.as<double>(), s))) return rc;
  hfk_cheb* ch = g->cheb[l].get();
  ch->csr_colptr = L.colptr.as<int32_t>();
  ch->csr_rowval = L.rowval.as<int32_t>();
  ch->csr_val = L.nzval.as<double>();
  return HFK_OK;
}

int coarse_setup(hfk_pmg* g, hipStream_t s, bool* ok) {
  hfk_op* c = g->ops.back();
  int rc;
  if (!g->colptr.p) {
    if ((rc = op_build_pattern(c))) return rc;
    if ((int64_t)c->rowval.size() >= ((int64_t)1 << 31)) return set_err(g->ctx, HFK_EDIM, "HyperFEMKernels: coarse matrix too large");
    std::vector<int32_t> cp(c->colptr.size()), rv(c->rowval.size());
    for (size_t i = 0; i < cp.size(); ++i) cp[i] = (int32_t)(c->colptr[i] - 1);
    for (size_t i = 0; i < rv.size(); ++i) rv[i] = (int32_t)(c->rowval[i] - 1);
    g->nnz = (int64_t)rv.size();
    HIPCHK(g->ctx, g->colptr.alloc(cp.size() * sizeof(int32_t)));
    HIPCHK(g->ctx, hipMemcpy(g->colptr.p, cp.data(), cp.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    HIPCHK(g->ctx, g->rowval.alloc(std::max<size_t>(rv.size(), 1) * sizeof(int32_t)));
    HIPCHK(g->ctx, hipMemcpy(g->rowval.p, rv.data(), rv.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    HIPCHK(g->ctx, g->nzval.alloc(std::max<size_t>(rv.size(), 1) * sizeof(double)));
    if ((rc = op_build_blocks(c))) return rc;
    if ((rc = bell_build(g))) return rc;
  }
  if ((rc = op_run_jacobian(c, c->state.as<double>(), g->nzval.as<double>(), s))) return rc;
  HIPCHK(g->ctx, csc_diag_rsqrt(c->nfree, g->colptr.as<int32_t>(), g->rowval.as<int32_t>(), g->nzval.as<double>(),
                                g->cdscale.as<double>(), s));
  if (g->bell.built) {
    const size_t nv = (size_t)9 * g->bell.W * g->bell.nb;
    HIPCHK(g->ctx, hipMemsetAsync(g->bell.v64.p, 0, nv * sizeof(double), s));
    if (g->bell.has32) HIPCHK(g->ctx, hipMemsetAsync(g->bell.v32.p, 0, nv * sizeof(float), s));
    HIPCHK(g->ctx, bell_fill(g->nnz, g->bell.pos.as<int32_t>(), g->nzval.as<double>(), g->bell.v64.as<double>(),
                             g->bell.has32 ? g->bell.v32.as<float>() : nullptr, s));
  }
  if (g->coarse_kind == 1) {
    bool dok = false;
    if ((rc = coarse_direct_setup(g, s, &dok))) return rc;
    if (!dok) {
      *ok = false;
      return HFK_OK;
    }
  }
  NodeBlocks& nb = c->blocks;
  HIPCHK(g->ctx, csr_blocks(nb.nblocks, nb.dofs.as<int32_t>(), g->colptr.as<int32_t>(), g->rowval.as<int32_t>(),
                            g->nzval.as<double>(), nb.B.as<double>(), s));
  HIPCHK(g->ctx, hipMemsetAsync(nb.bad.p, 0, sizeof(int), s));
  HIPCHK(g->ctx, count_nonfinite(g->nnz, g->nzval.as<double>(), nb.bad.as<int>(), s));
  HIPCHK(g->ctx, blocks_invert(nb.nblocks, nb.dofs.as<int32_t>(), nb.B.as<double>(), nb.Binv.as<double>(), nb.bad.as<int>(), s));
  int bad = 0;
  HIPCHK(g->ctx, hipMemcpyAsync(&bad, nb.bad.p, sizeof(int), hipMemcpyDeviceToHost, s));
  HIPCHK(g->ctx, hipStreamSynchronize(s));
  *ok = bad == 0;
  return HFK_OK;
}

// Chebyshev smoother of the assembled coarse level (same degree, range and inflation as the matrix-free levels);
// set up at its first use after an update.  *usable = false when its interval is not usable.
int coarse_smoother(hfk_pmg* g, hipStream_t s, bool* usable) {
  hfk_op* c = g->ops.back();
  int rc;
  if (!g->ccheb) {
    const hfk_cheb* f = g->cheb[0].get();
    std::unique_ptr<hfk_cheb> ch(new hfk_cheb());
    if ((rc = cheb_init(c, g->ccheb_degree > 0 ? g->ccheb_degree : f->degree, f->range, f->inflate, f->eig_iters, ch.get())))
      return rc;
    ch->csr_colptr = g->colptr.as<int32_t>();
    ch->csr_rowval = g->rowval.as<int32_t>();
    ch->csr_val = g->nzval.as<double>();
    if (g->bell.built) {  // FP32 values when the p-multigrid smooths in FP32 (not for an FP32 coarse PCG alone)
      const bool f32 = g->bell.has32 && g->smoother_storage >= 0 && store_fp32(g->smoother_storage);
      ch->bell_nb = g->bell.nb;
      ch->bell_W = g->bell.W;
      ch->bell_nbr = g->bell.nbr.as<int32_t>();
      ch->bell_dofs = c->blocks.dofs.as<int32_t>();
      ch->bell_fp32 = f32;
      ch->bell_val = f32 ? g->bell.v32.p : g->bell.v64.p;
    }
    g->ccheb = std::move(ch);
  }
  if (g->ccheb_stale) {
    g->ccheb->blocks_ok = g->coarse_ok;
    if ((rc = cheb_update(g->ccheb.get(), s))) return rc;
    g->ccheb_stale = false;
  }
  *usable = g->ccheb->ok;
  return HFK_OK;
}

// the network's approximation z of A^-1 r on the coarse level: z = (s/h) N(r/s) (s: RMS of the input, h: the
// network's cell size); with net_scaled the network acts on the diagonally scaled system D^-1/2 A D^-1/2
// (D = diag(A), unit diagonal, so no 1/h factor): z = s D^-1/2 N(D^-1/2 r / s), s the RMS of D^-1/2 r
int pmg_net_apply(hfk_pmg* g, hfk_net* net, const double* r, double* z, hipStream_t s) {
  if (!g->net_scaled) return net_apply_dev(net, r, z, s);
  const int64_t n = g->ops.back()->nfree;
  const double* d = g->cdscale.as<double>();
  double* v = g->cv.as<double>();
  HIPCHK(g->ctx, vec_mul(n, d, r, v, s));
  int rc = net_apply_dev(net, v, z, s);  // (s/h) N(v/s)
  if (rc) return rc;
  HIPCHK(g->ctx, vec_mul(n, d, z, z, s));
  HIPCHK(g->ctx, vec_axpby(n, 0.0, z, net_h(net), z, s));  // times h
  return HFK_OK;
}

// one two-grid cycle on the coarse level with the network as its coarse-grid correction:
// x = S b; x += N(b - A x); x = x + S (b - A x)   (S: the coarse Chebyshev smoother)
int net_cycle(hfk_pmg* g, const double* b, double* x, hipStream_t s) {
  hfk_cheb* S = g->ccheb.get();
  const int64_t n = g->ops.back()->nfree;
  double* t = g->ct.as<double>();
  double* u = g->cu.as<double>();
  int rc;
  if ((rc = cheb_smooth(S, x, b, true, t, s))) return rc;
  if ((rc = pmg_net_apply(g, g->net, t, u, s))) return rc;
  HIPCHK(g->ctx, vec_axpby(n, 1.0, u, 1.0, x, s));
  return cheb_smooth(S, x, b, false, nullptr, s);
}

// coarse solve on the assembled coarse matrix from zero to coarse_rtol (relative).  Without the network:
// block-Jacobi PCG.  With it (set_coarse_network): mode 1 starts the PCG from the network's output, mode 2 takes
// the network's output as the solution, mode 3 runs flexible CG preconditioned by the network cycle (net_cycle),
// mode 4 takes one network cycle as the solution.  Modes 3 and 4 fall back to the PCG when the coarse smoother
// has no usable interval.
int coarse_solve(hfk_pmg* g, const double* b, double* x, hipStream_t s, bool allow_net, double rtol, int maxiter) {
  hfk_op* c = g->ops.back();
  const int64_t n = c->nfree;
  NodeBlocks& nb = c->blocks;
  double* r = g->cr.as<double>();
  double* z = g->cz.as<double>();
  double* p = g->cp.as<double>();
  double* q = g->cq.as<double>();
  double* sc = g->cs.as<double>();
  int mode = allow_net && g->net && !g->net_suspended ? g->net_mode : 0;
  int rc;
  g->last_inner = 0;
  if (mode >= 3) {
    bool usable = false;
    if ((rc = coarse_smoother(g, s, &usable))) return rc;
    if (!usable) mode = 0;
  }
  if (mode == 4) return net_cycle(g, b, x, s);
  if (mode == 0 && g->coarse_kind == 1) {  // direct: x = A^-1 b
    HIPCHK(g->ctx, dense_matvec(n, g->cinv.as<double>(), b, x, s));
    return HFK_OK;
  }
  const bool flexible = mode == 3;
  auto precondition = [&](const double* rin, double* zout) -> int {
    if (flexible) return net_cycle(g, rin, zout, s);
    HIPCHK(g->ctx, blocks_step(nb.nblocks, nb.dofs.as<int32_t>(), nb.Binv.as<double>(), rin, zout, nullptr, 0.0, 1.0, s));
    return HFK_OK;
  };
  if (mode == 1 || mode == 2) {
    if ((rc = pmg_net_apply(g, g->net, b, x, s))) return rc;
    if (mode == 2) return HFK_OK;
    HIPCHK(g->ctx, vec_copy(n, b, r, s));  // r = b - A x0
    HIPCHK(g->ctx, coarse_spmv(g, x, r, -1.0, 1.0, s));
  } else {
    HIPCHK(g->ctx, launch_fill(x, n, 0.0, s));
    HIPCHK(g->ctx, vec_copy(n, b, r, s));
  }
  if ((rc = precondition(r, z))) return rc;
  HIPCHK(g->ctx, vec_copy(n, z, p, s));
  double h[3];
  if ((rc = g->dots.run(g->ctx, s, n, r, z, h, b, b, r, r))) return rc;
  const double bn = std::sqrt(h[1]);
  HIPCHK(g->ctx, vec_dots(n, r, z, nullptr, nullptr, nullptr, nullptr, g->dots.part.as<double>(), sc + kRZ, s));
  if (bn == 0 || !(std::sqrt(h[2]) > rtol * bn)) return HFK_OK;
  if (!flexible && g->bell.built) {  // block-Jacobi PCG on the FP64 block-ELL matrix: three launches per iteration
    const int64_t G = (g->bell.nb + 255) / 256;
    HIPCHK(g->ctx, g->pcg_part.alloc((size_t)(3 * G) * sizeof(double)));
    HIPCHK(g->ctx, hipMemcpyAsync(sc + kRZ2 + 1, sc + kRZ, sizeof(double), hipMemcpyDeviceToDevice, s));
    for (int it = 1; it <= maxiter; ++it) {
      const bool f32 = g->coarse_fp32 && g->bell.has32;
      HIPCHK(g->ctx, pcg_fused_iteration(g->bell.nb, g->bell.W, g->bell.nbr.as<int32_t>(),
                                         f32 ? g->bell.v32.p : g->bell.v64.p, f32, nb.dofs.as<int32_t>(), nb.Binv.as<double>(), x, r, z, p, q, sc,
                                         kRZ2 + (it & 1), kRZ2 + ((it + 1) & 1), g->pcg_part.as<double>(), s));
      g->last_inner = it;
      if (it % 5 == 0 || it == maxiter) {
        double rr = 0.0;
        HIPCHK(g->ctx, hipMemcpyAsync(&rr, sc + kRR, sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHK(g->ctx, hipStreamSynchronize(s));
        if (!(std::sqrt(rr) > rtol * bn)) break;  // converged (or NaN: leave it to the outer check)
      }
    }
    g->total_inner += g->last_inner;
    return HFK_OK;
  }
  for (int it = 1; it <= maxiter; ++it) {
    HIPCHK(g->ctx, coarse_spmv(g, p, q, 1.0, 0.0, s));
    HIPCHK(g->ctx, vec_dots(n, p, q, nullptr, nullptr, nullptr, nullptr, g->dots.part.as<double>(), sc + kPQ, s));
    HIPCHK(g->ctx, cg_update(n, sc, p, q, x, r, s));
    if ((rc = precondition(r, z))) return rc;
    // sc[kRZN] = r.z, sc[kRZN+1] = r.r, sc[kRZN+2] = z.q (the flexible direction's projection)
    HIPCHK(g->ctx, vec_dots(n, r, z, r, r, flexible ? z : nullptr, flexible ? q : nullptr, g->dots.part.as<double>(),
                            sc + kRZN, s));
    g->last_inner = it;
    if (flexible || it % 5 == 0 || it == maxiter) {
      double rr[2];
      HIPCHK(g->ctx, hipMemcpyAsync(rr, sc + kRZN, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
      HIPCHK(g->ctx, hipStreamSynchronize(s));
      if (!(std::sqrt(rr[1]) > rtol * bn)) break;  // converged (or NaN: leave it to the outer check)
    }
    if (flexible) HIPCHK(g->ctx, hipMemcpyAsync(sc + kZQ, sc + kRZN + 2, sizeof(double), hipMemcpyDeviceToDevice, s));
    HIPCHK(g->ctx, cg_direction(n, sc, z, p, flexible ? 1 : 0, s));
  }
  g->total_inner += g->last_inner;
  return HFK_OK;
}

// vcycle! (src/pmultigrid.jl:316-329)
int vcycle(hfk_pmg* g, size_t l, double* x, const double* b, hipStream_t s) {
  hfk_cheb* S = g->cheb[l].get();
  int rc;
  if ((rc = cheb_smooth(S, x, b, true, g->r[l]->as<double>(), s))) return rc;
  if ((rc = do_restrict(g->P[l].get(), g->r[l]->as<double>(), g->bc[l]->as<double>(), s))) return rc;
  if (l + 2 == g->ops.size()) {
    if ((rc = coarse_solve(g, g->bc[l]->as<double>(), g->xc[l]->as<double>(), s, true, g->coarse_rtol,
                           g->coarse_maxiter)))
      return rc;
  } else {
    if ((rc = vcycle(g, l + 1, g->xc[l]->as<double>(), g->bc[l]->as<double>(), s))) return rc;
  }
  if ((rc = do_prolongate(g->P[l].get(), g->xc[l]->as<double>(), g->e[l]->as<double>(), s))) return rc;
  HIPCHK(g->ctx, vec_axpby(g->ops[l]->nfree, 1.0, g->e[l]->as<double>(), 1.0, x, s));
  return cheb_smooth(S, x, b, false, nullptr, s);
}

// update_levels! + smoother_update! (src/pmultigrid.jl:168-221, 299-314)
int pmg_update(hfk_pmg* g, hipStream_t s) {
  hfk_op* f = g->ops[0];
  if (f->storage < 0) return set_err(g->ctx, HFK_ESTATE, "HyperFEMKernels: p-multigrid update before prepare! of the working operator");
  g->nupdate++;
  int rc;
  bool ok = true;
  // the working level's smoother: its own tangents, or a copy at the same state in the smoother's storage
  const bool second = g->smoother_storage >= 0 && g->smoother_storage != f->storage;
  if (second) {
    const int st = g->smoother_storage;
    const size_t bytes = (size_t)f->nbatches * store_entries(st) * f->C * f->np * store_bytes(st);
    HIPCHK(g->ctx, f->tangents_s.alloc(bytes));
    if ((rc = op_prepare_into(f, st, f->tangents_s.p, f->state.as<double>(), s))) return rc;
    f->storage_s = st;
  }
  g->cheb[0]->secondary = second;
  for (size_t l = 1; l < g->ops.size(); ++l) {
    hfk_op* pf = g->ops[l - 1];
    hfk_op* pc = g->ops[l];
    if ((rc = do_inject(g->P[l - 1].get(), pf->state.as<double>(), pf->dvals.as<double>(), pc->state.as<double>(),
                        pc->dvals.as<double>(), s)))
      return rc;
    if (l + 1 < g->ops.size()) {
      if ((rc = op_run_prepare(pc, pc->state.as<double>(), g->storage, s))) return rc;
      if (g->lcsc[l] && (rc = level_csc_update(g, l, s))) return rc;
    } else {
      // the coarse state is injected at every update; the matrix is re-assembled and set up on the first update,
      // every coarse_every updates, and after one that was not finite (src/pmultigrid.jl:200-216)
      g->since++;
      if (!g->coarse_built || g->since >= g->coarse_every) {
        bool cok = false;
        if ((rc = coarse_setup(g, s, &cok))) return rc;
        g->coarse_ok = cok;
        g->ccheb_stale = true;
        g->coarse_built = cok;
        if (cok) {
          g->since = 0;
          g->coarse_builds++;
        }
      }
      ok = ok && g->coarse_ok;
    }
  }
  bool indef = false;
  for (auto& ch : g->cheb) {
    if ((rc = cheb_update(ch.get(), s))) return rc;
    ok = ok && ch->ok;
    // smoother_indefinite (src/smoothers.jl:48): finite blocks and Ritz values, a non-positive one
    if (!ch->ok && ch->blocks_ok && std::isfinite(ch->lmax) && std::isfinite(ch->lmin) && (ch->lmin <= 0 || ch->lmax <= 0))
      indef = true;
  }
  g->ok = ok;
  g->indefinite = !ok && indef;
  g->built = true;
  return HFK_OK;
}

// Stagnation watch of a Krylov solve preconditioned by a p-multigrid that carries a coarse network: when the
// residual norm has not fallen by fb_factor over the last fb_window iterations (or the preconditioned direction is
// not finite), the network is set aside for the rest of the solve and the plain cycle preconditions from there on,
// so the solve converges wherever the plain p-multigrid does.  The next solve starts with the network again.
struct StagnationWatch {
  hfk_pmg* g = nullptr;
  bool armed = false;
  std::vector<double> hist;
  StagnationWatch(hfk_pmg* g_, int precond) {
    if (precond == 2 && g_) {
      g = g_;
      g->net_suspended = false;
      armed = g->net && g->net_mode > 0 && g->fb_window > 0;
    }
  }
  ~StagnationWatch() {
    if (g) g->net_suspended = false;
  }
  bool fall_back() {
    if (!armed) return false;
    armed = false;
    g->net_suspended = true;
    g->fallbacks++;
    return true;
  }
  // true when the network was set aside at this residual norm
  bool step(double rn) {
    if (!armed) return false;
    hist.push_back(rn);
    const size_t w = (size_t)g->fb_window;
    if (hist.size() > w && !(rn <= g->fb_factor * hist[hist.size() - 1 - w])) return fall_back();
    return false;
  }
};

int pmg_apply(hfk_pmg* g, const double* r, double* z, hipStream_t s) {
  if (!g->ok) {  // ldiv! returns NaN: the Krylov solve reports a failure (src/pmultigrid.jl:331-335)
    HIPCHK(g->ctx, launch_fill(z, g->ops[0]->nfree, NAN, s));
    return HFK_OK;
  }
  g->ncycle++;
  return vcycle(g, 0, z, r, s);
}

// MatrixFreeSolver solve! (src/solvers.jl:111-158) as CG (kind 0) or flexible CG (kind 1) from x = 0 on device,
// preconditioned by nothing (0), the working operator's nodal blocks (1) or one p-multigrid V-cycle (2).
// Stops when ||r||_2 <= atol + rtol ||b||_2 (recurrence residual), or on a non-positive curvature p'Ap
// (indefinite), or at maxiter; the final decision uses the true residual b - A x.
struct KrylovWork {
  DevBuf r, z, p, q, sc;
  Dots dots;
};

int krylov(hfk_op* op, hfk_pmg* g, int kind, int precond, const double* b, double* x, double rtol, double atol,
           int maxiter, int* status, double* rnorm, hipStream_t s) {
  const int64_t n = op->nfree;
  KrylovWork w;
  int rc;
  if ((rc = vec_alloc(op->ctx, w.r, n)) || (rc = vec_alloc(op->ctx, w.z, n)) || (rc = vec_alloc(op->ctx, w.p, n)) ||
      (rc = vec_alloc(op->ctx, w.q, n)) || (rc = w.dots.init(op->ctx)))
    return rc;
  HIPCHK(op->ctx, w.sc.alloc(kNS * sizeof(double)));
  double *r = w.r.as<double>(), *z = w.z.as<double>(), *p = w.p.as<double>(), *q = w.q.as<double>(), *sc = w.sc.as<double>();
  status[0] = 0;
  status[1] = 0;
  status[2] = 0;
  if (precond == 2 && (!g || g->ops[0] != op)) return set_err(op->ctx, HFK_EARG, "HyperFEMKernels: the p-multigrid is not built on this operator");
  if (precond == 2 && !g->built) return set_err(op->ctx, HFK_ESTATE, "HyperFEMKernels: p-multigrid used before its update");
  if (precond == 2 && !g->ok) {  // src/solvers.jl:114-121: a failed preconditioner returns at once, not converged
    status[2] = g->indefinite ? 1 : 0;
    rnorm[0] = rnorm[1] = NAN;
    return HFK_OK;
  }
  auto M = [&](const double* rin, double* zout) -> int {
    if (precond == 0) {
      HIPCHK(op->ctx, vec_copy(n, rin, zout, s));
      return HFK_OK;
    }
    if (precond == 1) return blocks_apply(op, rin, zout, s);
    return pmg_apply(g, rin, zout, s);
  };
  HIPCHK(op->ctx, launch_fill(x, n, 0.0, s));
  HIPCHK(op->ctx, vec_copy(n, b, r, s));
  double h[3];
  if ((rc = w.dots.run(op->ctx, s, n, b, b, h))) return rc;
  const double bn = std::sqrt(h[0]);
  rnorm[1] = bn;
  if (bn == 0) {
    status[0] = 1;
    rnorm[0] = 0;
    return HFK_OK;
  }
  const double tol = atol + rtol * bn;
  StagnationWatch watch(g, precond);
  // p = z = M r and sc[kRZ] = r.z: the first direction, and a restart after the fallback to the plain cycle
  auto restart = [&]() -> int {
    int rc2;
    if ((rc2 = M(r, z))) return rc2;
    HIPCHK(op->ctx, vec_copy(n, z, p, s));
    HIPCHK(op->ctx, vec_dots(n, r, z, nullptr, nullptr, nullptr, nullptr, w.dots.part.as<double>(), sc + kRZ, s));
    return HFK_OK;
  };
  if ((rc = restart())) return rc;
  int it = 0;
  bool conv = false, indef = false;
  while (it < maxiter) {
    if ((rc = op_run_apply(op, p, q, s, 1.0, false))) return rc;
    double pq;
    HIPCHK(op->ctx, vec_dots(n, p, q, nullptr, nullptr, nullptr, nullptr, w.dots.part.as<double>(), sc + kPQ, s));
    HIPCHK(op->ctx, hipMemcpyAsync(&pq, sc + kPQ, sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(op->ctx, hipStreamSynchronize(s));
    if (!std::isfinite(pq) && watch.fall_back()) {  // the network's direction is not finite: x and r still are
      if ((rc = restart())) return rc;
      continue;
    }
    if (!(pq > 0)) {  // negative curvature (or NaN): stop, as cg with linesearch
      indef = std::isfinite(pq);
      break;
    }
    HIPCHK(op->ctx, cg_update(n, sc, p, q, x, r, s));
    ++it;
    double rr;
    if ((rc = w.dots.run(op->ctx, s, n, r, r, &rr))) return rc;
    if (!std::isfinite(rr)) break;
    if (std::sqrt(rr) <= tol) {
      conv = true;
      break;
    }
    if (watch.step(std::sqrt(rr))) {  // stagnation with the network: the plain cycle from here on
      if ((rc = restart())) return rc;
      continue;
    }
    if ((rc = M(r, z))) return rc;
    HIPCHK(op->ctx, vec_dots(n, r, z, z, q, nullptr, nullptr, w.dots.part.as<double>(), sc + kRZN, s));
    // kRZN and kZQ: vec_dots writes consecutive outputs; kZQ is read from sc[kRZN + 1]
    HIPCHK(op->ctx, hipMemcpyAsync(sc + kZQ, sc + kRZN + 1, sizeof(double), hipMemcpyDeviceToDevice, s));
    HIPCHK(op->ctx, cg_direction(n, sc, z, p, kind == 1 ? 1 : 0, s));
  }
  // true residual
  if ((rc = op_run_apply(op, x, r, s, 1.0, false))) return rc;
  HIPCHK(op->ctx, vec_axpby(n, 1.0, b, -1.0, r, s));
  double rt;
  if ((rc = w.dots.run(op->ctx, s, n, r, r, &rt))) return rc;
  rnorm[0] = std::sqrt(rt);
  // a stop on non-positive curvature is never converged (src/solvers.jl:141-146: converged = false, indefinite)
  status[0] = !indef && (conv || rnorm[0] <= tol) && std::isfinite(rnorm[0]) ? 1 : 0;
  status[1] = it;
  status[2] = indef ? 1 : 0;
  return HFK_OK;
}

// GMRES / FGMRES on the hexahedral operator, preconditioned by nothing (0), its nodal blocks (1) or one
// p-multigrid V-cycle (2)
int gmres(hfk_op* op, hfk_pmg* g, bool flexible, int precond, const double* b, double* x, double rtol, double atol,
          int maxiter, int restart, int* status, double* rnorm, hipStream_t s) {
  const int64_t n = op->nfree;
  auto A = [&](const double* v, double* y) -> int { return op_run_apply(op, v, y, s, 1.0, false); };
  auto M = [&](const double* rin, double* zout) -> int {
    if (precond == 0) {
      HIPCHK(op->ctx, vec_copy(n, rin, zout, s));
      return HFK_OK;
    }
    if (precond == 1) return blocks_apply(op, rin, zout, s);
    return pmg_apply(g, rin, zout, s);
  };
  StagnationWatch watch(g, precond);
  return gmres_core(op->ctx, n, A, M, watch, flexible, !flexible && precond == 2, b, x, rtol, atol, maxiter, restart,
                    status, rnorm, s);
}

}  // namespace

extern "C" {

// ---- BlockJacobi ----------------------------------------------------------------------------------------
int hfk_block_jacobi_update(hfk_op* op, int* ok) {
  int rc = op_bind(op);
  if (rc) return rc;
  bool b = false;
  if ((rc = blocks_update(op, op->tangents.p, op->storage, op->ctx->stream, &b))) return rc;
  if (ok) *ok = b ? 1 : 0;
  return HFK_OK;
}

int hfk_block_jacobi_apply_dev(hfk_op* op, const double* r_dev, double* y_dev, void* stream) {
  int rc = op_bind(op);
  if (rc) return rc;
  return blocks_apply(op, r_dev, y_dev, op->stream(stream));
}

int hfk_block_jacobi_apply(hfk_op* op, const double* r, double* y) {
  int rc = op_bind(op);
  if (rc) return rc;
  if (!r || !y) return set_err(op->ctx, HFK_EARG, "HyperFEMKernels: null vector");
  const int64_t n = op->nfree;
  DevBuf rd, yd;
  if ((rc = vec_alloc(op->ctx, rd, n)) || (rc = vec_alloc(op->ctx, yd, n))) return rc;
  hipStream_t s = op->ctx->stream;
  HIPCHK(op->ctx, hipMemcpyAsync(rd.p, r, n * sizeof(double), hipMemcpyHostToDevice, s));
  HIPCHK(op->ctx, launch_fill(yd.as<double>(), n, 0.0, s));
  if ((rc = blocks_apply(op, rd.as<double>(), yd.as<double>(), s))) return rc;
  HIPCHK(op->ctx, hipMemcpyAsync(y, yd.p, n * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(op->ctx, hipStreamSynchronize(s));
  return HFK_OK;
}

int hfk_block_jacobi_blocks(hfk_op* op, int64_t* nblocks, int32_t* dofs, double* Binv) {
  int rc = op_bind(op);
  if (rc) return rc;
  if ((rc = op_build_blocks(op))) return rc;
  if (nblocks) *nblocks = op->blocks.nblocks;
  if (dofs) std::memcpy(dofs, op->blocks.dofs_h.data(), op->blocks.dofs_h.size() * sizeof(int32_t));
  if (Binv) {  // the device stores entry-major [e][block]; the ABI returns [block][e]
    const int64_t nb = op->blocks.nblocks;
    std::vector<double> t((size_t)nb * 9);
    HIPCHK(op->ctx, hipMemcpy(t.data(), op->blocks.Binv.p, t.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int64_t k = 0; k < nb; ++k)
      for (int e = 0; e < 9; ++e) Binv[9 * k + e] = t[(size_t)e * nb + k];
  }
  return HFK_OK;
}

// ---- Chebyshev ------------------------------------------------------------------------------------------
int hfk_cheb_create(hfk_op* op, int degree, double range, double inflate, int eig_iters, hfk_cheb** out) {
  int rc = op_bind(op);
  if (rc) return rc;
  if (!out) return set_err(op->ctx, HFK_EARG, "HyperFEMKernels: null output pointer");
  if ((rc = check_cheb_args(op->ctx, degree, range, inflate, eig_iters))) return rc;
  std::unique_ptr<hfk_cheb> ch(new hfk_cheb());
  if ((rc = cheb_init(op, degree, range, inflate, eig_iters, ch.get()))) return rc;
  *out = ch.release();
  return HFK_OK;
}

int hfk_cheb_destroy(hfk_cheb* ch) {
  delete ch;
  return HFK_OK;
}

int hfk_cheb_update(hfk_cheb* ch, int* status, double* interval) {
  if (!ch) return set_err(nullptr, HFK_EARG, "HyperFEMKernels: null smoother");
  int rc = op_bind(ch->op);
  if (rc) return rc;
  if ((rc = cheb_update(ch, ch->op->ctx->stream))) return rc;
  if (status) {
    status[0] = ch->ok ? 1 : 0;
    status[1] = (!ch->ok && ch->blocks_ok && std::isfinite(ch->lmax) && std::isfinite(ch->lmin) &&
                 (ch->lmin <= 0 || ch->lmax <= 0)) ? 1 : 0;
  }
  if (interval) {
    interval[0] = ch->lmin;
    interval[1] = ch->lmax;
    interval[2] = ch->a;
    interval[3] = ch->b;
  }
  return HFK_OK;
}

int hfk_cheb_smooth_dev(hfk_cheb* ch, double* x, const double* rhs, int zero_guess, double* residual, void* stream) {
  if (!ch) return set_err(nullptr, HFK_EARG, "HyperFEMKernels: null smoother");
  int rc = op_bind(ch->op);
  if (rc) return rc;
  return cheb_smooth(ch, x, rhs, zero_guess != 0, residual, ch->op->stream(stream));
}

int hfk_cheb_smooth(hfk_cheb* ch, double* x, const double* rhs, int zero_guess, double* residual) {
  if (!ch || !x || !rhs) return set_err(ch ? ch->op->ctx : nullptr, HFK_EARG, "HyperFEMKernels: null argument");
  hfk_op* op = ch->op;
  int rc = op_bind(op);
  if (rc) return rc;
  const int64_t n = op->nfree;
  hipStream_t s = op->ctx->stream;
  DevBuf xd, bd, rd;
  if ((rc = vec_alloc(op->ctx, xd, n)) || (rc = vec_alloc(op->ctx, bd, n)) || (rc = vec_alloc(op->ctx, rd, n))) return rc;
  HIPCHK(op->ctx, hipMemcpyAsync(xd.p, x, n * sizeof(double), hipMemcpyHostToDevice, s));
  HIPCHK(op->ctx, hipMemcpyAsync(bd.p, rhs, n * sizeof(double), hipMemcpyHostToDevice, s));
  if ((rc = cheb_smooth(ch, xd.as<double>(), bd.as<double>(), zero_guess != 0, residual ? rd.as<double>() : nullptr, s))) return rc;
  HIPCHK(op->ctx, hipMemcpyAsync(x, xd.p, n * sizeof(double), hipMemcpyDeviceToHost, s));
  if (residual) HIPCHK(op->ctx, hipMemcpyAsync(residual, rd.p, n * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(op->ctx, hipStreamSynchronize(s));
  return HFK_OK;
}

int hfk_eigenvalue_bounds(hfk_op* op, int k, double* lmin, double* lmax) {
  int rc = op_bind(op);
  if (rc) return rc;
  if (k < 1) return set_err(op->ctx, HFK_EARG, "HyperFEMKernels: eigenvalue_bounds: k must be >= 1");
  hfk_cheb ch;
  if ((rc = cheb_init(op, 2, 20.0, 1.1, std::max(k, 2), &ch))) return rc;
  if (!op->blocks.built) {
    bool ok;
    if ((rc = blocks_update(op, op->tangents.p, op->storage, op->ctx->stream, &ok))) return rc;
  }
  return arnoldi_bounds(op, &ch, k, op->ctx->stream, lmin, lmax);
}

// ---- transfers ------------------------------------------------------------------------------------------
int hfk_transfer_create(hfk_op* coarse, hfk_op* fine, hfk_transfer** out) {
  if (!coarse || !fine || !out) return set_err(nullptr, HFK_EARG, "HyperFEMKernels: null argument");
  int rc = op_bind(fine);
  if (rc) return rc;
  std::unique_ptr<hfk_transfer> T(new hfk_transfer());
  if ((rc = transfer_build(coarse, fine, T.get()))) return rc;
  *out = T.release();
  return HFK_OK;
}

int hfk_transfer_destroy(hfk_transfer* T) {
  delete T;
  return HFK_OK;
}

namespace {
int transfer_host(hfk_transfer* T, int which, const double* a, const double* b, double* c, double* d) {
  if (!T) return set_err(nullptr, HFK_EARG, "HyperFEMKernels: null transfer");
  hfk_op* f = T->fine;
  hfk_op* co = T->coarse;
  int rc = op_bind(f);
  if (rc) return rc;
  hipStream_t s = f->ctx->stream;
  DevBuf A, B, Cb, Db;
  const int64_t nf = f->nfree, nc = co->nfree;
  if (which == 0) {  // prolongate: a = xc -> c = yf
    if ((rc = vec_alloc(f->ctx, A, nc)) || (rc = vec_alloc(f->ctx, Cb, nf))) return rc;
    HIPCHK(f->ctx, hipMemcpyAsync(A.p, a, nc * sizeof(double), hipMemcpyHostToDevice, s));
    if ((rc = do_prolongate(T, A.as<double>(), Cb.as<double>(), s))) return rc;
    HIPCHK(f->ctx, hipMemcpyAsync(c, Cb.p, nf * sizeof(double), hipMemcpyDeviceToHost, s));
  } else if (which == 1) {  // restrict: a = yf -> c = yc
    if ((rc = vec_alloc(f->ctx, A, nf)) || (rc = vec_alloc(f->ctx, Cb, nc))) return rc;
    HIPCHK(f->ctx, hipMemcpyAsync(A.p, a, nf * sizeof(double), hipMemcpyHostToDevice, s));
    if ((rc = do_restrict(T, A.as<double>(), Cb.as<double>(), s))) return rc;
    HIPCHK(f->ctx, hipMemcpyAsync(c, Cb.p, nc * sizeof(double), hipMemcpyDeviceToHost, s));
  } else {  // inject: (a = xf, b = df) -> (c = xc, d = dc)
    if ((rc = vec_alloc(f->ctx, A, nf)) || (rc = vec_alloc(f->ctx, B, f->ndir)) || (rc = vec_alloc(f->ctx, Cb, nc)) ||
        (rc = vec_alloc(f->ctx, Db, co->ndir)))
      return rc;
    HIPCHK(f->ctx, hipMemcpyAsync(A.p, a, nf * sizeof(double), hipMemcpyHostToDevice, s));
    if (f->ndir) HIPCHK(f->ctx, hipMemcpyAsync(B.p, b, f->ndir * sizeof(double), hipMemcpyHostToDevice, s));
    if ((rc = do_inject(T, A.as<double>(), B.as<double>(), Cb.as<double>(), Db.as<double>(), s))) return rc;
    HIPCHK(f->ctx, hipMemcpyAsync(c, Cb.p, nc * sizeof(double), hipMemcpyDeviceToHost, s));
    if (co->ndir) HIPCHK(f->ctx, hipMemcpyAsync(d, Db.p, co->ndir * sizeof(double), hipMemcpyDeviceToHost, s));
  }
  HIPCHK(f->ctx, hipStreamSynchronize(s));
  return HFK_OK;
}
}  // namespace

int hfk_prolongate(hfk_transfer* T, const double* xc, double* yf) { return transfer_host(T, 0, xc, nullptr, yf, nullptr); }
int hfk_restrict(hfk_transfer* T, const double* yf, double* yc) { return transfer_host(T, 1, yf, nullptr, yc, nullptr); }
int hfk_inject(hfk_transfer* T, const double* xf, const double* df, double* xc, double* dc) {
  return transfer_host(T, 2, xf, df, xc, dc);
}

// ---- p-multigrid ----------------------------------------------------------------------------------------
int hfk_pmg_create(hfk_op** levels, int nlevels, int degree, double range, double inflate, int eig_iters,
                   int smoother_storage, int coarse_kind, double coarse_rtol, int coarse_maxiter, hfk_pmg** out) {
  if (!levels || !out || nlevels < 2)
    return set_err(nullptr, HFK_EARG, "HyperFEMKernels: PMultigrid needs at least two levels");
  hfk_ctx* ctx = levels[0]->ctx;
  int rc = op_bind(levels[0]);
  if (rc) return rc;
  if ((rc = check_cheb_args(ctx, degree, range, inflate, eig_iters))) return rc;
  if ((coarse_kind != 0 && coarse_kind != 1) || !(coarse_rtol > 0) || coarse_maxiter < 1)
    return set_err(ctx, HFK_EARG, "HyperFEMKernels: PMultigrid: unsupported coarse solver settings");
  if (coarse_kind == 1 && levels[nlevels - 1] && levels[nlevels - 1]->nfree > kMaxDirect)
    return set_err(ctx, HFK_EDIM, "HyperFEMKernels: PMultigrid: the direct coarse solve takes at most " +
                                      std::to_string(kMaxDirect) + " coarse dofs (add coarser levels)");
  if (smoother_storage < -1 || smoother_storage > 7)
    return set_err(ctx, HFK_EARG, "HyperFEMKernels: PMultigrid: unknown smoother storage");
  std::unique_ptr<hfk_pmg> g(new hfk_pmg());
  g->ctx = ctx;
  g->smoother_storage = smoother_storage;
  g->coarse_kind = coarse_kind;
  g->coarse_rtol = coarse_rtol;
  g->coarse_maxiter = coarse_maxiter;
  for (int l = 0; l < nlevels; ++l) {
    if (!levels[l] || levels[l]->ctx != ctx) return set_err(ctx, HFK_EARG, "HyperFEMKernels: levels on different contexts");
    g->ops.push_back(levels[l]);
  }
  for (int l = 0; l + 1 < nlevels; ++l) {
    std::unique_ptr<hfk_transfer> T(new hfk_transfer());
    if ((rc = transfer_build(g->ops[l + 1], g->ops[l], T.get()))) return rc;
    g->P.push_back(std::move(T));
    std::unique_ptr<hfk_cheb> ch(new hfk_cheb());
    if ((rc = cheb_init(g->ops[l], degree, range, inflate, eig_iters, ch.get()))) return rc;
    g->cheb.push_back(std::move(ch));
    for (auto* v : {&g->r, &g->e, &g->bc, &g->xc}) v->push_back(std::unique_ptr<DevBuf>(new DevBuf()));
    if ((rc = vec_alloc(ctx, *g->r.back(), g->ops[l]->nfree)) || (rc = vec_alloc(ctx, *g->e.back(), g->ops[l]->nfree)) ||
        (rc = vec_alloc(ctx, *g->bc.back(), g->ops[l + 1]->nfree)) || (rc = vec_alloc(ctx, *g->xc.back(), g->ops[l + 1]->nfree)))
      return rc;
  }
  for (int l = 1; l < nlevels; ++l) HIPCHK(ctx, g->ops[l]->state.alloc((size_t)g->ops[l]->nfree * sizeof(double)));
  g->lcsc.resize(nlevels);
  for (int l = 1; l + 1 < nlevels; ++l)
    if (g->ops[l]->nfree <= kAssembledLevel) {
      g->lcsc[l].reset(new hfk_pmg::LevelCsc());
      g->cheb[l]->csr_tangent_blocks = true;
    }
  const int64_t nc = g->ops.back()->nfree;
  if ((rc = vec_alloc(ctx, g->cr, nc)) || (rc = vec_alloc(ctx, g->cz, nc)) || (rc = vec_alloc(ctx, g->cp, nc)) ||
      (rc = vec_alloc(ctx, g->cq, nc)) || (rc = vec_alloc(ctx, g->ct, nc)) || (rc = vec_alloc(ctx, g->cu, nc)) ||
      (rc = vec_alloc(ctx, g->cv, nc)) || (rc = vec_alloc(ctx, g->cdscale, nc)) || (rc = g->dots.init(ctx)))
    return rc;
  HIPCHK(ctx, g->cs.alloc(kNS * sizeof(double)));
  *out = g.release();
  return HFK_OK;
}

int hfk_pmg_destroy(hfk_pmg* g) {
  delete g;
  return HFK_OK;
}

int hfk_pmg_set_coarse_every(hfk_pmg* g, int coarse_every) {
  if (!g) return set_err(nullptr, HFK_EARG, "HyperFEMKernels: null p-multigrid");
  if (coarse_every < 1) return set_err(g->ctx, HFK_EARG, "HyperFEMKernels: PMultigrid: coarse_every must be >= 1");
  g->coarse_every = coarse_every;
  return HFK_OK;
}

int hfk_pmg_set_coarse_precision(hfk_pmg* g, int bits) {
  if (!g) return set_err(nullptr, HFK_EARG, "HyperFEMKernels: null p-multigrid");
  if (bits != 32 && bits != 64) return set_err(g->ctx, HFK_EARG, "HyperFEMKernels: coarse precision must be 32 or 64");
  if (g->bell.built && bits == 32 && !g->bell.has32)
    return set_err(g->ctx, HFK_ESTATE, "HyperFEMKernels: FP32 coarse values must be chosen before the first update");
  g->coarse_fp32 = bits == 32;
  return HFK_OK;
}

int hfk_pmg_coarse_builds(hfk_pmg* g, int64_t* builds) {
  if (!g || !builds) return set_err(g ? g->ctx : nullptr, HFK_EARG, "HyperFEMKernels: null argument");
  *builds = g->coarse_builds;
  return HFK_OK;
}

int hfk_pmg_update(hfk_pmg* g, int* status) {
  if (!g) return set_err(nullptr, HFK_EARG, "HyperFEMKernels: null p-multigrid");
  int rc = op_bind(g->ops[0]);
  if (rc) return rc;
  g->storage = g->smoother_storage >= 0 ? g->smoother_storage : (g->ops[0]->storage < 0 ? 0 : g->ops[0]->storage);
  if ((rc = pmg_update(g, g->ctx->stream))) return rc;
  if (status) {
    status[0] = g->ok ? 1 : 0;
    status[1] = g->indefinite ? 1 : 0;
  }
  return HFK_OK;
}

int hfk_pmg_vcycle_dev(hfk_pmg* g, const double* r_dev, double* z_dev, void* stream) {
  if (!g) return set_err(nullptr, HFK_EARG, "HyperFEMKernels: null p-multigrid");
  int rc = op_bind(g->ops[0]);
  if (rc) return rc;
  if (!g->built) return set_err(g->ctx, HFK_ESTATE, "HyperFEMKernels: p-multigrid used before its update");
  return pmg_apply(g, r_dev, z_dev, g->ops[0]->stream(stream));
}

int hfk_pmg_vcycle(hfk_pmg* g, const double* r, double* z) {
  if (!g || !r || !z) return set_err(g ? g->ctx : nullptr, HFK_EARG, "HyperFEMKernels: null argument");
  int rc = op_bind(g->ops[0]);
  if (rc) return rc;
  if (!g->built) return set_err(g->ctx, HFK_ESTATE, "HyperFEMKernels: p-multigrid used before its update");
  const int64_t n = g->ops[0]->nfree;
  hipStream_t s = g->ctx->stream;
  DevBuf rd, zd;
  if ((rc = vec_alloc(g->ctx, rd, n)) || (rc = vec_alloc(g->ctx, zd, n))) return rc;
  HIPCHK(g->ctx, hipMemcpyAsync(rd.p, r, n * sizeof(double), hipMemcpyHostToDevice, s));
  if ((rc = pmg_apply(g, rd.as<double>(), zd.as<double>(), s))) return rc;
  HIPCHK(g->ctx, hipMemcpyAsync(z, zd.p, n * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(g->ctx, hipStreamSynchronize(s));
  return HFK_OK;
}

int hfk_pmg_info(hfk_pmg* g, double* info) {
  if (!g || !info) return set_err(g ? g->ctx : nullptr, HFK_EARG, "HyperFEMKernels: null argument");
  info[0] = (double)g->ops.size();
  info[1] = g->ok ? 1.0 : 0.0;
  info[2] = (double)g->ncycle;
  info[3] = (double)g->last_inner;
  info[4] = (double)g->total_inner;
  info[5] = g->coarse_ok ? 1.0 : 0.0;
  for (size_t l = 0; l < g->cheb.size(); ++l) {
    double* o = info + 6 + 6 * l;
    o[0] = g->cheb[l]->lmin;
    o[1] = g->cheb[l]->lmax;
    o[2] = g->cheb[l]->a;
    o[3] = g->cheb[l]->b;
    o[4] = g->cheb[l]->ok ? 1.0 : 0.0;
    o[5] = g->cheb[l]->blocks_ok ? 1.0 : 0.0;
  }
  return HFK_OK;
}

int hfk_pmg_coarse_matvec(hfk_pmg* g, const double* x, double* y, int scalar_csc) {
  if (!g || !x || !y) return set_err(g ? g->ctx : nullptr, HFK_EARG, "HyperFEMKernels: null argument");
  if (!g->built || !g->colptr.p) return set_err(g->ctx, HFK_ESTATE, "HyperFEMKernels: p-multigrid not updated");
  int rc = op_bind(g->ops[0]);
  if (rc) return rc;
  hipStream_t s = g->ctx->stream;
  const int64_t n = g->ops.back()->nfree;
  DevBuf xd, yd;
  if ((rc = vec_alloc(g->ctx, xd, n)) || (rc = vec_alloc(g->ctx, yd, n))) return rc;
  HIPCHK(g->ctx, hipMemcpyAsync(xd.p, x, n * sizeof(double), hipMemcpyHostToDevice, s));
  if (scalar_csc)
    HIPCHK(g->ctx, csc_t_spmv(n, g->colptr.as<int32_t>(), g->rowval.as<int32_t>(), g->nzval.as<double>(), xd.as<double>(),
                              yd.as<double>(), 1.0, 0.0, s));
  else
    HIPCHK(g->ctx, coarse_spmv(g, xd.as<double>(), yd.as<double>(), 1.0, 0.0, s));
  HIPCHK(g->ctx, hipMemcpyAsync(y, yd.p, n * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(g->ctx, hipStreamSynchronize(s));
  return HFK_OK;
}

int hfk_pmg_set_fallback(hfk_pmg* g, int window, double factor) {
  if (!g) return set_err(nullptr, HFK_EARG, "HyperFEMKernels: null p-multigrid");
  if (window < 0 || !(factor > 0 && factor < 1))
    return set_err(g->ctx, HFK_EARG, "HyperFEMKernels: fallback window must be >= 0 and factor in (0, 1)");
  g->fb_window = window;
  g->fb_factor = factor;
  return HFK_OK;
}

int hfk_pmg_fallbacks(hfk_pmg* g, int64_t* count) {
  if (!g || !count) return set_err(g ? g->ctx : nullptr, HFK_EARG, "HyperFEMKernels: null argument");
  *count = g->fallbacks;
  return HFK_OK;
}

int hfk_pmg_set_coarse_network(hfk_pmg* g, hfk_net* net, int mode, int smoother_degree, int scaled) {
  if (!g) return set_err(nullptr, HFK_EARG, "HyperFEMKernels: null p-multigrid");
  if (scaled != 0 && scaled != 1) return set_err(g->ctx, HFK_EARG, "HyperFEMKernels: scaled must be 0 or 1");
  g->net_scaled = scaled == 1;
  if (smoother_degree < 0) return set_err(g->ctx, HFK_EARG, "HyperFEMKernels: smoother degree must be >= 0");
  if (mode < 0 || mode > 4) return set_err(g->ctx, HFK_EARG, "HyperFEMKernels: coarse network mode must be 0 to 4");
  if (mode > 0 && (!net || net_ncoarse(net) != g->ops.back()->nfree))
    return set_err(g->ctx, HFK_EDIM, "HyperFEMKernels: the network is not built for this coarse level");
  g->net = mode > 0 ? net : nullptr;
  g->net_mode = mode;
  const int deg = smoother_degree > 0 ? smoother_degree : g->cheb[0]->degree;
  if (g->ccheb && g->ccheb->degree != deg) g->ccheb->degree = deg;  // the interval does not depend on the degree
  g->ccheb_degree = smoother_degree;
  return HFK_OK;
}

// Training pairs from the plain cycle (two-level view: the working level and the assembled coarse level): the
// pre-smoothed residual of b_k restricted to the coarse level and its coarse solution.
int hfk_pmg_coarse_pairs_dev(hfk_pmg* g, const double* b_dev, int nsamples, double* bc_dev, double* ec_dev,
                             int coarse_smoothed, double rtol, int maxiter, void* stream) {
  if (!g || !b_dev || !bc_dev || nsamples < 1 || !(rtol > 0) || maxiter < 1)
    return set_err(g ? g->ctx : nullptr, HFK_EARG, "HyperFEMKernels: invalid arguments");
  if (g->ops.size() != 2) return set_err(g->ctx, HFK_ESTATE, "HyperFEMKernels: coarse pairs need a two-level hierarchy");
  if (!g->built || !g->ok) return set_err(g->ctx, HFK_ESTATE, "HyperFEMKernels: p-multigrid not updated or not ok");
  int rc = op_bind(g->ops[0]);
  if (rc) return rc;
  hipStream_t s = g->ops[0]->stream(stream);
  const int64_t nf = g->ops[0]->nfree, nc = g->ops[1]->nfree;
  DevBuf xw;
  if ((rc = vec_alloc(g->ctx, xw, nf))) return rc;
  if (coarse_smoothed) {
    bool usable = false;
    if ((rc = coarse_smoother(g, s, &usable))) return rc;
    if (!usable) return set_err(g->ctx, HFK_ESTATE, "HyperFEMKernels: the coarse smoother has no usable interval");
  }
  for (int k = 0; k < nsamples; ++k) {
    const double* b = b_dev + (size_t)k * nf;
    double* bc = bc_dev + (size_t)k * nc;
    if ((rc = cheb_smooth(g->cheb[0].get(), xw.as<double>(), b, true, g->r[0]->as<double>(), s))) return rc;
    if ((rc = do_restrict(g->P[0].get(), g->r[0]->as<double>(), bc, s))) return rc;
    if (coarse_smoothed) {  // the network cycle's input: the residual of the coarse pre-smoothing
      if ((rc = cheb_smooth(g->ccheb.get(), g->cu.as<double>(), bc, true, g->ct.as<double>(), s))) return rc;
      HIPCHK(g->ctx, vec_copy(nc, g->ct.as<double>(), bc, s));
    }
    if (ec_dev && (rc = coarse_solve(g, bc, ec_dev + (size_t)k * nc, s, false, rtol, maxiter))) return rc;
  }
  HIPCHK(g->ctx, hipStreamSynchronize(s));
  return HFK_OK;
}

// Energy training of the coarse network on coarse right-hand sides r_k (device, [k][nc]): minimise
// L(w) = w'(A/h) w - 2 w'x, x = r/s, whose minimiser is the coarse solve in the network's units (z = (s/h) w
// = A^-1 r); in the scaled mode A/h -> D^-1/2 A D^-1/2 and r -> D^-1/2 r (pmg_net_apply).  Needs one product
// with the assembled coarse matrix per sample and no exact solves.
// one sample of the energy loss: w = N(r), q = (A/h) w - r/s, the weight gradient of L = (w'(A/h)w - 2 w'x)/n
// accumulated into the network's gradient; *L receives the sample's loss
int energy_sample(hfk_pmg* g, hfk_net* net, const double* r, DevBuf& w, DevBuf& q, double* L, hipStream_t s) {
  const int64_t n = g->ops.back()->nfree;
  const double invh = g->net_scaled ? 1.0 : 1.0 / net_h(net);  // the loss's operator: D^-1/2 A D^-1/2 or A/h
  int rc;
  if (g->net_scaled) {  // train on D^-1/2 A D^-1/2 and D^-1/2 r
    HIPCHK(g->ctx, vec_mul(n, g->cdscale.as<double>(), r, g->cv.as<double>(), s));
    r = g->cv.as<double>();
  }
  if ((rc = net_forward_vec(net, r, w.as<double>(), s))) return rc;
  const double* wa = w.as<double>();
  if (g->net_scaled) {
    HIPCHK(g->ctx, vec_mul(n, g->cdscale.as<double>(), w.as<double>(), g->cu.as<double>(), s));
    wa = g->cu.as<double>();
  }
  HIPCHK(g->ctx, coarse_spmv(g, wa, q.as<double>(), invh, 0.0, s));
  if (g->net_scaled) HIPCHK(g->ctx, vec_mul(n, g->cdscale.as<double>(), q.as<double>(), q.as<double>(), s));
  double h3[3];
  if ((rc = g->dots.run(g->ctx, s, n, w.as<double>(), q.as<double>(), h3, w.as<double>(), r, r, r))) return rc;
  const double sc = std::sqrt(h3[2] / (double)n);
  const double sinv = sc > 0 ? 1.0 / sc : 1.0;
  *L = (h3[0] - 2.0 * h3[1] * sinv) / (double)n;
  HIPCHK(g->ctx, vec_axpby(n, -sinv, r, 1.0, q.as<double>(), s));  // q = (A/h) w - x
  return net_backward_vec(net, q.as<double>(), 2.0 / (double)n, s);
}

int hfk_pmg_train_coarse_network(hfk_pmg* g, hfk_net* net, const double* bc_dev, int nsamples, int steps, int batch,
                                 double lr, uint64_t seed, double* history, void* stream) {
  if (!g || !net || !bc_dev || nsamples < 1 || steps < 0 || batch < 1 || !(lr > 0))
    return set_err(g ? g->ctx : nullptr, HFK_EARG, "HyperFEMKernels: invalid training arguments");
  if (!g->built || !g->colptr.p) return set_err(g->ctx, HFK_ESTATE, "HyperFEMKernels: p-multigrid not updated");
  const int64_t n = g->ops.back()->nfree;
  if (net_ncoarse(net) != n) return set_err(g->ctx, HFK_EDIM, "HyperFEMKernels: the network is not built for this coarse level");
  int rc = op_bind(g->ops[0]);
  if (rc) return rc;
  hipStream_t s = g->ops[0]->stream(stream);
  DevBuf w, q;
  if ((rc = vec_alloc(g->ctx, w, n)) || (rc = vec_alloc(g->ctx, q, n))) return rc;
  uint64_t st = seed * 6364136223846793005ull + 1442695040888963407ull;
  for (int step = 0; step < steps; ++step) {
    double Lsum = 0;
    for (int b = 0; b < batch; ++b) {
      st = st * 6364136223846793005ull + 1442695040888963407ull;
      const int k = (int)((st >> 33) % (uint64_t)nsamples);
      double L;
      if ((rc = energy_sample(g, net, bc_dev + (size_t)k * n, w, q, &L, s))) return rc;
      Lsum += L;
    }
    if ((rc = net_adam(net, lr, batch, s))) return rc;
    if (history) history[step] = Lsum / batch;
  }
  HIPCHK(g->ctx, hipStreamSynchronize(s));
  return HFK_OK;
}

// The energy loss summed over the samples bc_k (device) and its weight gradient (host, FP32), without a training
// step: the quantity hfk_pmg_train_coarse_network descends, for checking it against another implementation.
int hfk_pmg_coarse_network_energy_gradient(hfk_pmg* g, hfk_net* net, const double* bc_dev, int nsamples, float* grad,
                                           double* loss, void* stream) {
  if (!g || !net || !bc_dev || nsamples < 1)
    return set_err(g ? g->ctx : nullptr, HFK_EARG, "HyperFEMKernels: invalid arguments");
  if (!g->built || !g->colptr.p) return set_err(g->ctx, HFK_ESTATE, "HyperFEMKernels: p-multigrid not updated");
  const int64_t n = g->ops.back()->nfree;
  if (net_ncoarse(net) != n) return set_err(g->ctx, HFK_EDIM, "HyperFEMKernels: the network is not built for this coarse level");
  int rc = op_bind(g->ops[0]);
  if (rc) return rc;
  hipStream_t s = g->ops[0]->stream(stream);
  DevBuf w, q;
  if ((rc = vec_alloc(g->ctx, w, n)) || (rc = vec_alloc(g->ctx, q, n))) return rc;
  if ((rc = net_zero_grad(net, s))) return rc;
  double total = 0;
  for (int k = 0; k < nsamples; ++k) {
    double L;
    if ((rc = energy_sample(g, net, bc_dev + (size_t)k * n, w, q, &L, s))) return rc;
    total += L;
  }
  if (grad && (rc = net_copy_grad(net, grad, s))) return rc;
  HIPCHK(g->ctx, hipStreamSynchronize(s));
  if (loss) *loss = total;
  return HFK_OK;
}

// Relative energy-norm error of the network's coarse correction against the coarse solve to 1e-10, averaged over
// the right-hand sides r_k (device): sqrt((z - e)' A (z - e) / e' A e)
int hfk_pmg_coarse_network_error(hfk_pmg* g, hfk_net* net, const double* bc_dev, int nsamples, double* err,
                                 void* stream) {
  if (!g || !net || !bc_dev || nsamples < 1 || !err)
    return set_err(g ? g->ctx : nullptr, HFK_EARG, "HyperFEMKernels: invalid arguments");
  const int64_t n = g->ops.back()->nfree;
  if (net_ncoarse(net) != n) return set_err(g->ctx, HFK_EDIM, "HyperFEMKernels: the network is not built for this coarse level");
  int rc = op_bind(g->ops[0]);
  if (rc) return rc;
  hipStream_t s = g->ops[0]->stream(stream);
  DevBuf e, z, q;
  if ((rc = vec_alloc(g->ctx, e, n)) || (rc = vec_alloc(g->ctx, z, n)) || (rc = vec_alloc(g->ctx, q, n))) return rc;
  double total = 0;
  for (int k = 0; k < nsamples; ++k) {
    const double* r = bc_dev + (size_t)k * n;
    if ((rc = coarse_solve(g, r, e.as<double>(), s, false, 1e-10, 20000))) return rc;
    if ((rc = pmg_net_apply(g, net, r, z.as<double>(), s))) return rc;
    HIPCHK(g->ctx, vec_axpby(n, -1.0, e.as<double>(), 1.0, z.as<double>(), s));  // z - e
    double h2[2];
    HIPCHK(g->ctx, coarse_spmv(g, z.as<double>(), q.as<double>(), 1.0, 0.0, s));
    if ((rc = g->dots.run(g->ctx, s, n, z.as<double>(), q.as<double>(), h2, e.as<double>(), r))) return rc;
    total += std::sqrt(std::max(h2[0], 0.0) / std::max(h2[1], 1e-300));
  }
  *err = total / nsamples;
  return HFK_OK;
}

// ---- Krylov ---------------------------------------------------------------------------------------------
int hfk_krylov_solve_dev(hfk_op* op, hfk_pmg* pmg, int kind, int precond, const double* b_dev, double* x_dev, double rtol,
                         double atol, int maxiter, int restart, int* status, double* rnorm, void* stream) {
  int rc = op_bind(op);
  if (rc) return rc;
  if (!b_dev || !x_dev || !status || !rnorm) return set_err(op->ctx, HFK_EARG, "HyperFEMKernels: null argument");
  if (kind < 0 || kind > 3 || precond < 0 || precond > 2 || maxiter < 0 || rtol < 0 || atol < 0)
    return set_err(op->ctx, HFK_EARG, "HyperFEMKernels: MatrixFreeSolver: invalid settings");
  if (kind >= 2 && (restart < 1 || restart > 1024))
    return set_err(op->ctx, HFK_EARG, "HyperFEMKernels: MatrixFreeSolver: restart must be in 1..1024");
  if (op->storage < 0) return set_err(op->ctx, HFK_ESTATE, "HyperFEMKernels: solve! before prepare!");
  if (precond == 1 && !op->blocks.built) {
    bool ok;
    if ((rc = blocks_update(op, op->tangents.p, op->storage, op->stream(stream), &ok))) return rc;
  }
  if (kind >= 2) {
    if (precond == 2 && (!pmg || pmg->ops[0] != op))
      return set_err(op->ctx, HFK_EARG, "HyperFEMKernels: the p-multigrid is not built on this operator");
    if (precond == 2 && !pmg->built) return set_err(op->ctx, HFK_ESTATE, "HyperFEMKernels: p-multigrid used before its update");
    if (precond == 2 && !pmg->ok) {  // src/solvers.jl:114-121
      status[0] = status[1] = 0;
      status[2] = pmg->indefinite ? 1 : 0;
      rnorm[0] = rnorm[1] = NAN;
      return HFK_OK;
    }
    return gmres(op, pmg, kind == 3, precond, b_dev, x_dev, rtol, atol, maxiter, restart, status, rnorm,
                 op->stream(stream));
  }
  return krylov(op, pmg, kind, precond, b_dev, x_dev, rtol, atol, maxiter, status, rnorm, op->stream(stream));
}

int hfk_krylov_solve(hfk_op* op, hfk_pmg* pmg, int kind, int precond, const double* b, double* x, double rtol,
                     double atol, int maxiter, int restart, int* status, double* rnorm) {
  int rc = op_bind(op);
  if (rc) return rc;
  if (!b || !x) return set_err(op->ctx, HFK_EARG, "HyperFEMKernels: null vector");
  const int64_t n = op->nfree;
  hipStream_t s = op->ctx->stream;
  DevBuf bd, xd;
  if ((rc = vec_alloc(op->ctx, bd, n)) || (rc = vec_alloc(op->ctx, xd, n))) return rc;
  HIPCHK(op->ctx, hipMemcpyAsync(bd.p, b, n * sizeof(double), hipMemcpyHostToDevice, s));
  if ((rc = hfk_krylov_solve_dev(op, pmg, kind, precond, bd.as<double>(), xd.as<double>(), rtol, atol, maxiter, restart,
                                 status, rnorm, nullptr)))
    return rc;
  HIPCHK(op->ctx, hipMemcpyAsync(x, xd.p, n * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(op->ctx, hipStreamSynchronize(s));
  return HFK_OK;
}

}  // extern "C"
