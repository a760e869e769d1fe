// GMRES / FGMRES on the device over any operator and preconditioner, and the small helpers of the Krylov solves
// (deterministic dot products read back to the host, vector allocation).  Internal; not part of the C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "hfk_op.hpp"
#include "hfk_solver.hpp"

namespace hfk {

inline int vec_alloc(hfk_ctx* ctx, DevBuf& b, int64_t n) {
  HIPCHK(ctx, b.alloc((size_t)std::max<int64_t>(n, 1) * sizeof(double)));
  return HFK_OK;
}

struct Dots {
  DevBuf part, out;
  int init(hfk_ctx* ctx) {
    HIPCHK(ctx, part.alloc(dot_partial_doubles() * sizeof(double)));
    HIPCHK(ctx, out.alloc(4 * sizeof(double)));
    return HFK_OK;
  }
  // blocking: computes up to 3 dots and reads them back
  int run(hfk_ctx* ctx, hipStream_t s, int64_t n, const double* a0, const double* b0, double* res,
          const double* a1 = nullptr, const double* b1 = nullptr, const double* a2 = nullptr, const double* b2 = nullptr) {
    HIPCHK(ctx, vec_dots(n, a0, b0, a1, b1, a2, b2, part.as<double>(), out.as<double>(), s));
    const int nd = a2 ? 3 : (a1 ? 2 : 1);
    HIPCHK(ctx, hipMemcpyAsync(res, out.p, nd * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipStreamSynchronize(s));
    return HFK_OK;
  }
};

// a Watch that never sets the preconditioner aside
struct NoWatch {
  bool fall_back() { return false; }
  bool step(double) { return false; }
};

// GMRES / FGMRES over any operator A (y = A x on the device) and preconditioner M (z = M^-1 r), both returning a
// status; Watch is the learned cycle's stagnation watch (fall_back(), step(rn)) or a no-op.
// Krylov.jl's gmres! / fgmres! as MatrixFreeSolver calls them (src/solvers.jl:128-133): right preconditioning,
// restarted every `restart` iterations, converged when the residual estimate |g_{j+1}| <= atol + rtol ||b|| (the
// true residual of right preconditioning).  Arnoldi with classical Gram-Schmidt done twice (two batched
// projections per iteration), Givens rotations on the host.  GMRES stores the basis V and forms
// x += M^-1 (V y) at each restart; FGMRES stores Z_j = M^-1 v_j as well (a variable preconditioner, e.g. the
// p-multigrid with an inexact coarse solve) and forms x += Z y.  When the method stops without its test (iteration
// limit, breakdown), the decision is taken on the true residual (src/solvers.jl:149-157).
template <class Apply, class Precond, class Watch>
int gmres_core(hfk_ctx* ctx, int64_t n, Apply A, Precond M, Watch& watch, bool flexible, bool variable, const double* b,
               double* x, double rtol, double atol, int maxiter, int restart, int* status, double* rnorm, hipStream_t s) {
  const int m = restart;
  DevBuf V, Z, t, r, hd, part;
  Dots dots;
  int rc;
  if ((rc = dots.init(ctx))) return rc;
  HIPCHK(ctx, V.alloc((size_t)(m + 1) * n * sizeof(double)));
  if (flexible) HIPCHK(ctx, Z.alloc((size_t)m * n * sizeof(double)));
  if ((rc = vec_alloc(ctx, t, n)) || (rc = vec_alloc(ctx, r, n))) return rc;
  HIPCHK(ctx, hd.alloc((size_t)(m + 1) * sizeof(double)));
  HIPCHK(ctx, part.alloc(multi_dot_partial_doubles(m + 1) * sizeof(double)));
  double* Vp = V.as<double>();
  double* Zp = Z.as<double>();
  double* tp = t.as<double>();
  double* rp = r.as<double>();
  double* hp = hd.as<double>();
  status[0] = status[1] = status[2] = 0;
  HIPCHK(ctx, launch_fill(x, n, 0.0, s));
  double h[1];
  if ((rc = dots.run(ctx, s, n, b, b, h))) return rc;
  const double bn = std::sqrt(h[0]);
  rnorm[1] = bn;
  if (bn == 0) {
    status[0] = 1;
    rnorm[0] = 0;
    return HFK_OK;
  }
  const double tol = atol + rtol * bn;
  std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), gv(m + 1), y(m), hcol(m + 1);
  int it = 0;
  bool solved = false, stop = false;
  HIPCHK(ctx, vec_copy(n, b, rp, s));  // r = b - A x, x = 0
  double beta = bn;
  while (!stop && it < maxiter) {
    HIPCHK(ctx, vec_axpby(n, 1.0 / beta, rp, 0.0, Vp, s));  // v_0 = r / beta
    std::fill(gv.begin(), gv.end(), 0.0);
    gv[0] = beta;
    int k = 0;  // columns of this cycle
    for (int j = 0; j < m && it < maxiter; ++j) {
      double* vj = Vp + (size_t)j * n;
      double* w = Vp + (size_t)(j + 1) * n;
      double* zj = flexible ? Zp + (size_t)j * n : tp;
      if ((rc = M(vj, zj))) return rc;
      if ((rc = A(zj, w))) return rc;
      // two passes of classical Gram-Schmidt against v_0..v_j
      std::fill(hcol.begin(), hcol.end(), 0.0);
      for (int pass = 0; pass < 2; ++pass) {
        HIPCHK(ctx, multi_dot(n, Vp, n, j + 1, w, part.as<double>(), hp, s));
        HIPCHK(ctx, multi_axpy(n, Vp, n, j + 1, hp, -1.0, w, s));
        std::vector<double> hh(j + 1);
        HIPCHK(ctx, hipMemcpyAsync(hh.data(), hp, (j + 1) * sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHK(ctx, hipStreamSynchronize(s));
        for (int i = 0; i <= j; ++i) hcol[i] += hh[i];
      }
      double ww;
      if ((rc = dots.run(ctx, s, n, w, w, &ww))) return rc;
      const double hn = std::sqrt(ww);
      if (!std::isfinite(hn) && watch.fall_back()) {  // the network's column is not finite: redo it plainly
        --j;
        continue;
      }
      ++it;
      if (!std::isfinite(hn)) {
        stop = true;  // NaN: keep the columns before it, leave the decision to the true residual
        break;
      }
      k = j + 1;
      // apply the previous rotations, then the new one
      for (int i = 0; i < j; ++i) {
        const double a0 = hcol[i], a1 = hcol[i + 1];
        hcol[i] = cs[i] * a0 + sn[i] * a1;
        hcol[i + 1] = -sn[i] * a0 + cs[i] * a1;
      }
      const double den = std::hypot(hcol[j], hn);
      cs[j] = den > 0 ? hcol[j] / den : 1.0;
      sn[j] = den > 0 ? hn / den : 0.0;
      hcol[j] = den;
      gv[j + 1] = -sn[j] * gv[j];
      gv[j] = cs[j] * gv[j];
      for (int i = 0; i <= j; ++i) H[(size_t)i * m + j] = hcol[i];
      const double est = std::fabs(gv[j + 1]);
      if (est <= tol) {
        // GMRES with a p-multigrid preconditioner: the cycle (iterative coarse solve) differs from one
        // application to the next, so x += M^-1 (V y) does not realise the estimate; the restart below decides
        // on the true residual and continues when it is not met
        if (!variable) {
          solved = true;
          stop = true;
        }
        break;
      }
      if (!(hn > 1e-14 * den)) {  // breakdown: the Krylov space holds the solution
        stop = true;
        break;
      }
      HIPCHK(ctx, vec_axpby(n, 0.0, w, 1.0 / hn, w, s));  // v_{j+1} = w / hn
      // stagnation with the network: the plain cycle for the rest of the solve (FGMRES keeps Z_j of every
      // preconditioner; GMRES decides on the true residual at the restart)
      (void)watch.step(est);
    }
    if (k == 0) break;
    // y = R^-1 g (upper triangular k x k), then x += Z y (FGMRES) or x += M^-1 (V y) (GMRES)
    for (int i = k - 1; i >= 0; --i) {
      double v = gv[i];
      for (int l = i + 1; l < k; ++l) v -= H[(size_t)i * m + l] * y[l];
      y[i] = v / H[(size_t)i * m + i];
    }
    HIPCHK(ctx, hipMemcpyAsync(hp, y.data(), k * sizeof(double), hipMemcpyHostToDevice, s));
    if (flexible) {
      HIPCHK(ctx, multi_axpy(n, Zp, n, k, hp, 1.0, x, s));
    } else {
      HIPCHK(ctx, launch_fill(rp, n, 0.0, s));
      HIPCHK(ctx, multi_axpy(n, Vp, n, k, hp, 1.0, rp, s));
      if ((rc = M(rp, tp))) return rc;
      HIPCHK(ctx, vec_axpby(n, 1.0, tp, 1.0, x, s));
    }
    HIPCHK(ctx, hipStreamSynchronize(s));  // y is read from host memory by the copy above
    if (stop) break;
    // restart from the true residual
    if ((rc = A(x, rp))) return rc;
    HIPCHK(ctx, vec_axpby(n, 1.0, b, -1.0, rp, s));
    double rr;
    if ((rc = dots.run(ctx, s, n, rp, rp, &rr))) return rc;
    beta = std::sqrt(rr);
    if (!(beta > tol)) {
      solved = std::isfinite(beta);
      break;
    }
  }
  // the true residual
  if ((rc = A(x, rp))) return rc;
  HIPCHK(ctx, vec_axpby(n, 1.0, b, -1.0, rp, s));
  double rt;
  if ((rc = dots.run(ctx, s, n, rp, rp, &rt))) return rc;
  rnorm[0] = std::sqrt(rt);
  status[0] = (solved || rnorm[0] <= tol) && std::isfinite(rnorm[0]) ? 1 : 0;
  if (variable) status[0] = rnorm[0] <= tol ? 1 : 0;  // only the true residual counts
  status[1] = it;
  status[2] = 0;
  return HFK_OK;
}

}  // namespace hfk
