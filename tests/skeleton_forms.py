"""Face point forms given as HIP source for the skeleton and boundary-form tests.

NONLINEAR_FACET_SOURCE is the device twin of oracle/skeleton.py's NonlinearFacet: energy
e = gamma h (|J|^2 / 2 + kappa |J|^4 / 4) + beta |J n|^2 / 2 + alpha |um|^2 |J|^2 / 2 with J = grad u+ - grad u-
(J[3c + K] = du_c/dx_K jump), um = (u+ + u-) / 2 and p = (gamma, kappa, beta, alpha); residual = de/ds, jvp its
derivative.  NONLINEAR_BOUNDARY_SOURCE is the twin of NonlinearBoundary: e = k |u|^2 / 2 + kappa |u|^4 / 4 +
beta |grad u n|^2 / 2 + gamma x_0 |u|^2 / 2, p = (k, kappa, beta, gamma).
"""

NONLINEAR_FACET_SOURCE = r"""
struct NonlinearFacet {
  __device__ static void parts(const double* s, double* J, double* um) {
    for (int k = 0; k < 9; ++k) J[k] = s[3 + k] - s[15 + k];
    for (int c = 0; c < 3; ++c) um[c] = 0.5 * (s[c] + s[12 + c]);
  }
  __device__ static double energy(const double* s, const double* x, const double* n, double h, const double* p) {
    double J[9], um[3];
    parts(s, J, um);
    double jj = 0.0, jn2 = 0.0, uu = 0.0;
    for (int k = 0; k < 9; ++k) jj += J[k] * J[k];
    for (int c = 0; c < 3; ++c) {
      const double jn = J[3 * c] * n[0] + J[3 * c + 1] * n[1] + J[3 * c + 2] * n[2];
      jn2 += jn * jn;
      uu += um[c] * um[c];
    }
    return p[0] * h * (0.5 * jj + 0.25 * p[1] * jj * jj) + 0.5 * p[2] * jn2 + 0.5 * p[3] * uu * jj;
  }
  __device__ static void residual(const double* s, const double* x, const double* n, double h, const double* p,
                                  double* R) {
    double J[9], um[3];
    parts(s, J, um);
    double jj = 0.0, uu = 0.0;
    for (int k = 0; k < 9; ++k) jj += J[k] * J[k];
    for (int c = 0; c < 3; ++c) uu += um[c] * um[c];
    for (int c = 0; c < 3; ++c) {
      const double jn = J[3 * c] * n[0] + J[3 * c + 1] * n[1] + J[3 * c + 2] * n[2];
      for (int K = 0; K < 3; ++K) {
        const double d = p[0] * h * (1.0 + p[1] * jj) * J[3 * c + K] + p[2] * jn * n[K] + p[3] * uu * J[3 * c + K];
        R[3 + 3 * c + K] = d;
        R[15 + 3 * c + K] = -d;
      }
      R[c] = 0.5 * p[3] * jj * um[c];
      R[12 + c] = 0.5 * p[3] * jj * um[c];
    }
  }
  __device__ static void jvp(const double* s, const double* ds, const double* x, const double* n, double h,
                             const double* p, double* dR) {
    double J[9], um[3], dJ[9], dum[3];
    parts(s, J, um);
    parts(ds, dJ, dum);
    double jj = 0.0, djj = 0.0, uu = 0.0, duu = 0.0;
    for (int k = 0; k < 9; ++k) {
      jj += J[k] * J[k];
      djj += 2.0 * J[k] * dJ[k];
    }
    for (int c = 0; c < 3; ++c) {
      uu += um[c] * um[c];
      duu += 2.0 * um[c] * dum[c];
    }
    for (int c = 0; c < 3; ++c) {
      const double dn = dJ[3 * c] * n[0] + dJ[3 * c + 1] * n[1] + dJ[3 * c + 2] * n[2];
      for (int K = 0; K < 3; ++K) {
        const double d = p[0] * h * ((1.0 + p[1] * jj) * dJ[3 * c + K] + p[1] * djj * J[3 * c + K]) + p[2] * dn * n[K] +
                         p[3] * (uu * dJ[3 * c + K] + duu * J[3 * c + K]);
        dR[3 + 3 * c + K] = d;
        dR[15 + 3 * c + K] = -d;
      }
      dR[c] = 0.5 * p[3] * (djj * um[c] + jj * dum[c]);
      dR[12 + c] = 0.5 * p[3] * (djj * um[c] + jj * dum[c]);
    }
  }
};
"""

NONLINEAR_BOUNDARY_SOURCE = r"""
struct NonlinearBoundary {
  __device__ static double energy(const double* s, const double* x, const double* n, double h, const double* p) {
    double uu = 0.0, gn2 = 0.0;
    for (int c = 0; c < 3; ++c) {
      uu += s[c] * s[c];
      const double gn = s[3 + 3 * c] * n[0] + s[4 + 3 * c] * n[1] + s[5 + 3 * c] * n[2];
      gn2 += gn * gn;
    }
    return 0.5 * p[0] * uu + 0.25 * p[1] * uu * uu + 0.5 * p[2] * gn2 + 0.5 * p[3] * x[0] * uu;
  }
  __device__ static void residual(const double* s, const double* x, const double* n, double h, const double* p,
                                  double* R) {
    double uu = 0.0;
    for (int c = 0; c < 3; ++c) uu += s[c] * s[c];
    const double a = p[0] + p[1] * uu + p[3] * x[0];
    for (int c = 0; c < 3; ++c) {
      R[c] = a * s[c];
      const double gn = s[3 + 3 * c] * n[0] + s[4 + 3 * c] * n[1] + s[5 + 3 * c] * n[2];
      for (int K = 0; K < 3; ++K) R[3 + 3 * c + K] = p[2] * gn * n[K];
    }
  }
  __device__ static void jvp(const double* s, const double* ds, const double* x, const double* n, double h,
                             const double* p, double* dR) {
    double uu = 0.0, udu = 0.0;
    for (int c = 0; c < 3; ++c) {
      uu += s[c] * s[c];
      udu += s[c] * ds[c];
    }
    const double a = p[0] + p[1] * uu + p[3] * x[0];
    for (int c = 0; c < 3; ++c) {
      dR[c] = a * ds[c] + 2.0 * p[1] * udu * s[c];
      const double gn = ds[3 + 3 * c] * n[0] + ds[4 + 3 * c] * n[1] + ds[5 + 3 * c] * n[2];
      for (int K = 0; K < 3; ++K) dR[3 + 3 * c + K] = p[2] * gn * n[K];
    }
  }
};
"""
