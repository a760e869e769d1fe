"""Pointwise constitutive laws compiled at run time (hiprtc, gfx950).

The reference compiles the pointwise integrand of a captured weak form into a Julia function (``PointForm``,
src/codegen.jl:166-198) and reuses the compiled code for captures of the same structure (src/tree.jl:244-283).
Here the pointwise part is the constitutive law -- HIP source of a struct with

    __device__ void setup(const double* F, double lam, double mu);   // F[c + 3J] = d x_c / d X_J
    __device__ double stress(int c, int J) const;                     // first Piola-Kirchhoff P_cJ
    __device__ double tangent(int c, int J, int k, int L) const;     // d P_cJ / d F_kL
    __device__ double energy() const;                                 // W (optional: the energy functional)

-- compiled by libhfk (``hfk_law_create``) into the pointwise kernels of ``prepare!`` (stored tangents in the
operator's layout) and ``residual!`` (point coefficients), between the library's own staging and contraction
kernels.  Code objects are cached per process by a hash of the source.  ``hfk_det3(a)`` and
``hfk_inv3(a, r, det)`` (column-major 3x3) are in scope.

``SVK_SOURCE``, ``NEOHOOKE_SOURCE`` and ``LINEAR_SOURCE`` are the built-in laws written as such sources
(templates for new laws); ``NEOHOOKE_QUADVOL_SOURCE`` is a law the library does not build in.
"""
from __future__ import annotations

import ctypes as C

from . import _lib
from ._lib import check

SVK_SOURCE = r"""
struct SVK {  // P = F S, S = lam tr(E) I + 2 mu E, E = (F'F - I)/2
  double F[9], S[9], lam, mu;
  __device__ void setup(const double* Fin, double l, double m) {
    lam = l;
    mu = m;
    for (int i = 0; i < 9; ++i) F[i] = Fin[i];
    double E[9];
    for (int I = 0; I < 3; ++I)
      for (int J = 0; J < 3; ++J) {
        double s = 0;
        for (int k = 0; k < 3; ++k) s += F[k + 3 * I] * F[k + 3 * J];
        E[I + 3 * J] = 0.5 * (s - (I == J ? 1.0 : 0.0));
      }
    const double tr = E[0] + E[4] + E[8];
    for (int I = 0; I < 3; ++I)
      for (int J = 0; J < 3; ++J) S[I + 3 * J] = (I == J ? lam * tr : 0.0) + 2 * mu * E[I + 3 * J];
  }
  __device__ double stress(int c, int J) const {
    double s = 0;
    for (int I = 0; I < 3; ++I) s += F[c + 3 * I] * S[I + 3 * J];
    return s;
  }
  __device__ double tangent(int c, int J, int k, int L) const {
    double b = 0;
    for (int I = 0; I < 3; ++I) b += F[c + 3 * I] * F[k + 3 * I];
    return (c == k ? S[L + 3 * J] : 0.0) + lam * F[c + 3 * J] * F[k + 3 * L] +
           mu * (F[c + 3 * L] * F[k + 3 * J] + (J == L ? b : 0.0));
  }
  __device__ double energy() const {  // lam/2 tr(E)^2 + mu E:E
    double tr = 0, ee = 0;
    for (int I = 0; I < 3; ++I)
      for (int J = 0; J < 3; ++J) {
        double s = 0;
        for (int k = 0; k < 3; ++k) s += F[k + 3 * I] * F[k + 3 * J];
        const double e = 0.5 * (s - (I == J ? 1.0 : 0.0));
        ee += e * e;
        if (I == J) tr += e;
      }
    return 0.5 * lam * tr * tr + mu * ee;
  }
};
"""

NEOHOOKE_SOURCE = r"""
struct NeoHooke {  // W = mu/2 (I1 - 3) - mu lnJ + lam/2 lnJ^2
  double F[9], A[9], lam, mu, lnJ;  // A = F^-1
  __device__ void setup(const double* Fin, double l, double m) {
    lam = l;
    mu = m;
    for (int i = 0; i < 9; ++i) F[i] = Fin[i];
    const double d = hfk_det3(F);
    hfk_inv3(F, A, d);
    lnJ = log(d);
  }
  __device__ double stress(int c, int J) const { return mu * F[c + 3 * J] + (lam * lnJ - mu) * A[J + 3 * c]; }
  __device__ double tangent(int c, int J, int k, int L) const {
    return (c == k && J == L ? mu : 0.0) + (mu - lam * lnJ) * A[J + 3 * k] * A[L + 3 * c] +
           lam * A[J + 3 * c] * A[L + 3 * k];
  }
  __device__ double energy() const {
    double I1 = 0;
    for (int i = 0; i < 9; ++i) I1 += F[i] * F[i];
    return 0.5 * mu * (I1 - 3.0) - mu * lnJ + 0.5 * lam * lnJ * lnJ;
  }
};
"""

LINEAR_SOURCE = r"""
struct Linear {  // sigma = lam tr(G) I + mu (G + G'), G = F - I
  double G[9], lam, mu, tr;
  __device__ void setup(const double* Fin, double l, double m) {
    lam = l;
    mu = m;
    for (int c = 0; c < 3; ++c)
      for (int J = 0; J < 3; ++J) G[c + 3 * J] = Fin[c + 3 * J] - (c == J ? 1.0 : 0.0);
    tr = G[0] + G[4] + G[8];
  }
  __device__ double stress(int c, int J) const { return (c == J ? lam * tr : 0.0) + mu * (G[c + 3 * J] + G[J + 3 * c]); }
  __device__ double tangent(int c, int J, int k, int L) const {
    return (c == J && k == L ? lam : 0.0) + (c == k && J == L ? mu : 0.0) + (J == k && c == L ? mu : 0.0);
  }
  __device__ double energy() const {  // lam/2 tr^2 + mu eps:eps
    double ee = 0;
    for (int c = 0; c < 3; ++c)
      for (int J = 0; J < 3; ++J) ee += 0.25 * (G[c + 3 * J] + G[J + 3 * c]) * (G[c + 3 * J] + G[J + 3 * c]);
    return 0.5 * lam * tr * tr + mu * ee;
  }
};
"""

# W = mu/2 (I1 - 3) - mu lnJ + lam/2 (J - 1)^2: P = mu (F - F^-T) + g(J) F^-T with g = lam (J - 1) J;
# dP_cJ/dF_kL = mu d_ck d_JL + (mu - g) A_Jk A_Lc + g'(J) J A_Jc A_Lk, g' J = lam (2J - 1) J.
# At F = I it reduces to the linear law's tangent (g(1) = 0, g'(1) = lam).
NEOHOOKE_QUADVOL_SOURCE = r"""
struct NeoHookeQuadVol {
  double F[9], A[9], lam, mu, g, gJ, J;
  __device__ void setup(const double* Fin, double l, double m) {
    lam = l;
    mu = m;
    for (int i = 0; i < 9; ++i) F[i] = Fin[i];
    const double d = hfk_det3(F);
    hfk_inv3(F, A, d);
    J = d;
    g = lam * (d - 1.0) * d;
    gJ = lam * (2.0 * d - 1.0) * d;
  }
  __device__ double energy() const {
    double I1 = 0;
    for (int i = 0; i < 9; ++i) I1 += F[i] * F[i];
    return 0.5 * mu * (I1 - 3.0) - mu * log(J) + 0.5 * lam * (J - 1.0) * (J - 1.0);
  }
  __device__ double stress(int c, int J) const { return mu * F[c + 3 * J] + (g - mu) * A[J + 3 * c]; }
  __device__ double tangent(int c, int J, int k, int L) const {
    return (c == k && J == L ? mu : 0.0) + (mu - g) * A[J + 3 * k] * A[L + 3 * c] + gJ * A[J + 3 * c] * A[L + 3 * k];
  }
};
"""

SOURCES = {"svk": (SVK_SOURCE, "SVK"), "neohookean": (NEOHOOKE_SOURCE, "NeoHooke"),
           "linear": (LINEAR_SOURCE, "Linear"), "neohookean_quadvol": (NEOHOOKE_QUADVOL_SOURCE, "NeoHookeQuadVol")}


def check_law(source: str, name: str) -> str:
    """Compile ``source`` without a device (``hfk_law_check``); returns the compiler log, raises ``ValueError``
    with it when the law does not compile."""
    lib = _lib.load()
    buf = C.create_string_buffer(1 << 16)
    st = lib.hfk_law_check(source.encode(), name.encode(), buf, len(buf))
    if st != _lib.HFK_OK:
        msg = lib.hfk_last_error(None)
        raise ValueError(msg.decode() if msg else "HyperFEMKernels: the law does not compile")
    return buf.value.decode()


class CompiledLaw:
    """A law compiled for the device of ``ctx`` (``hfk_law_create``).  Attach it with
    ``KernelOperator.set_law``; keep it alive while an operator uses it."""

    def __init__(self, source: str, name: str, ctx=None):
        from .operator import default_context

        self.ctx = ctx or default_context()
        self.lib = self.ctx.lib
        self.source, self.name = source, name
        h = C.c_void_p()
        st = self.lib.hfk_law_create(self.ctx.handle, source.encode(), name.encode(), C.byref(h))
        if st != _lib.HFK_OK:
            msg = self.lib.hfk_last_error(self.ctx.handle)
            raise ValueError(msg.decode() if msg else "HyperFEMKernels: the law does not compile")
        self.handle = h
        key, cached = C.c_uint64(), C.c_int()
        check(self.lib.hfk_law_info(h, C.byref(key), C.byref(cached)), self.ctx.handle)
        self.key, self.from_cache = int(key.value), bool(cached.value)

    @classmethod
    def builtin(cls, law: str, ctx=None) -> "CompiledLaw":
        src, name = SOURCES[law]
        return cls(src, name, ctx)

    def close(self):
        if getattr(self, "handle", None):
            self.lib.hfk_law_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
