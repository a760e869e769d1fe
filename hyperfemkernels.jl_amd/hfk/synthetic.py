"""Portable synthetic inputs (SplitMix64), identical in C++, Python and Julia.

* SplitMix64 (Steele, Lea, Flood, OOPSLA 2014): ``state += 0x9E3779B97F4A7C15``, then the two xor-shift-multiply
  rounds of the published finaliser.  Output ``i`` of seed ``s`` is ``mix(s + (i + 1) * golden)``.
* ``u01 = (x >> 11) * 2^-53`` in [0, 1).
* N(0, 1) by Box--Muller, cosine branch, from consecutive pairs ``(x_2i, x_2i+1)`` with
  ``u1 = ((x_2i >> 11) + 1) * 2^-53`` in (0, 1] and ``u2 = u01(x_2i+1)``.

The displacement states and material fields are the recipes of the benchmark workloads:

* smooth state: ``u_c(x) = amp / M * sum_m sin(2 pi k_cm . x + phi_cm)``, ``k_cm`` in {1,2,3}^3 uniform,
  ``phi_cm`` ~ U[0, 2 pi);
* lognormal shear modulus ``mu(x) = exp(sigma g(x))`` with ``g`` a unit-variance Gaussian field of squared
  exponential correlation length ``ell`` from ``M`` random Fourier features:
  ``g = sqrt(2/M) sum cos(w . x + b)``, ``w ~ N(0, ell^-2 I)``, ``b ~ U[0, 2 pi)``.
"""
from __future__ import annotations

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, n: int, offset: int = 0) -> np.ndarray:
    """Outputs ``offset .. offset + n - 1`` of the SplitMix64 stream of ``seed`` (uint64)."""
    with np.errstate(over="ignore"):
        i = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
        z = np.uint64(seed) + i * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def uniform(seed: int, n: int, offset: int = 0) -> np.ndarray:
    return (splitmix64(seed, n, offset) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53


def normal(seed: int, n: int, offset: int = 0) -> np.ndarray:
    x = splitmix64(seed, 2 * n, 2 * offset)
    u1 = ((x[0::2] >> np.uint64(11)).astype(np.float64) + 1.0) * 2.0 ** -53
    u2 = (x[1::2] >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


def smooth_field(seed: int, modes: int = 8, amp: float = 1e-2):
    """The smooth random displacement ``u(x)`` (callable on (m, 3) coordinates -> (m, 3))."""
    r = splitmix64(seed, 3 * modes * 4)
    kk = (r[: 9 * modes] % np.uint64(3)).astype(np.float64).reshape(3, modes, 3) + 1.0
    phi = (r[9 * modes: 12 * modes] >> np.uint64(11)).astype(np.float64).reshape(3, modes) * 2.0 ** -53 * 2 * np.pi

    def u(xyz):
        xyz = np.asarray(xyz, dtype=np.float64)
        out = np.zeros((xyz.shape[0], 3))
        for c in range(3):
            arg = 2 * np.pi * (xyz @ kk[c].T) + phi[c][None, :]
            out[:, c] = amp / modes * np.sin(arg).sum(axis=1)
        return out

    return u


def lognormal_features(seed: int, ell: float = 0.1, features: int = 1024):
    """Random Fourier features of the unit-variance Gaussian field with squared-exponential correlation length
    ``ell``: g(x) = sqrt(2/M) sum_m cos(w_m . x + b_m), w ~ N(0, ell^-2 I), b ~ U[0, 2 pi)."""
    w = normal(seed, 3 * features).reshape(features, 3) / ell
    b = uniform(seed + 1, features) * 2 * np.pi
    return w, b


def lognormal_field(seed: int, sigma: float = 1.0, ell: float = 0.1, features: int = 1024):
    """``mu_factor(x) = exp(sigma g(x))`` (callable on (m, 3) -> (m,))."""
    w, b = lognormal_features(seed, ell, features)

    def f(xyz):
        xyz = np.asarray(xyz, dtype=np.float64)
        out = np.empty(xyz.shape[0])
        for s in range(0, xyz.shape[0], 1 << 16):
            g = np.sqrt(2.0 / features) * np.cos(xyz[s:s + (1 << 16)] @ w.T + b[None, :]).sum(axis=1)
            out[s:s + (1 << 16)] = np.exp(sigma * g)
        return out

    return f


# mu(x) = p[0] exp(p[1] sqrt(2/M) sum_m cos(w_m . x + b_m)), M = p[2], w = p[3 .. 3+3M) (m-major), b = p[3+3M ..):
# the lognormal field above as a function of x for KernelOperator.set_mu_function (evaluated on the device)
LOGNORMAL_MU_SOURCE = r"""
__device__ double lognormal_mu(const double* x, const double* p) {
  const int M = (int)p[2];
  const double* w = p + 3;
  const double* b = p + 3 + 3 * M;
  double g = 0.0;
  for (int m = 0; m < M; ++m) g += cos(w[3 * m] * x[0] + w[3 * m + 1] * x[1] + w[3 * m + 2] * x[2] + b[m]);
  return p[0] * exp(p[1] * sqrt(2.0 / M) * g);
}
"""


def lognormal_mu_params(seed: int, sigma: float = 1.0, ell: float = 0.1, features: int = 1024, mu0: float = 1.0):
    """The parameter array of LOGNORMAL_MU_SOURCE for mu0 * lognormal_field(seed, sigma, ell, features)."""
    w, b = lognormal_features(seed, ell, features)
    return np.concatenate([[mu0, sigma, float(features)], w.reshape(-1), b])
