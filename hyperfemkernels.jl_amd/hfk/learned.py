"""Learned coarse correction (no counterpart in the reference): host side over libhfk.

``CoarseNetwork`` wraps the device U-Net of hfk_net.hip (BF16 on MFMA, FP32 master weights, trained on the
device with the library's own forward/backward kernels and Adam).  It acts on the assembled Q1 level of a
``PMultigrid``: ``set_coarse_network(pmg, net, mode)`` makes it the coarse correction (mode "replace"), the initial guess of the
coarse PCG (mode "warmstart"), or the coarse-grid correction of a two-grid cycle on the assembled coarse level,
between Chebyshev pre- and post-smoothing on that level: one such cycle as the coarse correction (mode "cycle") or
as the preconditioner of a flexible coarse CG (mode "precondition").  The outer flexible CG's tolerance fixes the
answer in every mode.
Training pairs come from the plain p-multigrid itself (``coarse_pairs``).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, ptr

MODES = {"off": 0, "warmstart": 1, "replace": 2, "precondition": 3, "cycle": 4}


class CoarseNetwork:
    """U-Net on the vertex lattice of a Cartesian n^3 Q1 space (``coarse_space``: mesh.HexSpace of order 1).

    With ``coarsen`` the U-Net runs on every other lattice vertex (fed by full weighting, prolongated back
    trilinearly), which keeps its receptive field at the whole domain on fine coarse levels (n = 64)."""

    def __init__(self, ctx, coarse_space, seed: int = 0, coarsen: bool = False):
        n = coarse_space.n
        if not (n[0] == n[1] == n[2]) or coarse_space.order != 1:
            raise ValueError("HyperFEMKernels: the coarse network needs a cubic Q1 lattice")
        self.ctx, self.lib = ctx, ctx.lib
        self.N = n[0] + 1
        self.ncoarse = coarse_space.nfree
        dofs = np.ascontiguousarray(coarse_space.lattice_dofs(), dtype=np.int32)
        h = (coarse_space.domain[1] - coarse_space.domain[0]) / n[0]
        hd = C.c_void_p()
        check(self.lib.hfk_net_create(ctx.handle, self.N, int(bool(coarsen)), ptr(dofs, C.c_int32), self.ncoarse, h, seed,
                                      C.byref(hd)), ctx.handle)
        self.coarsen = bool(coarsen)
        self.handle = hd
        self.h = h
        self.nparams = int(self.lib.hfk_net_num_params(hd))

    def set_aux(self, aux_vertex):
        """Auxiliary input channel per lattice vertex (e.g. normalised log shear modulus), or None."""
        if aux_vertex is None:
            check(self.lib.hfk_net_set_aux(self.handle, None), self.ctx.handle)
            return
        a = np.ascontiguousarray(aux_vertex, dtype=np.float32)
        if a.size != self.N ** 3:
            raise ValueError("HyperFEMKernels: one auxiliary value per lattice vertex")
        check(self.lib.hfk_net_set_aux(self.handle, a.ctypes.data_as(C.POINTER(C.c_float))), self.ctx.handle)

    def weights(self) -> np.ndarray:
        w = np.zeros(self.nparams, dtype=np.float32)
        check(self.lib.hfk_net_get_weights(self.handle, w.ctypes.data_as(C.POINTER(C.c_float))), self.ctx.handle)
        return w

    def set_weights(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        if w.size != self.nparams:
            raise ValueError("HyperFEMKernels: wrong number of network weights")
        check(self.lib.hfk_net_set_weights(self.handle, w.ctypes.data_as(C.POINTER(C.c_float))), self.ctx.handle)

    def apply(self, r):
        r = np.ascontiguousarray(r, dtype=np.float64)
        z = np.zeros(self.ncoarse)
        check(self.lib.hfk_net_apply(self.handle, ptr(r, C.c_double), ptr(z, C.c_double)), self.ctx.handle)
        return z

    def gradient(self, r, e):
        """(loss, dloss/dweights) summed over the samples (rows of r and e)."""
        r = np.ascontiguousarray(np.atleast_2d(r), dtype=np.float64)
        e = np.ascontiguousarray(np.atleast_2d(e), dtype=np.float64)
        g = np.zeros(self.nparams, dtype=np.float32)
        loss = C.c_double()
        check(self.lib.hfk_net_gradient(self.handle, ptr(r, C.c_double), ptr(e, C.c_double), r.shape[0],
                                        g.ctypes.data_as(C.POINTER(C.c_float)), C.byref(loss)), self.ctx.handle)
        return loss.value, g

    def train_dev(self, r_ptr: int, e_ptr: int, nsamples: int, steps: int, batch: int = 4, lr: float = 1e-3,
                  seed: int = 0, stream: int = 0) -> np.ndarray:
        """Adam on device-resident pairs ([k][ncoarse] doubles at r_ptr/e_ptr); returns the loss per step."""
        hist = np.zeros(max(steps, 1))
        check(self.lib.hfk_net_train_dev(self.handle, C.c_void_p(r_ptr), C.c_void_p(e_ptr), nsamples, steps, batch,
                                         lr, seed, ptr(hist, C.c_double), C.c_void_p(stream)), self.ctx.handle)
        return hist[:steps]

    def __del__(self):
        try:
            self.lib.hfk_net_destroy(self.handle)
        except Exception:
            pass


def coarse_pairs_dev(pmg, b_ptr: int, nsamples: int, bc_ptr: int, ec_ptr: int | None, coarse_smoothed: bool = False,
                     rtol: float = 1e-10, maxiter: int = 5000, stream: int = 0):
    """Training data from the plain p-multigrid (device arrays [k][n]): restricted pre-smoothed residuals (with
    ``coarse_smoothed``, the residuals after the coarse Chebyshev pre-smoothing: the network cycle's input), and
    their coarse solutions when ec_ptr is given.  The work runs on ``stream`` (0: the context's own stream, which
    does not wait for other streams): pass the stream that wrote the inputs and the output buffers, or synchronise
    first."""
    check(pmg.lib.hfk_pmg_coarse_pairs_dev(pmg.handle, C.c_void_p(b_ptr), nsamples, C.c_void_p(bc_ptr),
                                           C.c_void_p(ec_ptr) if ec_ptr else None, int(bool(coarse_smoothed)), rtol,
                                           maxiter, C.c_void_p(stream)),
          pmg.levels[0].ctx.handle)


def train_energy_dev(pmg, net: "CoarseNetwork", bc_ptr: int, nsamples: int, steps: int, batch: int = 4,
                     lr: float = 1e-3, seed: int = 0, stream: int = 0) -> np.ndarray:
    """Energy training of the network on the coarse level of ``pmg`` (no exact solves); loss per step."""
    hist = np.zeros(max(steps, 1))
    check(pmg.lib.hfk_pmg_train_coarse_network(pmg.handle, net.handle, C.c_void_p(bc_ptr), nsamples, steps, batch, lr,
                                               seed, ptr(hist, C.c_double), C.c_void_p(stream)),
          pmg.levels[0].ctx.handle)
    return hist[:steps]


def train_energy_problems(problems, net: "CoarseNetwork", steps: int, chunk: int = 25, batch: int = 4,
                          lr: float = 1e-3, seed: int = 0, stream: int = 0, decay: str = "none", step0: int = 0,
                          total: int | None = None) -> np.ndarray:
    """Energy training of one network over several operators: ``problems`` is a list of (pmg, bc_ptr, nsamples,
    aux) -- the p-multigrid whose coarse matrix defines the energy, its device samples, and the auxiliary vertex
    field of that operator (or None).  The problems take turns, ``chunk`` Adam steps each (the optimiser state
    carries over); ``decay="cosine"`` lowers the step size from ``lr`` to lr/20 along a half cosine over a run of
    ``total`` steps (default ``steps``) of which this call does steps ``step0 .. step0 + steps`` (a long run can be
    done in slices).  Returns the loss per step."""
    if decay not in ("none", "cosine"):
        raise ValueError(f"HyperFEMKernels: unknown learning-rate decay {decay!r}")
    total = steps if total is None else total
    hist = []
    done, turn = 0, step0 // chunk
    while done < steps:
        pmg, bc_ptr, ns, aux = problems[turn % len(problems)]
        k = min(chunk, steps - done)
        net.set_aux(aux)
        step_lr = lr
        if decay == "cosine":
            step_lr = lr * (0.05 + 0.95 * 0.5 * (1.0 + np.cos(np.pi * (step0 + done + 0.5 * k) / total)))
        hist.append(train_energy_dev(pmg, net, bc_ptr, ns, k, batch=batch, lr=step_lr, seed=seed + turn,
                                     stream=stream))
        done += k
        turn += 1
    return np.concatenate(hist) if hist else np.zeros(0)


def energy_gradient_dev(pmg, net: "CoarseNetwork", bc_ptr: int, nsamples: int, stream: int = 0):
    """The energy loss the training descends, summed over the device samples, and its weight gradient (FP32)."""
    g = np.zeros(net.nparams, dtype=np.float32)
    loss = C.c_double()
    check(pmg.lib.hfk_pmg_coarse_network_energy_gradient(pmg.handle, net.handle, C.c_void_p(bc_ptr), nsamples,
                                                         g.ctypes.data_as(C.POINTER(C.c_float)), C.byref(loss),
                                                         C.c_void_p(stream)), pmg.levels[0].ctx.handle)
    return loss.value, g


def coarse_error_dev(pmg, net: "CoarseNetwork", bc_ptr: int, nsamples: int, stream: int = 0) -> float:
    """Mean relative energy-norm error of the network's coarse correction (against the coarse solve)."""
    err = C.c_double()
    check(pmg.lib.hfk_pmg_coarse_network_error(pmg.handle, net.handle, C.c_void_p(bc_ptr), nsamples, C.byref(err),
                                               C.c_void_p(stream)), pmg.levels[0].ctx.handle)
    return err.value


def set_coarse_network(pmg, net: CoarseNetwork | None, mode: str = "warmstart", smoother_degree: int = 0,
                       scaled: bool = True):
    """Attach ``net`` to the coarse level of ``pmg`` (modes in MODES); ``smoother_degree`` is the degree of the
    coarse Chebyshev smoother of the network cycle (0: the p-multigrid's own degree); with ``scaled`` the network
    acts on the diagonally scaled coarse system D^-1/2 A D^-1/2 (training and error included)."""
    if mode not in MODES:
        raise ValueError(f"HyperFEMKernels: unknown coarse network mode {mode!r}")
    check(pmg.lib.hfk_pmg_set_coarse_network(pmg.handle, None if net is None else net.handle, MODES[mode],
                                             int(smoother_degree), int(bool(scaled))), pmg.levels[0].ctx.handle)
    pmg.coarse_network = (net, mode)
