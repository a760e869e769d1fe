"""Velocity-pressure point forms generated from a captured integrand, as HIP source for ``hfk_oseen_set_form``.

The reference captures the weak form ``a((u,p),(v,q))`` and ``l((v,q))`` of an operator, differentiates the
integrand with respect to the test and trial slots at every quadrature point and emits the pointwise code
(``emit!``/``PointForm``, src/codegen.jl:122-198; test x trial factorisation, src/factorforms.jl:108-233).  Here the
host writes the integrand with the scalar DAG of ``hfk.codegen`` over symbolic slots -- the trial function u, its
gradient, p and its gradient; the same for the test function -- and over the point data of the step (w, un, f,
lap), the cell size h and the parameters rho, mu, c0 (plus the form's own).  From it:

* the residual coefficients: co_s = d(a(x; v) - l(v)) / d v_s for the 16 test slots, at the state slots of x;
* the tangent T[s][r] = d^2 a / d v_s d u_r (16 x 16), projected onto the 50 factors of the factorised class that the
  device kernel contracts (hfk_oseen_form.hpp); the projection is checked on random points -- the reconstructed
  T must equal the derived one entry by entry (as the reference's check_factorisation does) -- and a form
  outside the class, or one whose tangent depends on the state, is refused.

Both are emitted as one straight-line function each (one temporary per DAG node) in a struct for hiprtc.  The
integrand of the Oseen-BDF2 step with its SUPG/PSPG/LSIC terms (test/common.jl:165-228) is ``oseen_supg``.
"""
from __future__ import annotations

import math
from typing import Callable

import numpy as np

from . import codegen as cg
from .codegen import Node, const, diff, sqrt, var, wrap

# slot numbering, test and trial alike: value 0..2, d/dx_J of component i at 3 + 3i + J, p at 12, dp/dx_J at 13 + J
NSLOT = 16


def grad_slot(i: int, J: int) -> int:
    return 3 + 3 * i + J


class Vec3:
    def __init__(self, e):
        self.e = [wrap(v) for v in e]

    def __getitem__(self, i):
        return self.e[i]

    def __add__(self, o):
        return Vec3([a + b for a, b in zip(self.e, o.e)])

    def __sub__(self, o):
        return Vec3([a - b for a, b in zip(self.e, o.e)])

    def __mul__(self, s):
        s = wrap(s)
        return Vec3([a * s for a in self.e])

    __rmul__ = __mul__

    def __neg__(self):
        return Vec3([-a for a in self.e])


def dot(a: Vec3, b: Vec3) -> Node:
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]


def norm(a: Vec3) -> Node:
    return sqrt(dot(a, a))


class Field:
    """A velocity-pressure function at a point: u (Vec3), grad (Mat3 with grad[J, i] = d u_i / d x_J, Gridap's
    convention), p and gp (Vec3)."""

    def __init__(self, slots):
        self.slots = slots
        self.u = Vec3(slots[0:3])
        self.grad = cg.Mat3([slots[grad_slot(i, J)] for i in range(3) for J in range(3)])  # [J + 3 i]
        self.p = slots[12]
        self.gp = Vec3(slots[13:16])

    def div(self) -> Node:
        return cg.tr(self.grad)

    def conv(self, w: Vec3) -> Vec3:
        """(grad u)' . w: component i = sum_J d u_i/dx_J w_J."""
        return Vec3([sum((self.grad[J, i] * w[J] for J in range(3)), const(0.0)) for i in range(3)])

    def eps(self) -> cg.Mat3:
        return (self.grad + self.grad.T()) * 0.5


class Data:
    """The point data and parameters a form may use: w, un, f, lap (Vec3), h, rho, mu, c0 and the form's own
    parameters by name (``param``)."""

    def __init__(self, extra=()):
        self.w = Vec3([var(f"d{i}") for i in range(3)])
        self.un = Vec3([var(f"d{3 + i}") for i in range(3)])
        self.f = Vec3([var(f"d{6 + i}") for i in range(3)])
        self.lap = Vec3([var(f"d{9 + i}") for i in range(3)])
        self.h = var("h")
        self.rho, self.mu, self.c0 = var("prm0"), var("prm1"), var("prm2")
        self.extra = list(extra)
        if len(self.extra) > 5:
            raise ValueError("HyperFEMKernels: a point form takes at most 5 parameters of its own")
        self._extra = {n: var(f"prm{3 + k}") for k, n in enumerate(self.extra)}

    def param(self, name: str) -> Node:
        return self._extra[name]

    def variables(self):
        return [*self.w.e, *self.un.e, *self.f.e, *self.lap.e, self.h, self.rho, self.mu, self.c0,
                *self._extra.values()]


def _evaluate(outputs, binding: dict):
    """Values of the nodes `outputs` (numpy, FP64) with the variables bound by id in `binding`."""
    val = dict(binding)

    def ev(n):
        stack = [(n, False)]
        while stack:
            m, done = stack.pop()
            if id(m) in val:
                continue
            if m.op == "const":
                val[id(m)] = m.value
                continue
            if m.op == "var":
                raise ValueError(f"HyperFEMKernels: unbound variable {m.value}")
            if not done:
                stack.append((m, True))
                stack.extend((a, False) for a in m.args if id(a) not in val)
                continue
            a = [val[id(x)] for x in m.args]
            val[id(m)] = {"add": lambda: a[0] + a[1], "sub": lambda: a[0] - a[1], "mul": lambda: a[0] * a[1],
                          "div": lambda: a[0] / a[1], "neg": lambda: -a[0], "log": lambda: math.log(a[0]),
                          "exp": lambda: math.exp(a[0]), "sqrt": lambda: math.sqrt(a[0])}[m.op]()
        return val[id(n)]

    return [ev(o) for o in outputs]


def _emit(outputs, names: dict, out_name: str):
    """Straight-line HIP statements computing `outputs` into out_name[k] (one temporary per node)."""
    order, seen = [], set()
    for o in outputs:
        stack = [(o, False)]
        while stack:
            n, done = stack.pop()
            if id(n) in seen or n.op in ("const", "var"):
                continue
            if done:
                seen.add(id(n))
                order.append(n)
                continue
            stack.append((n, True))
            stack.extend((a, False) for a in reversed(n.args) if id(a) not in seen)
    loc = dict(names)
    body = []

    def ref(n):
        if n.op == "const":
            return repr(float(n.value)) if math.isfinite(n.value) else "(0.0 / 0.0)"
        return loc[id(n)]

    for n in order:
        a = [ref(x) for x in n.args]
        expr = {"add": f"{a[0]} + {a[-1]}", "sub": f"{a[0]} - {a[-1]}", "mul": f"{a[0]} * {a[-1]}",
                "div": f"{a[0]} / {a[-1]}", "neg": f"-{a[0]}", "log": f"log({a[0]})", "exp": f"exp({a[0]})",
                "sqrt": f"sqrt({a[0]})"}[n.op]
        t = f"t{len(body)}"
        loc[id(n)] = t
        body.append(f"    const double {t} = {expr};")
    body += [f"    {out_name}[{k}] = {ref(n)};" for k, n in enumerate(outputs)]
    return body


class OseenPointForm:
    """A velocity-pressure point form from the integrand ``a(U, V, D)`` (bilinear in the trial U and the test V)
    and the load ``l(V, D)``: U and V are ``Field``s of symbolic slots, D the ``Data``.  ``params``: names of the
    form's own parameters (values given to ``OseenOperator.set_form``)."""

    def __init__(self, a: Callable, l: Callable, params=(), name: str = "GeneratedOseenForm"):
        self.name = name
        self.params = tuple(params)
        self.D = D = Data(self.params)
        self.trial = [var(f"u{k}") for k in range(NSLOT)]
        self.test = [var(f"v{k}") for k in range(NSLOT)]
        self.state = [var(f"s{k}") for k in range(NSLOT)]
        U, V, X = Field(self.trial), Field(self.test), Field(self.state)
        aw = wrap(a(U, V, D))
        # residual: a(x; v) - l(v), coefficient of every test slot at the state slots
        rw = wrap(a(X, V, D)) - wrap(l(V, D))
        self.co = [diff(rw, self.test[s], {}) for s in range(NSLOT)]
        # tangent: d^2 a / d v_s d u_r
        # (diff memoises per node for one variable: a fresh memo for every variable)
        ds = [diff(aw, self.test[s], {}) for s in range(NSLOT)]
        self.T = [[None] * NSLOT for _ in range(NSLOT)]
        for r in range(NSLOT):
            memo = {}
            for s in range(NSLOT):
                self.T[s][r] = diff(ds[s], self.trial[r], memo)
        self.k = self._project()
        self._check()

    # -- the 50 factors of hfk_oseen_form.hpp ---------------------------------------------------------------
    def _project(self):
        T, g = self.T, grad_slot
        m = T[0][0]
        cu = [T[0][g(0, L)] for L in range(3)]
        cv = [T[g(0, J)][0] for J in range(3)]
        kt = T[g(0, 1)][g(1, 0)]
        kd = T[g(0, 0)][g(1, 1)]
        K = [None] * 9
        for L in range(3):
            for J in range(3):
                e = T[g(0, J)][g(0, L)]
                K[J + 3 * L] = e - kt - kd if (J == 0 and L == 0) else e
        zv = [T[i][12] for i in range(3)]
        zg = T[0][13]
        zd = T[g(0, 0)][12]
        zw = [T[g(0, J)][13] for J in range(3)]
        yv = [T[12][j] for j in range(3)]
        yd = T[12][g(0, 0)]
        yg = T[13][0]
        yw = [T[13][g(0, L)] for L in range(3)]
        x0 = T[12][12]
        xg = [T[12][13 + L] for L in range(3)]
        xh = [T[13 + J][12] for J in range(3)]
        X = [T[13 + J][13 + L] for L in range(3) for J in range(3)]  # [J + 3L]
        return [m, *cu, *cv, *K, kt, kd, *zv, zg, zd, *zw, *yv, yd, yg, *yw, x0, *xg, *xh, *X]

    @staticmethod
    def reconstruct(k) -> np.ndarray:
        """T (16 x 16) from the 50 factors (numpy): the factorised class of hfk_oseen_form.hpp."""
        k = np.asarray(k, dtype=np.float64)
        m, cu, cv, K, kt, kd = k[0], k[1:4], k[4:7], k[7:16].reshape(3, 3, order="F"), k[16], k[17]
        zv, zg, zd, zw = k[18:21], k[21], k[22], k[23:26]
        yv, yd, yg, yw = k[26:29], k[29], k[30], k[31:34]
        x0, xg, xh, X = k[34], k[35:38], k[38:41], k[41:50].reshape(3, 3, order="F")
        T = np.zeros((NSLOT, NSLOT))
        d = np.eye(3)
        g = grad_slot
        for i in range(3):
            for j in range(3):
                T[i, j] = d[i, j] * m
                for L in range(3):
                    T[i, g(j, L)] = d[i, j] * cu[L]
                    T[g(i, L), j] = d[i, j] * cv[L]
                for J in range(3):
                    for L in range(3):
                        T[g(i, J), g(j, L)] = d[i, j] * K[J, L] + kt * d[i, L] * d[J, j] + kd * d[i, J] * d[j, L]
            T[i, 12] = zv[i]
            T[12, i] = yv[i]
            for L in range(3):
                T[i, 13 + L] = d[i, L] * zg
                T[g(i, L), 12] = d[i, L] * zd
                T[13 + L, i] = d[L, i] * yg
                T[12, g(i, L)] = d[i, L] * yd
                for J in range(3):
                    T[g(i, J), 13 + L] = d[i, L] * zw[J]
                    T[13 + J, g(i, L)] = d[J, i] * yw[L]
        T[12, 12] = x0
        T[12, 13:16] = xg
        T[13:16, 12] = xh
        T[13:16, 13:16] = X
        return T

    def _bind(self, rng, state=None):
        b = {id(v): x for v, x in zip(self.D.variables(), self.sample_data(rng))}
        st = rng.standard_normal(NSLOT) if state is None else state
        b.update({id(v): x for v, x in zip(self.state, st)})
        b.update({id(v): x for v, x in zip(self.trial, rng.standard_normal(NSLOT))})
        b.update({id(v): x for v, x in zip(self.test, rng.standard_normal(NSLOT))})
        return b

    def sample_data(self, rng):
        """Random point data and positive parameters (w, un, f, lap, h, rho, mu, c0, own parameters)."""
        return np.concatenate([rng.standard_normal(12), rng.uniform(0.05, 1.0, 4 + len(self.params))])

    def _check(self, trials: int = 4):
        rng = np.random.default_rng(12345)
        flat = [t for row in self.T for t in row]
        for _ in range(trials):
            b = self._bind(rng)
            T = np.array(_evaluate(flat, b)).reshape(NSLOT, NSLOT)
            scale = max(np.abs(T).max(), 1e-300)
            b2 = dict(b)
            b2.update({id(v): x for v, x in zip(self.trial, rng.standard_normal(NSLOT))})
            T2 = np.array(_evaluate(flat, b2)).reshape(NSLOT, NSLOT)
            if np.abs(T - T2).max() > 1e-12 * scale:
                raise ValueError("HyperFEMKernels: the form is not bilinear (its tangent depends on the trial state)")
            Tr = self.reconstruct(_evaluate(self.k, b))
            if np.abs(T - Tr).max() > 1e-12 * scale:
                raise ValueError("HyperFEMKernels: the form's tangent is not of the factorised velocity-pressure class "
                                 "(isotropic in the components, one advecting direction per term)")

    def blocks(self) -> np.ndarray:
        """The (test field, trial field) blocks that carry a term (2 x 2, velocity 0, pressure 1): a block whose
        tangent entries are all the constant 0 carries none (form_blocks, src/assembly.jl:288-303: by structure, so
        a term multiplied by a parameter that happens to be 0 keeps its block)."""
        fld = lambda s: 0 if s < 12 else 1
        m = np.zeros((2, 2), dtype=bool)
        for s in range(NSLOT):
            for r in range(NSLOT):
                t = self.T[s][r]
                if not (t.op == "const" and t.value == 0.0):
                    m[fld(s), fld(r)] = True
        return m

    # -- evaluation and code --------------------------------------------------------------------------------
    def evaluate(self, state, data, h, prm):
        """(co (16), T (16 x 16), k (50)) at one point with numpy: state slots, data (12), h, prm (rho, mu, c0,
        own parameters)."""
        b = {id(v): float(x) for v, x in zip(self.D.variables(), [*data, h, *prm])}
        b.update({id(v): float(x) for v, x in zip(self.state, state)})
        co = np.array(_evaluate(self.co, b))
        T = np.array(_evaluate([t for row in self.T for t in row], b)).reshape(NSLOT, NSLOT)
        k = np.array(_evaluate(self.k, b))
        return co, T, k

    def source(self) -> str:
        names = {id(v): f"s[{k}]" for k, v in enumerate(self.state)}
        for k in range(12):
            names[id(self.D.variables()[k])] = f"d[{k}]"
        names[id(self.D.h)] = "h"
        for k in range(3 + len(self.params)):
            names[id(self.D.variables()[13 + k])] = f"prm[{k}]"
        res = _emit(self.co, names, "co")
        tan = _emit(self.k, names, "k")
        mask = sum(1 << i for i, n in enumerate(self.k) if not (n.op == "const" and n.value == 0.0))
        return "\n".join([
            f"struct {self.name} {{  // generated by hfk.formgen: {len(res)} + {len(tan)} statements",
            f"  static constexpr unsigned long long kNonzero = {mask:#x}ull;  // factors that are not the constant 0",
            "  __device__ static void residual(const double* s, const double* d, double h, const double* prm, "
            "double* co) {",
            "    (void)s; (void)d; (void)h; (void)prm;",
            *res,
            "  }",
            "  __device__ static void tangent(const double* d, double h, const double* prm, double* k) {",
            "    (void)d; (void)h; (void)prm;",
            *tan,
            "  }",
            "};",
            ""])


# ---------------------------------------------------------------------------------------------------------------
# the Oseen-BDF2 step of test/common.jl:165-228


def tau_reference(D: Data):
    """ts = ((2|un|/h)^2 + 9 (4 mu / (rho h^2))^2)^(-1/2), tp = ts / rho, tl = rho |un| h / 2 (test/common.jl:201-205)."""
    unn = norm(D.un)
    r1 = 2.0 * unn / D.h
    r2 = 3.0 * (4.0 * D.mu / (D.rho * D.h * D.h))
    ts = 1.0 / sqrt(r1 * r1 + r2 * r2)
    return ts, ts / D.rho, D.rho * unn * D.h * 0.5


def oseen_supg(tau: Callable = tau_reference, params=(), name: str = "OseenSupgForm") -> OseenPointForm:
    """The stabilised Oseen integrand (test/common.jl:207-220):
        a((u,p),(v,q)) = rho c u.v + 2 mu eps(u):eps(v) + rho ((grad u)'w).v - p div v - q div u
                       + ts ((grad v)'w).(Ru + grad p) - tp grad q.(Ru + grad p) + tl div v div u,
        l((v,q))       = rho f.v + ts ((grad v)'w).rk - tp grad q.rk,
    Ru = rho c u + rho (grad u)'w, rk = rho f + mu lap, with the stabilisation parameters (ts, tp, tl) = tau(D)."""

    def a(U, V, D):
        ts, tp, tl = tau(D)
        rho, mu, c = D.rho, D.mu, D.c0
        Ru = U.u * (rho * c) + U.conv(D.w) * rho
        return (rho * c * dot(U.u, V.u) + 2.0 * mu * cg.ddot(U.eps(), V.eps()) + rho * dot(U.conv(D.w), V.u)
                - U.p * V.div() - V.p * U.div()
                + ts * dot(V.conv(D.w), Ru + U.gp) - tp * dot(V.gp, Ru + U.gp) + tl * V.div() * U.div())

    def l(V, D):
        ts, tp, tl = tau(D)
        rk = D.f * D.rho + D.lap * D.mu
        return D.rho * dot(D.f, V.u) + ts * dot(V.conv(D.w), rk) - tp * dot(V.gp, rk)

    return OseenPointForm(a, l, params=params, name=name)


def check_compiles(form: OseenPointForm) -> None:
    """Compile the form's source with hiprtc (no device needed); ValueError with the compiler log."""
    import ctypes as C

    from . import _lib

    lib = _lib.load()
    log = C.create_string_buffer(1 << 16)
    if lib.hfk_oseen_form_check(form.source().encode(), form.name.encode(), log, len(log)) != 0:
        raise ValueError("HyperFEMKernels: the point form does not compile:\n" + log.value.decode(errors="replace"))
