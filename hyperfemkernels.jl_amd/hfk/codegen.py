"""Pointwise constitutive code generated from an expression tree, as HIP source for ``laws.CompiledLaw``.

The reference captures the integrand as an expression tree and emits a straight-line function from it, one
temporary per operation, memoised on (operation, arguments) (``emit!``, src/codegen.jl:122-163), compiled once per
structure (``PointForm``, :166-198).  Here the host records the constitutive law the same way: a Python function of
the deformation gradient written with the operations below (``Mat3`` products, ``tr``, ``det``, ``inv``, ``log``,
``sqrt``, ...) builds a hash-consed scalar DAG over the nine entries of F and the parameters lam and mu.  From a
strain energy W(F) the stress P = dW/dF and the tangent T = dP/dF are derived symbolically (from a stress P(F),
only T); the DAG is simplified, emitted as one temporary per node, and wrapped in the law interface that
``hfk_law_create`` compiles with hiprtc (setup / stress / tangent).  lam and mu stay runtime arguments, so one
compiled law serves every modulus; other numbers are part of the structure (and of the source hash that keys the
compile cache).

The same DAG evaluates on the host with numpy (``evaluate``), which the CPU tests use to check the derivatives
against finite differences and closed forms.
"""
from __future__ import annotations

import math
from typing import Callable

import numpy as np

# ---------------------------------------------------------------------------------------------------------------
# scalar DAG, hash-consed: equal (op, args) give the same node, so common subexpressions are shared


class Node:
    __slots__ = ("op", "args", "value", "key", "_hash")

    def __init__(self, op, args=(), value=None):
        self.op, self.args, self.value = op, tuple(args), value
        self.key = (op, tuple(id(a) for a in self.args), value)
        self._hash = hash(self.key)

    # arithmetic builds new (shared) nodes
    def __add__(self, o):
        return add(self, wrap(o))

    def __radd__(self, o):
        return add(wrap(o), self)

    def __sub__(self, o):
        return sub(self, wrap(o))

    def __rsub__(self, o):
        return sub(wrap(o), self)

    def __mul__(self, o):
        return mul(self, wrap(o))

    def __rmul__(self, o):
        return mul(wrap(o), self)

    def __truediv__(self, o):
        return div(self, wrap(o))

    def __rtruediv__(self, o):
        return div(wrap(o), self)

    def __neg__(self):
        return neg(self)

    def __pow__(self, k):
        if not isinstance(k, int) or k < 0:
            raise ValueError("HyperFEMKernels: only non-negative integer powers are supported")
        out = const(1.0)
        for _ in range(k):
            out = out * self
        return out


_table: dict = {}


def _node(op, args=(), value=None) -> Node:
    n = Node(op, args, value)
    hit = _table.get(n.key)
    if hit is not None:
        return hit
    _table[n.key] = n
    return n


def const(v: float) -> Node:
    return _node("const", (), float(v))


def var(name: str) -> Node:
    return _node("var", (), name)


def wrap(x) -> Node:
    return x if isinstance(x, Node) else const(x)


def is_const(a: Node, v: float | None = None) -> bool:
    return a.op == "const" and (v is None or a.value == v)


def add(a: Node, b: Node) -> Node:
    if is_const(a, 0.0):
        return b
    if is_const(b, 0.0):
        return a
    if is_const(a) and is_const(b):
        return const(a.value + b.value)
    return _node("add", (a, b))


def sub(a: Node, b: Node) -> Node:
    if is_const(b, 0.0):
        return a
    if is_const(a) and is_const(b):
        return const(a.value - b.value)
    if a is b:
        return const(0.0)
    if is_const(a, 0.0):
        return neg(b)
    return _node("sub", (a, b))


def mul(a: Node, b: Node) -> Node:
    if is_const(a, 0.0) or is_const(b, 0.0):
        return const(0.0)
    if is_const(a, 1.0):
        return b
    if is_const(b, 1.0):
        return a
    if is_const(a) and is_const(b):
        return const(a.value * b.value)
    if is_const(a, -1.0):
        return neg(b)
    if is_const(b, -1.0):
        return neg(a)
    return _node("mul", (a, b))


def div(a: Node, b: Node) -> Node:
    if is_const(b, 0.0):
        raise ZeroDivisionError("HyperFEMKernels: division by the constant 0 in a law")
    if is_const(a, 0.0):
        return const(0.0)
    if is_const(b, 1.0):
        return a
    if is_const(a) and is_const(b):
        return const(a.value / b.value)
    return _node("div", (a, b))


def neg(a: Node) -> Node:
    if is_const(a):
        return const(-a.value)
    if a.op == "neg":
        return a.args[0]
    return _node("neg", (a,))


def log(a) -> Node:
    a = wrap(a)
    if is_const(a):
        return const(math.log(a.value))
    return _node("log", (a,))


def exp(a) -> Node:
    a = wrap(a)
    if is_const(a):
        return const(math.exp(a.value))
    return _node("exp", (a,))


def sqrt(a) -> Node:
    a = wrap(a)
    if is_const(a):
        return const(math.sqrt(a.value))
    return _node("sqrt", (a,))


def diff(f: Node, x: Node, memo: dict | None = None) -> Node:
    """d f / d x (x a variable), memoised per node."""
    memo = {} if memo is None else memo
    k = id(f)
    if k in memo:
        return memo[k]
    op = f.op
    if op == "const":
        d = const(0.0)
    elif op == "var":
        d = const(1.0 if f is x else 0.0)
    else:
        a = f.args[0]
        da = diff(a, x, memo)
        if op == "neg":
            d = neg(da)
        elif op == "log":
            d = div(da, a)
        elif op == "exp":
            d = mul(da, f)
        elif op == "sqrt":
            d = div(da, mul(const(2.0), f))
        else:
            b = f.args[1]
            db = diff(b, x, memo)
            if op == "add":
                d = add(da, db)
            elif op == "sub":
                d = sub(da, db)
            elif op == "mul":
                d = add(mul(da, b), mul(a, db))
            elif op == "div":
                d = sub(div(da, b), div(mul(f, db), b))
            else:
                raise ValueError(f"HyperFEMKernels: cannot differentiate {op}")
    memo[k] = d
    return d


# ---------------------------------------------------------------------------------------------------------------
# 3x3 tensors of nodes, column-major A[i + 3 j] (F[c + 3 J] = d x_c / d X_J, as the law interface)


class Mat3:
    def __init__(self, e):
        e = [wrap(v) for v in e]
        if len(e) != 9:
            raise ValueError("HyperFEMKernels: a Mat3 has 9 entries")
        self.e = e

    def __getitem__(self, ij):
        i, j = ij
        return self.e[i + 3 * j]

    @staticmethod
    def identity():
        return Mat3([1.0 if i == j else 0.0 for j in range(3) for i in range(3)])

    def T(self):
        return Mat3([self[j, i] for j in range(3) for i in range(3)])

    def __matmul__(self, o):
        return Mat3([sum((self[i, k] * o[k, j] for k in range(3)), const(0.0)) for j in range(3) for i in range(3)])

    def __add__(self, o):
        return Mat3([a + b for a, b in zip(self.e, o.e)])

    def __sub__(self, o):
        return Mat3([a - b for a, b in zip(self.e, o.e)])

    def __mul__(self, s):
        s = wrap(s)
        return Mat3([a * s for a in self.e])

    __rmul__ = __mul__


def tr(A: Mat3) -> Node:
    return A[0, 0] + A[1, 1] + A[2, 2]


def det(A: Mat3) -> Node:
    return (A[0, 0] * (A[1, 1] * A[2, 2] - A[1, 2] * A[2, 1]) - A[0, 1] * (A[1, 0] * A[2, 2] - A[1, 2] * A[2, 0]) +
            A[0, 2] * (A[1, 0] * A[2, 1] - A[1, 1] * A[2, 0]))


def inv(A: Mat3) -> Mat3:
    d = det(A)
    c = [[A[(i + 1) % 3, (j + 1) % 3] * A[(i + 2) % 3, (j + 2) % 3] -
          A[(i + 1) % 3, (j + 2) % 3] * A[(i + 2) % 3, (j + 1) % 3] for j in range(3)] for i in range(3)]
    # inverse = adjugate / det, adj[i][j] = cofactor[j][i]
    return Mat3([c[j][i] / d for j in range(3) for i in range(3)])


def ddot(A: Mat3, B: Mat3) -> Node:
    return sum((a * b for a, b in zip(A.e, B.e)), const(0.0))


# ---------------------------------------------------------------------------------------------------------------


class PointLaw:
    """The derived point quantities of a law: P (9 nodes, P[c + 3J]) and T (81 nodes, T[(c + 3J) + 9 (k + 3L)] =
    dP_cJ / dF_kL) over the variables F (9), lam and mu; W the energy node of a law given by its energy (else
    None)."""

    def __init__(self, F, lam, mu, P, T, W=None):
        self.F, self.lam, self.mu, self.P, self.T, self.W = F, lam, mu, P, T, W

    def outputs(self):
        return self.P + self.T + ([self.W] if self.W is not None else [])

    def nodes(self):
        """Every node reachable from the outputs, in an order where arguments come first."""
        seen, order = set(), []
        stack = [(n, False) for n in reversed(self.outputs())]
        while stack:
            n, done = stack.pop()
            if id(n) in seen:
                continue
            if done:
                seen.add(id(n))
                order.append(n)
                continue
            stack.append((n, True))
            for a in reversed(n.args):
                if id(a) not in seen:
                    stack.append((a, False))
        return order

    def evaluate(self, F: np.ndarray, lam: float, mu: float):
        """P (3x3, P[c, J]) and T (3x3x3x3, T[c, J, k, L]) at F (3x3, F[c, J]) with numpy, in FP64."""
        F = np.asarray(F, dtype=np.float64)
        val = {id(self.F.e[i + 3 * j]): F[i, j] for j in range(3) for i in range(3)}
        val[id(self.lam)], val[id(self.mu)] = float(lam), float(mu)
        for n in self.nodes():
            if n.op == "const":
                val[id(n)] = n.value
            elif n.op == "var":
                if id(n) not in val:
                    raise ValueError(f"HyperFEMKernels: unbound variable {n.value}")
            else:
                a = [val[id(x)] for x in n.args]
                val[id(n)] = {"add": lambda: a[0] + a[1], "sub": lambda: a[0] - a[1], "mul": lambda: a[0] * a[1],
                              "div": lambda: a[0] / a[1], "neg": lambda: -a[0], "log": lambda: math.log(a[0]),
                              "exp": lambda: math.exp(a[0]), "sqrt": lambda: math.sqrt(a[0])}[n.op]()
        self._last_energy = val[id(self.W)] if self.W is not None else None
        P = np.array([[val[id(self.P[c + 3 * J])] for J in range(3)] for c in range(3)])
        T = np.zeros((3, 3, 3, 3))
        for c in range(3):
            for J in range(3):
                for k in range(3):
                    for L in range(3):
                        T[c, J, k, L] = val[id(self.T[(c + 3 * J) + 9 * (k + 3 * L)])]
        return P, T

    def energy(self, F: np.ndarray, lam: float, mu: float) -> float:
        """W at F with numpy (a law given by its energy)."""
        if self.W is None:
            raise ValueError("HyperFEMKernels: the law was not given by an energy")
        self.evaluate(F, lam, mu)
        return self._last_energy

    def source(self, name: str) -> str:
        """HIP source of the law struct ``name`` (setup / stress / tangent, and energy for a law given by its energy)
        for ``laws.CompiledLaw``."""
        names = {}
        for j in range(3):
            for i in range(3):
                names[id(self.F.e[i + 3 * j])] = f"Fin[{i + 3 * j}]"
        names[id(self.lam)], names[id(self.mu)] = "lam", "mu"
        body = []

        def ref(n):
            if n.op == "const":
                return repr(float(n.value)) if math.isfinite(n.value) else "(0.0 / 0.0)"
            return names[id(n)]

        for n in self.nodes():
            if n.op in ("const", "var"):
                continue
            a = [ref(x) for x in n.args]
            expr = {"add": f"{a[0]} + {a[-1]}", "sub": f"{a[0]} - {a[-1]}", "mul": f"{a[0]} * {a[-1]}",
                    "div": f"{a[0]} / {a[-1]}", "neg": f"-{a[0]}", "log": f"log({a[0]})", "exp": f"exp({a[0]})",
                    "sqrt": f"sqrt({a[0]})"}[n.op]
            t = f"t{len(body)}"
            names[id(n)] = t
            body.append(f"    const double {t} = {expr};")
        outs = [f"    P[{i}] = {ref(n)};" for i, n in enumerate(self.P)]
        outs += [f"    T[{i}] = {ref(n)};" for i, n in enumerate(self.T)]
        members = "  double P[9], T[81];"
        extra = []
        if self.W is not None:
            outs.append(f"    W = {ref(self.W)};")
            members = "  double P[9], T[81], W;"
            extra = ["  __device__ double energy() const { return W; }"]
        return "\n".join([
            f"struct {name} {{  // generated by hfk.codegen: {len(body)} operations",
            members,
            "  __device__ void setup(const double* Fin, double lam, double mu) {",
            *body, *outs,
            "  }",
            "  __device__ double stress(int c, int J) const { return P[c + 3 * J]; }",
            "  __device__ double tangent(int c, int J, int k, int L) const { return T[(c + 3 * J) + 9 * (k + 3 * L)]; }",
            *extra,
            "};",
            ""])


def _variables():
    F = Mat3([var(f"F{i}") for i in range(9)])
    return F, var("lam"), var("mu")


def from_energy(W: Callable[[Mat3, Node, Node], Node]) -> PointLaw:
    """A law from its strain energy W(F, lam, mu): P = dW/dF, T = dP/dF, both derived from the tree."""
    F, lam, mu = _variables()
    w = wrap(W(F, lam, mu))
    P = [diff(w, F.e[i], {}) for i in range(9)]
    law = _with_tangent(F, lam, mu, P)
    law.W = w
    return law


def from_stress(Pf: Callable[[Mat3, Node, Node], Mat3]) -> PointLaw:
    """A law from its first Piola-Kirchhoff stress P(F, lam, mu): T = dP/dF derived from the tree."""
    F, lam, mu = _variables()
    P = [wrap(v) for v in Pf(F, lam, mu).e]
    return _with_tangent(F, lam, mu, P)


def _with_tangent(F, lam, mu, P) -> PointLaw:
    T = [None] * 81
    for j in range(9):
        memo = {}
        for i in range(9):
            T[i + 9 * j] = diff(P[i], F.e[j], memo)
    return PointLaw(F, lam, mu, P, T)


# the laws of test/common.jl written as energies (templates)
def svk_energy(F: Mat3, lam: Node, mu: Node) -> Node:
    """St Venant-Kirchhoff: W = lam/2 tr(E)^2 + mu E:E, E = (F'F - I)/2 (test/common.jl:2-8)."""
    E = (F.T() @ F - Mat3.identity()) * 0.5
    return lam * 0.5 * tr(E) * tr(E) + mu * ddot(E, E)


def neohooke_energy(F: Mat3, lam: Node, mu: Node) -> Node:
    """Compressible neo-Hooke: W = mu/2 (tr(F'F) - 3) - mu ln J + lam/2 (ln J)^2."""
    lnJ = log(det(F))
    return mu * 0.5 * (tr(F.T() @ F) - 3.0) - mu * lnJ + lam * 0.5 * lnJ * lnJ
