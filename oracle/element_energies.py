"""IPC barriers, bending, Baraff-Witkin, StVK membrane and affine orthogonality (FP64).

Oracle module (test infrastructure only).

Two restatements of every energy, pinned against each other and against finite
differences, closed forms and invariances in ``tests/test_oracle_energies.py``:
* ``value_<kind>``: the energy written directly in numpy from the reference definitions
  (examples/one_bunny/helpers.py:74-120, examples/many_bunnies_one_cloth/helpers.py:
  193-259, examples/one_bunny_partial_abd/helpers.py:132-139; StVK membrane has no
  reference counterpart);
* ``derivatives(kind)``: the same energy as a SymPy expression, differentiated exactly
  and evaluated in float64 (numpy lambdify).
Element vectors x are [n, arity * width] (vertex rows concatenated), element parameters
p [n, P], scene parameters s [S] (Baraff-Witkin's two stretch coefficients are chosen per
element by ``scene_params``).
"""
from __future__ import annotations

from functools import lru_cache

import numpy as np

KINDS = ["point_point", "point_edge", "point_triangle", "edge_edge", "bending",
         "baraff_witkin", "stvk_membrane", "affine_orthogonality"]
# kind: (arity, width, element params, scene params)
SHAPES = {
    "point_point": (2, 3, 0, 2),
    "point_edge": (3, 3, 0, 2),
    "point_triangle": (4, 3, 0, 2),
    "edge_edge": (4, 3, 0, 2),
    "bending": (4, 3, 2, 1),
    "baraff_witkin": (3, 3, 5, 5),
    "stvk_membrane": (3, 3, 5, 3),
    "affine_orthogonality": (1, 9, 0, 1),
}


# ------------------------------------------------------------------ numpy restatement
def _rows(x, k):
    return x[:, 3 * k:3 * k + 3]


def _barrier(d, s):
    dhat, kappa = s[0], s[1]
    return kappa * (d - dhat) ** 2 * np.log(d / dhat) ** 2


def _dot(a, b):
    return np.einsum("ij,ij->i", a, b)


def value_point_point(x, p, s):
    e = _rows(x, 1) - _rows(x, 0)
    return _barrier(_dot(e, e), s)


def value_point_edge(x, p, s):
    p0, p1, p2 = _rows(x, 0), _rows(x, 1), _rows(x, 2)
    c = np.cross(p1 - p0, p2 - p0)
    e = p2 - p1
    return _barrier(_dot(c, c) / _dot(e, e), s)


def value_point_triangle(x, p, s):
    p0, p1, p2, p3 = (_rows(x, k) for k in range(4))
    b = np.cross(p2 - p1, p3 - p1)
    atb = _dot(p0 - p1, b)
    return _barrier(atb * atb / _dot(b, b), s)


def value_edge_edge(x, p, s):
    p0, p1, p2, p3 = (_rows(x, k) for k in range(4))
    b = np.cross(p1 - p0, p3 - p2)
    atb = _dot(p2 - p0, b)
    return _barrier(atb * atb / _dot(b, b), s)


def edge_theta(q0, q1, q2, q3):
    """Dihedral angle of helpers.py:193-203 (batched rows)."""
    n0 = np.cross(q0 - q2, q1 - q2)
    n1 = np.cross(q1 - q3, q0 - q3)
    axis = q1 - q0
    num = _dot(np.cross(n0, n1), axis) / np.linalg.norm(axis, axis=1)
    den = _dot(n0, n1) + np.linalg.norm(n0, axis=1) * np.linalg.norm(n1, axis=1)
    return 2.0 * np.arctan2(num, den)


def value_bending(x, p, s):
    t = edge_theta(*(_rows(x, k) for k in range(4)))
    return s[0] * (t - p[:, 0]) ** 2 * p[:, 1]


def membrane_f(x, p):
    """F = [x1 - x0, x2 - x0] Minv, [n, 3, 2]."""
    D = np.stack([_rows(x, 1) - _rows(x, 0), _rows(x, 2) - _rows(x, 0)], axis=2)
    Minv = p[:, :4].reshape(-1, 2, 2)
    return D @ Minv


def value_baraff_witkin(x, p, s):
    F = membrane_f(x, p)
    i5u = np.linalg.norm(F[:, :, 0], axis=1)
    i5v = np.linalg.norm(F[:, :, 1], axis=1)
    i6 = _dot(F[:, :, 0], F[:, :, 1])
    cu = np.where(i5u < 1.0, 0.01, 1.0)
    cv = np.where(i5v < 1.0, 0.01, 1.0)
    stretch = cu * (i5u - 1.0) ** 2 + cv * (i5v - 1.0) ** 2
    return (s[0] * stretch + s[1] * i6 * i6) * p[:, 4] * s[2] * s[2]


def value_stvk_membrane(x, p, s):
    F = membrane_f(x, p)
    E = 0.5 * (np.transpose(F, (0, 2, 1)) @ F - np.eye(2))
    tr = E[:, 0, 0] + E[:, 1, 1]
    return p[:, 4] * s[2] * s[2] * (s[0] * (E ** 2).sum(axis=(1, 2)) + 0.5 * s[1] * tr * tr)


def value_affine_orthogonality(x, p, s):
    A = x[:, :9].reshape(-1, 3, 3)
    D = np.transpose(A, (0, 2, 1)) @ A - np.eye(3)
    return s[0] * 0.5 * (D ** 2).sum(axis=(1, 2))


VALUES = {k: globals()[f"value_{k}"] for k in KINDS}


def scene_params(kind, x, p, s):
    """Per-element scene parameters [n, S] (Baraff-Witkin stretch branch coefficients)."""
    s = np.asarray(s, dtype=np.float64)
    out = np.tile(s, (x.shape[0], 1))
    if kind == "baraff_witkin":
        F = membrane_f(x, p)
        out[:, 3] = np.where(np.linalg.norm(F[:, :, 0], axis=1) < 1.0, 0.01, 1.0)
        out[:, 4] = np.where(np.linalg.norm(F[:, :, 1], axis=1) < 1.0, 0.01, 1.0)
    return out


# ------------------------------------------------------------------ rest data
def bending_rest(q0, q1, q2, q3):
    return np.stack([edge_theta(q0, q1, q2, q3), np.linalg.norm(q1 - q0, axis=1)], axis=1)


def triangle_rest(v0, v1, v2, thickness):
    """[Minv (4), thickness * area] with the rest frame of helpers.py:214-237."""
    n = np.cross(v1 - v0, v2 - v0)
    ln = np.linalg.norm(n, axis=1)
    n = n / ln[:, None]
    tgt = np.array([0.0, 1.0, 0.0])
    vec = np.cross(n, tgt)
    cs = n @ tgt
    K = np.zeros((len(n), 3, 3))
    K[:, 0, 1], K[:, 0, 2] = -vec[:, 2], vec[:, 1]
    K[:, 1, 0], K[:, 1, 2] = vec[:, 2], -vec[:, 0]
    K[:, 2, 0], K[:, 2, 1] = -vec[:, 1], vec[:, 0]
    R = np.eye(3) + K + (K @ K) / (1.0 + cs)[:, None, None]
    uv = [np.einsum("bij,bj->bi", R, v)[:, [0, 2]] for v in (v0, v1, v2)]
    M = np.stack([np.stack([uv[1][:, 0] - uv[0][:, 0], uv[2][:, 0] - uv[0][:, 0]], 1),
                  np.stack([uv[1][:, 1] - uv[0][:, 1], uv[2][:, 1] - uv[0][:, 1]], 1)], 1)
    Minv = np.linalg.inv(M)
    return np.concatenate([Minv.reshape(-1, 4), (thickness * ln / 2.0)[:, None]], axis=1)


# ------------------------------------------------------------------ SymPy restatement
def _sym_energy(kind, x, p, s):
    import sympy as sp

    def v(i):
        return sp.Matrix(x[3 * i:3 * i + 3])

    def bar(d):
        return s[1] * (d - s[0]) ** 2 * sp.log(d / s[0]) ** 2

    if kind == "point_point":
        e = v(1) - v(0)
        return bar(e.dot(e))
    if kind == "point_edge":
        c = (v(1) - v(0)).cross(v(2) - v(0))
        e = v(2) - v(1)
        return bar(c.dot(c) / e.dot(e))
    if kind == "point_triangle":
        b = (v(2) - v(1)).cross(v(3) - v(1))
        a = (v(0) - v(1)).dot(b)
        return bar(a * a / b.dot(b))
    if kind == "edge_edge":
        b = (v(1) - v(0)).cross(v(3) - v(2))
        a = (v(2) - v(0)).dot(b)
        return bar(a * a / b.dot(b))
    if kind == "bending":
        q0, q1, q2, q3 = (v(i) for i in range(4))
        n0 = (q0 - q2).cross(q1 - q2)
        n1 = (q1 - q3).cross(q0 - q3)
        ax = q1 - q0
        t = 2 * sp.atan2(n0.cross(n1).dot(ax) / sp.sqrt(ax.dot(ax)),
                         n0.dot(n1) + sp.sqrt(n0.dot(n0)) * sp.sqrt(n1.dot(n1)))
        return s[0] * (t - p[0]) ** 2 * p[1]
    if kind in ("baraff_witkin", "stvk_membrane"):
        D = sp.Matrix.hstack(v(1) - v(0), v(2) - v(0))
        F = D * sp.Matrix([[p[0], p[1]], [p[2], p[3]]])
        if kind == "baraff_witkin":
            fu, fv = F[:, 0], F[:, 1]
            st = s[3] * (sp.sqrt(fu.dot(fu)) - 1) ** 2 + s[4] * (sp.sqrt(fv.dot(fv)) - 1) ** 2
            return (s[0] * st + s[1] * fu.dot(fv) ** 2) * p[4] * s[2] ** 2
        E = (F.T * F - sp.eye(2)) / 2
        tr = E[0, 0] + E[1, 1]
        return p[4] * s[2] ** 2 * (s[0] * sum(E[i, j] ** 2 for i in range(2) for j in range(2))
                                   + s[1] / 2 * tr * tr)
    if kind == "affine_orthogonality":
        A = sp.Matrix(3, 3, list(x[:9]))
        D = A.T * A - sp.eye(3)
        return s[0] / 2 * sum(D[i, j] ** 2 for i in range(3) for j in range(3))
    raise KeyError(kind)


@lru_cache(maxsize=None)
def derivatives(kind):
    """(E, grad, Hessian) evaluators: f(x [n, N], p [n, P], s [n, S]) -> arrays."""
    import sympy as sp
    arity, width, npar, nsc = SHAPES[kind]
    n = arity * width
    x = sp.symbols(f"x0:{n}", real=True)
    p = sp.symbols(f"p0:{max(npar, 1)}", real=True)
    s = sp.symbols(f"s0:{nsc}", real=True)
    E = _sym_energy(kind, x, p, s)
    g = [sp.diff(E, xi) for xi in x]
    H = [[sp.diff(gi, xj) for xj in x] for gi in g]
    args = list(x) + list(p) + list(s)
    fE = sp.lambdify(args, E, "numpy")
    fg = sp.lambdify(args, g, "numpy")
    fH = sp.lambdify(args, H, "numpy")

    def call(f, X, P, S, shape):
        m = X.shape[0]
        P = np.zeros((m, 1)) if npar == 0 else P
        cols = [X[:, i] for i in range(n)] + [P[:, i] for i in range(max(npar, 1))] + \
            [S[:, i] for i in range(nsc)]
        out = f(*cols)
        return np.broadcast_to(np.asarray(np.broadcast_arrays(*_flatten(out)), dtype=np.float64).T,
                               (m,) + (int(np.prod(shape)),)).reshape((m,) + shape) if shape else \
            np.broadcast_to(np.asarray(out, dtype=np.float64), (m,)).copy()

    return (lambda X, P, S: call(fE, X, P, S, ()),
            lambda X, P, S: call(fg, X, P, S, (n,)),
            lambda X, P, S: call(fH, X, P, S, (n, n)))


def _flatten(out):
    if isinstance(out, (list, tuple)):
        res = []
        for o in out:
            res.extend(_flatten(o))
        return res
    return [out]


def evaluate(kind, X, P, s):
    """(E [n], g [n, N], H [n, N, N]) in FP64 for elements X with parameters P."""
    S = scene_params(kind, X, P, s)
    fE, fg, fH = derivatives(kind)
    return fE(X, P, S), fg(X, P, S), fH(X, P, S)
