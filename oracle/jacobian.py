"""Mixed second derivatives d g / d theta of the element energies (FP64 numpy).

Oracle module (test infrastructure only).

Restates what the reference's rectangular second-order Jacobians compute for the energies
of its adjoint example (``differentiator/differentiator.pyx:367-489``,
``kernel/Jacobian/secondOrderJacobianKernelNoProject.pyx``, used by
``examples/bunny_material_optimization/bunny_material_optimization.py:605-618``):

* Stable Neo-Hookean (``examples/one_bunny/helpers.py:42-67``),
  E = s vol (mu/2 (I_C - 3) - mu/2 log(I_C + 1) + lam/2 (J - alpha)^2), alpha = 1 + 3/4 mu/lam,
  s = dt^2. By hand, with dI_C = d I_C / d x = 2 F g_a and dJ = d J / d x = cof(F) g_a:
    d g / d lam = s vol (J - 1) dJ                 (the alpha terms cancel)
    d g / d mu  = s vol (1/2 (1 - 1/(I_C + 1)) dI_C - 3/4 dJ)
    d g / d dt  = 2 g / dt
* Lamé coefficients of the example (``bunny_material_optimization.py:355-359``),
  mu = Y / (2 (1 + nu)), lam = Y nu / ((1 + nu)(1 - 2 nu)), and their derivatives
    d mu / d Y = 1 / (2 (1 + nu)),            d mu / d nu = -Y / (2 (1 + nu)^2),
    d lam / d Y = nu / ((1 + nu)(1 - 2 nu)),  d lam / d nu = Y (1 + 2 nu^2) / ((1 + nu)(1 - 2 nu))^2.
* Inertia (``helpers.py:69-71``), g = m (x - x~), x~ = x_prev + v dt - (0, 9.8 dt^2, 0):
    d g / d m = x - x~,  d g / d x_prev = -m I,  d g / d v = -m dt I,
    d g / d dt = -m (v - (0, 19.6 dt, 0));
  with the example's velocity v = (x_prev - x_prev2) / dt (``:340``) the chain rule gives
    d g / d x_prev = -2 m I,  d g / d x_prev2 = m I.
* The generated kinds (IPC barriers, bending, Baraff-Witkin, StVK, affine orthogonality):
  d g / d s from the SymPy gradient of ``oracle.element_energies`` (exact differentiation,
  float64 evaluation).

Pinned in ``tests/test_jacobian_host.py`` by central differences of the gradients w.r.t.
each parameter, by homogeneity (g is linear in (mu, lam) jointly, in Y, in m, in kappa, in
the bending stiffness, in (ks, kh) and in (mu, lam) of StVK: sum theta d g / d theta = g)
and by the closed forms above for the cases they reduce to.
"""
from __future__ import annotations

from functools import lru_cache

import numpy as np

from . import element_energies as ee
from .energies import _shape_gradients, _snh_terms, inertia_target


# ------------------------------------------------------------------ Lamé coefficients
def lame(young, poisson):
    y = np.asarray(young, dtype=np.float64)
    nu = np.asarray(poisson, dtype=np.float64)
    return y / (2.0 * (1.0 + nu)), y * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))


def lame_derivatives(young, poisson):
    """(d mu / d Y, d mu / d nu, d lam / d Y, d lam / d nu)."""
    y = np.asarray(young, dtype=np.float64)
    nu = np.asarray(poisson, dtype=np.float64)
    d = (1.0 + nu) * (1.0 - 2.0 * nu)
    return (1.0 / (2.0 * (1.0 + nu)), -y / (2.0 * (1.0 + nu) ** 2), nu / d,
            y * (1.0 + 2.0 * nu * nu) / (d * d))


# ------------------------------------------------------------------ SNH
def snh_param_gradients(tet_positions, dm_inv, vol, mu, lam, dt):
    """(d g / d mu, d g / d lam, d g / d dt), each [T, 12] (vertex-major local DoF)."""
    F, ic, J, cof, _ = _snh_terms(tet_positions, dm_inv, mu, lam)
    G = _shape_gradients(dm_inv)                               # [T, 4, 3]
    dic = 2.0 * np.einsum("tij,taj->tai", F, G)                # d I_C / d x_a
    dj = np.einsum("tij,taj->tai", cof, G)                     # d J / d x_a
    s = dt * dt * np.asarray(vol, dtype=np.float64)
    T = F.shape[0]
    mu = np.broadcast_to(np.asarray(mu, dtype=np.float64), (T,))
    lam = np.broadcast_to(np.asarray(lam, dtype=np.float64), (T,))
    g_lam = (s * (J - 1.0))[:, None, None] * dj
    g_mu = s[:, None, None] * (0.5 * (1.0 - 1.0 / (ic + 1.0))[:, None, None] * dic - 0.75 * dj)
    alpha = 1.0 + 0.75 * mu / lam
    grad = s[:, None, None] * ((0.5 * mu * (1.0 - 1.0 / (ic + 1.0)))[:, None, None] * dic
                               + (lam * (J - alpha))[:, None, None] * dj)
    g_dt = 2.0 / dt * grad
    return g_mu.reshape(T, 12), g_lam.reshape(T, 12), g_dt.reshape(T, 12)


# ------------------------------------------------------------------ inertia
def inertia_param_gradients(position, last_position, velocity, mass, dt):
    """Dict of per-vertex derivatives of g = m (x - x~): "mass" [V,3], "x_prev" [V,3,3],
    "velocity" [V,3,3], "dt" [V,3]."""
    x = np.asarray(position, dtype=np.float64)
    v = np.asarray(velocity, dtype=np.float64)
    m = np.broadcast_to(np.asarray(mass, dtype=np.float64), (x.shape[0],))
    d = x - inertia_target(last_position, velocity, dt)
    eye = np.eye(3)[None]
    dv = v.copy()
    dv[:, 1] -= 19.6 * dt
    return {"mass": d, "x_prev": -m[:, None, None] * eye, "velocity": -(m * dt)[:, None, None] * eye,
            "dt": -m[:, None] * dv}


# ------------------------------------------------------------------ generated kinds
@lru_cache(maxsize=None)
def _gradient(kind):
    import sympy as sp
    arity, width, npar, nsc = ee.SHAPES[kind]
    n = arity * width
    x = sp.symbols(f"x0:{n}", real=True)
    p = sp.symbols(f"p0:{max(npar, 1)}", real=True)
    s = sp.symbols(f"s0:{nsc}", real=True)
    E = ee._sym_energy(kind, x, p, s)
    return sp.lambdify((x, p, s), [sp.diff(E, xi) for xi in x], "numpy")


def element_gradients(kind, X, P, s):
    """Element gradients [n, N] of a generated kind (the SymPy gradient of
    ``oracle.element_energies``, without its Hessian), scene parameters chosen per
    element as the kernels do."""
    arity, width, npar, nsc = ee.SHAPES[kind]
    X = np.asarray(X, dtype=np.float64)
    n = X.shape[0]
    P = np.zeros((n, 1)) if P is None or npar == 0 else np.asarray(P, dtype=np.float64)
    f = _gradient(kind)
    out = np.zeros((n, arity * width))
    for e in range(n):
        se = ee.scene_params(kind, X[e:e + 1], P[e:e + 1], np.asarray(s, dtype=np.float64))
        out[e] = np.array(f(X[e], P[e], np.asarray(se, dtype=np.float64).reshape(-1)),
                          dtype=np.float64).reshape(-1)
    return out


@lru_cache(maxsize=None)
def _scene_jacobian(kind):
    import sympy as sp
    arity, width, npar, nsc = ee.SHAPES[kind]
    n = arity * width
    x = sp.symbols(f"x0:{n}", real=True)
    p = sp.symbols(f"p0:{max(npar, 1)}", real=True)
    s = sp.symbols(f"s0:{nsc}", real=True)
    E = ee._sym_energy(kind, x, p, s)
    G = [[sp.diff(sp.diff(E, xi), sj) for sj in s] for xi in x]
    return sp.lambdify((x, p, s), G, "numpy")


def element_scene_gradients(kind, X, P, s):
    """d g / d s [n, N, S] of a generated kind at element vectors X [n, N], element
    parameters P [n, npar] and scene parameters s [S] (Baraff-Witkin's stretch
    coefficients chosen per element as the kernels do)."""
    arity, width, npar, nsc = ee.SHAPES[kind]
    X = np.asarray(X, dtype=np.float64)
    n = X.shape[0]
    P = np.zeros((n, 1)) if P is None or npar == 0 else np.asarray(P, dtype=np.float64)
    f = _scene_jacobian(kind)
    out = np.zeros((n, arity * width, nsc))
    for e in range(n):
        se = ee.scene_params(kind, X[e:e + 1], P[e:e + 1], np.asarray(s, dtype=np.float64))
        se = np.asarray(se, dtype=np.float64).reshape(-1)
        out[e] = np.array(f(X[e], P[e], se), dtype=np.float64).reshape(arity * width, nsc)
    return out


# ------------------------------------------------------------------ dense assembly
def dense_jacobian(element_rows, blocks, element_cols, n_rows, n_cols):
    """Dense [n_rows, n_cols] sum of element blocks [n, N, w]: DoF r of element e at row
    element_rows[e, r] (-1: not a row), column c at element_cols[e] + c."""
    J = np.zeros((n_rows, n_cols))
    rows = np.asarray(element_rows)
    cols = np.asarray(element_cols)
    B = np.asarray(blocks, dtype=np.float64)
    n, N, w = B.shape
    for e in range(n):
        for r in range(N):
            if rows[e, r] < 0:
                continue
            J[rows[e, r], cols[e]:cols[e] + w] += B[e, r]
    return J
