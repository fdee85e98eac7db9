"""Closed-form element energies, gradients and Hessians (FP64 numpy).

Oracle module (test infrastructure only).

Stable Neo-Hookean: ``examples/one_bunny/helpers.py:42-67``
  TB rows = rest edges (x1-x0, x2-x0, x3-x0); vol = det(TB^T)/6 (signed);
  IB = (TB^T)^-1; F rows = deformed edges; FI = F^T IB = Ds Dm^-1;
  E = s * vol * (mu/2 (I_C - 3) - mu/2 log(I_C + 1) + lam/2 (J - (1 + 3/4 mu/lam))^2)
  with I_C = tr(FI^T FI), J = det FI and s = dt^2.
Inertia: ``examples/one_bunny/helpers.py:69-71``
  x~ = x_prev + v dt - (0, 9.8 dt^2, 0);  E = 1/2 (x - x~)^T m (x - x~).

Derivatives are derived by hand (not generated):
  P  = mu (1 - 1/(I_C+1)) F + lam (J - alpha) cof(F)
  dP/dF = mu (1 - 1/(I_C+1)) I + 2 mu/(I_C+1)^2 f f^T + lam c c^T + lam (J - alpha) d2J/dF2
with f = vec(F), c = vec(cof F) (column stacking). With g_a (a=0..3) the rows of
G = [-(1^T Dm^-1); Dm^-1] the vertex a / component i derivative of column j of F is
g_a[j], so grad_a = s vol P g_a and H_ab = s vol sum_jn g_a[j] g_b[n] K_(j)(n).
"""
from __future__ import annotations

import numpy as np


def _cross_matrix(v):
    """[v]x for a batch of 3-vectors: [...,3] -> [...,3,3]."""
    z = np.zeros(v.shape[:-1])
    return np.stack([
        np.stack([z, -v[..., 2], v[..., 1]], -1),
        np.stack([v[..., 2], z, -v[..., 0]], -1),
        np.stack([-v[..., 1], v[..., 0], z], -1)], -2)


def snh_rest_data(rest_tet_positions):
    """Per-tet (Dm^-1 [T,3,3], vol [T]) from rest positions [T,4,3].

    Dm has the rest edges as columns (Dm = TB^T of helpers.py:49-52); vol is signed.
    """
    p = np.asarray(rest_tet_positions, dtype=np.float64)
    dm = np.stack([p[:, 1] - p[:, 0], p[:, 2] - p[:, 0], p[:, 3] - p[:, 0]], axis=2)
    vol = np.linalg.det(dm) / 6.0
    return np.linalg.inv(dm), vol


def _snh_terms(tet_positions, dm_inv, mu, lam):
    x = np.asarray(tet_positions, dtype=np.float64)
    ds = np.stack([x[:, 1] - x[:, 0], x[:, 2] - x[:, 0], x[:, 3] - x[:, 0]], axis=2)
    F = ds @ dm_inv                                   # [T,3,3]
    ic = np.einsum("tij,tij->t", F, F)
    J = np.linalg.det(F)
    f0, f1, f2 = F[:, :, 0], F[:, :, 1], F[:, :, 2]
    cof = np.stack([np.cross(f1, f2), np.cross(f2, f0), np.cross(f0, f1)], axis=2)  # columns
    alpha = 1.0 + 0.75 * mu / lam
    return F, ic, J, cof, alpha


def snh_energy(tet_positions, dm_inv, vol, mu, lam, dt):
    F, ic, J, cof, alpha = _snh_terms(tet_positions, dm_inv, mu, lam)
    psi = 0.5 * mu * (ic - 3.0) - 0.5 * mu * np.log(ic + 1.0) + 0.5 * lam * (J - alpha) ** 2
    return dt * dt * vol * psi


def _shape_gradients(dm_inv):
    """G [T,4,3]: g_0 = -(sum of Dm^-1 rows), g_a = row a-1 of Dm^-1."""
    g = np.concatenate([-dm_inv.sum(axis=1, keepdims=True), dm_inv], axis=1)
    return g


def snh_gradient_hessian(tet_positions, dm_inv, vol, mu, lam, dt):
    """Return (gradient [T,12], Hessian [T,12,12]) before any projection.

    Local DoF order is vertex-major: (v0x, v0y, v0z, v1x, ...), the row-major
    flattening of the 4x3 ``tet_position`` attribute.
    """
    F, ic, J, cof, alpha = _snh_terms(tet_positions, dm_inv, mu, lam)
    T = F.shape[0]
    lam_t = np.broadcast_to(np.asarray(lam, dtype=np.float64), (T,))   # scalar or per tet
    s = dt * dt * vol
    a1 = mu * (1.0 - 1.0 / (ic + 1.0))
    a2 = 2.0 * mu / (ic + 1.0) ** 2
    a3 = lam * (J - alpha)
    P = a1[:, None, None] * F + a3[:, None, None] * cof
    G = _shape_gradients(dm_inv)                       # [T,4,3]
    grad = np.einsum("tij,taj->tai", P, G) * s[:, None, None]
    # K as [T, col j, row i, col n, row m] -> blocks K_(j)(n) [3x3]
    f = np.transpose(F, (0, 2, 1))                     # f[:, j] = column j of F
    c = np.transpose(cof, (0, 2, 1))
    K = np.zeros((T, 3, 3, 3, 3))
    eye = np.eye(3)
    for j in range(3):
        K[:, j, :, j, :] += a1[:, None, None] * eye
        for n in range(3):
            K[:, j, :, n, :] += a2[:, None, None] * f[:, j, :, None] * f[:, n, None, :]
            K[:, j, :, n, :] += lam_t[:, None, None] * c[:, j, :, None] * c[:, n, None, :]
    # d2J: block (0,1) = -[f2]x, (0,2) = [f1]x, (1,2) = -[f0]x, transposes below
    cx = [_cross_matrix(f[:, k]) for k in range(3)]
    hj = {(0, 1): -cx[2], (0, 2): cx[1], (1, 2): -cx[0]}
    for (j, n), blk in hj.items():
        K[:, j, :, n, :] += a3[:, None, None] * blk
        K[:, n, :, j, :] += a3[:, None, None] * np.transpose(blk, (0, 2, 1))
    H = np.einsum("taj,tbn,tjink->taibk", G, G, K)    # [T,4,3,4,3]
    H = H.reshape(T, 12, 12) * s[:, None, None]
    return grad.reshape(T, 12), H


def inertia_target(last_position, velocity, dt):
    x = np.asarray(last_position, dtype=np.float64) + np.asarray(velocity, dtype=np.float64) * dt
    x = x.copy()
    x[:, 1] -= 9.8 * dt * dt
    return x


def inertia_energy(position, last_position, velocity, mass, dt):
    d = np.asarray(position, dtype=np.float64) - inertia_target(last_position, velocity, dt)
    return 0.5 * np.asarray(mass, dtype=np.float64) * np.einsum("vi,vi->v", d, d)


def inertia_gradient_hessian(position, last_position, velocity, mass, dt):
    m = np.asarray(mass, dtype=np.float64)
    d = np.asarray(position, dtype=np.float64) - inertia_target(last_position, velocity, dt)
    grad = m[:, None] * d
    H = m[:, None, None] * np.eye(3)[None]
    return grad, H


# ---------------------------------------------------------------------------
# Closed-form block evaluation (used by the timed CPU baseline; pinned against the
# K-tensor contraction of ``snh_gradient_hessian`` in the tests).
# ---------------------------------------------------------------------------
_U4 = np.array([[1.0, 1.0, 1.0], [-1.0, 1.0, 1.0], [0.0, -2.0, 1.0], [0.0, 0.0, -3.0]]) \
    / np.sqrt([2.0, 6.0, 12.0])


def _snh_block_terms(F, cof, ic, J, mu, lam, alpha, G, s):
    a1 = mu * (1.0 - 1.0 / (ic + 1.0))
    a2 = 2.0 * mu / (ic + 1.0) ** 2
    a3 = lam * (J - alpha)
    u = np.einsum("tij,taj->tai", F, G)
    w = np.einsum("tij,taj->tai", cof, G)
    return a1, a2, a3, u, w


def _blocks_from_vectors(F, G, u, w, a1, a2, lam, a3, s):
    """H_ab = s (a1 (g_a.g_b) I + a2 u_a u_b^T + lam w_a w_b^T - a3 [F (g_a x g_b)]x)."""
    T, V = G.shape[:2]
    gg = np.einsum("taj,tbj->tab", G, G)
    H = (a2[:, None, None, None, None] * u[:, :, None, :, None] * u[:, None, :, None, :]
         + lam * w[:, :, None, :, None] * w[:, None, :, None, :])
    H += (a1[:, None, None] * gg)[:, :, :, None, None] * np.eye(3)
    c = np.cross(G[:, :, None, :], G[:, None, :, :])          # [T,V,V,3]
    t = np.einsum("tij,tabj->tabi", F, c)
    z = np.zeros(t.shape[:-1])
    tx = np.stack([np.stack([z, -t[..., 2], t[..., 1]], -1),
                   np.stack([t[..., 2], z, -t[..., 0]], -1),
                   np.stack([-t[..., 1], t[..., 0], z], -1)], -2)
    H -= a3[:, None, None, None, None] * tx
    return H * s[:, None, None, None, None]                   # [T,V,V,3,3]


def snh_gradient_hessian_fast(tet_positions, dm_inv, vol, mu, lam, dt, method=2):
    """Gradient [T,12] and PSD-projected Hessian [T,12,12] with the closed form and the
    translation-free 9x9 reduction of the projection."""
    from .projection import is_psd_batch
    F, ic, J, cof, alpha = _snh_terms(tet_positions, dm_inv, mu, lam)
    G = _shape_gradients(dm_inv)
    s = dt * dt * vol
    a1, a2, a3, u, w = _snh_block_terms(F, cof, ic, J, mu, lam, alpha, G, s)
    grad = (a1[:, None, None] * u + a3[:, None, None] * w) * s[:, None, None]
    T = F.shape[0]
    Hb = _blocks_from_vectors(F, G, u, w, a1, a2, lam, a3, s)
    H = Hb.transpose(0, 1, 3, 2, 4).reshape(T, 12, 12)
    if method > 0:
        bad = ~is_psd_batch(H)
        if bad.any():
            Gc = np.einsum("ak,taj->tkj", _U4, G[bad])
            uc = np.einsum("ak,tai->tki", _U4, u[bad])
            wc = np.einsum("ak,tai->tki", _U4, w[bad])
            Mb = _blocks_from_vectors(F[bad], Gc, uc, wc, a1[bad], a2[bad], lam, a3[bad], s[bad])
            M = Mb.transpose(0, 1, 3, 2, 4).reshape(-1, 9, 9)
            lw, V = np.linalg.eigh(M)
            lw = np.where(lw < 0, np.abs(lw) if method == 1 else 0.0, lw)
            Mp = np.einsum("bik,bk,bjk->bij", V, lw, V).reshape(-1, 3, 3, 3, 3)
            Hp = np.einsum("ak,bl,tkilj->taibj", _U4, _U4, Mp).reshape(-1, 12, 12)
            H[bad] = Hp
    return grad.reshape(T, 12), H
