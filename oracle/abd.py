"""Dense gradient and Hessian of the affine-body cube scene (FP64, CPU).

Oracle module (test infrastructure only).

Restates the chain rule of the reference's affine-body setup
(examples/one_bunny_partial_abd/one_bunny_partial_abd.py:127-144, vertex x = t + A r
through a JOIN): every term is evaluated in vertex positions, PSD-projected there when
its projection method asks for it (the map is linear: differentiator.pyx:795-801), and
mapped to the targets [free positions (3 each), affine matrices (9 each, row-major),
translations (3 each)] with J = dx/dq (dx_j / dA_jl = r_l, dx / dt = I).
"""
from __future__ import annotations

import numpy as np

from . import energies, projection, element_energies


def dof_layout(sc):
    nfree, nb = sc.n_free, sc.n_bodies
    off_A = 3 * nfree
    off_t = off_A + 9 * nb
    return off_A, off_t, off_t + 3 * nb


def vertex_jacobian(sc, v):
    """(global dofs, J [3, len]) of vertex v (fixed vertices have no dofs)."""
    off_A, off_t, _ = dof_layout(sc)
    nf, na = sc.n_fixed, sc.n_abd
    if v < nf:
        return np.zeros(0, dtype=np.int64), np.zeros((3, 0))
    if v < nf + na:
        b = sc.body_of[v - nf]
        r = sc.cube.rest[v]
        dofs = np.concatenate([off_A + 9 * b + np.arange(9), off_t + 3 * b + np.arange(3)])
        J = np.zeros((3, 12))
        for j in range(3):
            J[j, 3 * j:3 * j + 3] = r
            J[j, 9 + j] = 1.0
        return dofs, J
    f = v - nf - na
    return 3 * f + np.arange(3), np.eye(3)


def assemble(sc, method=2):
    cs = sc.cube
    _, _, ndof = dof_layout(sc)
    g = np.zeros(ndof)
    H = np.zeros((ndof, ndof))
    x = sc.positions()
    dinv, vol = energies.snh_rest_data(cs.rest[cs.tets])
    gt, Ht = energies.snh_gradient_hessian(x[cs.tets], dinv, vol, cs.mu, cs.lam, cs.dt)
    if method > 0:
        Ht = projection.project_batch(Ht, method)
    for e, tet in enumerate(cs.tets):
        dofs, Js = [], []
        for a, v in enumerate(tet):
            d, J = vertex_jacobian(sc, v)
            Jf = np.zeros((12, d.size))
            Jf[3 * a:3 * a + 3] = J
            dofs.append(d)
            Js.append(Jf)
        d = np.concatenate(dofs)
        J = np.concatenate(Js, axis=1)
        np.add.at(g, d, J.T @ gt[e])
        np.add.at(H, (d[:, None], d[None, :]), J.T @ Ht[e] @ J)   # d repeats a body's dofs
    # inertia on abd and free vertices (projection -1: none)
    nf = sc.n_fixed
    gi, Hi = energies.inertia_gradient_hessian(x[nf:], cs.last_position[nf:], cs.velocity[nf:],
                                               cs.mass[nf:], cs.dt)
    for k in range(x.shape[0] - nf):
        d, J = vertex_jacobian(sc, nf + k)
        np.add.at(g, d, J.T @ gi[k])
        np.add.at(H, (d[:, None], d[None, :]), J.T @ Hi[k] @ J)
    # affine orthogonality per body (A row-major)
    off_A, _, _ = dof_layout(sc)
    Eg, ga, Ha = element_energies.evaluate("affine_orthogonality", sc.A.reshape(-1, 9),
                                           np.zeros((sc.n_bodies, 0)), [sc.affine_stiffness])
    if method > 0:
        Ha = projection.project_batch(Ha, method)
    for b in range(sc.n_bodies):
        d = off_A + 9 * b + np.arange(9)
        g[d] += ga[b]
        H[np.ix_(d, d)] += Ha[b]
    return g, H
