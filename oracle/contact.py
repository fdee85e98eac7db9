"""Dense gradient and Hessian of the contact scene (FP64, CPU).

Oracle module (test infrastructure only).

Every term is evaluated in vertex positions by the oracle's energy restatements
(``oracle/energies.py`` for SNH and inertia, ``oracle/element_energies.py`` for the
membrane, bending, barrier and affine terms), PSD-projected there when its projection
method asks for it (the vertex map x = t + A r is linear, so the reference projects the
position-space Hessian: differentiator.pyx:795-801), and mapped to the targets [moving
positions (3 each), affine matrices (9 each, row-major), translations (3 each)] with
J = dx/dq (dx_j / dA_jl = r_l, dx / dt = I), as
examples/one_bunny_partial_abd/one_bunny_partial_abd.py:127-144 sets them up. Barrier
pairs are dynamic terms whose blocks add to the static ones (matrix.pyx:45-85).
"""
from __future__ import annotations

import numpy as np

from . import energies, projection, element_energies as ee


def dof_layout(cs):
    off_A = 3 * cs.n_free
    off_t = off_A + 9 * cs.n_bodies
    return off_A, off_t, off_t + 3 * cs.n_bodies


def vertex_jacobian(cs, v):
    """(global dofs, J [3, len]) of union vertex v (fixed vertices have none)."""
    off_A, off_t, _ = dof_layout(cs)
    nf, na = cs.n_fixed, cs.n_abd
    if v < nf:
        return np.zeros(0, dtype=np.int64), np.zeros((3, 0))
    if v < nf + na:
        b = cs.body_of[v - nf]
        r = cs.rest[v]
        dofs = np.concatenate([off_A + 9 * b + np.arange(9), off_t + 3 * b + np.arange(3)])
        J = np.zeros((3, 12))
        for j in range(3):
            J[j, 3 * j:3 * j + 3] = r
            J[j, 9 + j] = 1.0
        return dofs, J
    f = v - nf - na
    return 3 * f + np.arange(3), np.eye(3)


def _scatter(cs, g, H, conn, ge, He):
    """Chain rule of per-element position-space (gradient, Hessian) onto the targets."""
    cache = {}
    for e, verts in enumerate(conn):
        dofs, Js = [], []
        for a, v in enumerate(verts):
            v = int(v)
            if v not in cache:
                cache[v] = vertex_jacobian(cs, v)
            d, J = cache[v]
            Jf = np.zeros((3 * len(verts), d.size))
            Jf[3 * a:3 * a + 3] = J
            dofs.append(d)
            Js.append(Jf)
        d = np.concatenate(dofs)
        if d.size == 0:
            continue
        J = np.concatenate(Js, axis=1)
        np.add.at(g, d, J.T @ ge[e])
        Hd = J.T @ He[e] @ J
        np.add.at(H, (d[:, None], d[None, :]), Hd)


def element_terms(cs):
    """Position-space (gradient, Hessian) of the cloth and barrier elements by kind."""
    x = cs.positions()
    out = {}
    R = cs.rest[cs.triangles]
    P = ee.triangle_rest(R[:, 0], R[:, 1], R[:, 2], cs.thickness)
    _, g, H = ee.evaluate("stvk_membrane", x[cs.triangles].reshape(-1, 9), P,
                          [cs.cloth_mu, cs.cloth_lam, cs.dt])
    out["stvk_membrane"] = (g, H)
    Q = cs.rest[cs.hinges]
    P = ee.bending_rest(Q[:, 0], Q[:, 1], Q[:, 2], Q[:, 3])
    _, g, H = ee.evaluate("bending", x[cs.hinges].reshape(-1, 12), P, [cs.bend])
    out["bending"] = (g, H)
    for kind, pairs in (("point_triangle", cs.pt), ("edge_edge", cs.ee)):
        _, g, H = ee.evaluate(kind, x[pairs].reshape(-1, 12), np.zeros((pairs.shape[0], 0)),
                              [cs.dhat, cs.kappa])
        out[kind] = (g, H)
    return out


def assemble(cs, method=2, terms=None, pt_rows=None, ee_rows=None):
    """Dense (gradient, Hessian) over the targets. ``terms``: element_terms(cs), possibly
    precomputed; ``pt_rows`` / ``ee_rows`` select the active pairs (default: all)."""
    if terms is None:
        terms = element_terms(cs)
    _, _, ndof = dof_layout(cs)
    g = np.zeros(ndof)
    H = np.zeros((ndof, ndof))
    x = cs.positions()

    def proj(He):
        return projection.project_batch(He, method) if method > 0 else He

    dinv, vol = energies.snh_rest_data(cs.rest[cs.tets])
    gt, Ht = energies.snh_gradient_hessian(x[cs.tets], dinv, vol, cs.mu, cs.lam, cs.dt)
    _scatter(cs, g, H, cs.tets, gt, proj(Ht))
    gs, Hs = terms["stvk_membrane"]
    _scatter(cs, g, H, cs.triangles, gs, proj(Hs))
    gb, Hb = terms["bending"]
    _scatter(cs, g, H, cs.hinges, gb, proj(Hb))
    for kind, pairs, rows in (("point_triangle", cs.pt, pt_rows), ("edge_edge", cs.ee, ee_rows)):
        rows = np.arange(pairs.shape[0]) if rows is None else np.asarray(rows)
        if rows.size == 0:
            continue
        gp, Hp = terms[kind]
        _scatter(cs, g, H, pairs[rows], gp[rows], proj(Hp[rows]))
    off_A, _, _ = dof_layout(cs)
    _, ga, Ha = ee.evaluate("affine_orthogonality", cs.A.reshape(-1, 9),
                            np.zeros((cs.n_bodies, 0)), [cs.affine_stiffness])
    Ha = proj(Ha)
    for b in range(cs.n_bodies):
        d = off_A + 9 * b + np.arange(9)
        g[d] += ga[b]
        H[np.ix_(d, d)] += Ha[b]
    nf = cs.n_fixed
    vs = np.arange(nf, cs.rest.shape[0])
    gi, Hi = energies.inertia_gradient_hessian(x[vs], cs.last_position[vs], cs.velocity[vs],
                                               cs.mass[vs], cs.dt)
    _scatter(cs, g, H, vs[:, None], gi, Hi)
    return g, H
