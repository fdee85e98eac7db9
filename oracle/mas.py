"""Multilevel additive Schwarz apply (FP64 numpy) for a given hierarchy.

Oracle module (test infrastructure only).

Restates the reference CPU path: Galerkin coarse operators by identity aggregation of
the fine blocks (``solver/mas/numeric_assembly.py:64-83``), domain matrices = principal
submatrices over a domain's nodes (``:47-61``), explicit FP64 inverses
(``solver/mas/local_inverse.py:40-78``), and the apply
z = sum_l w_l P_l B_l^-1 R_l r with level weights (1, ..., 1, coarsest)
(``solver/mas/preconditioner.py:26-154``).

Pinned by invariants only (the reference's METIS partition is not reproducible offline):
R_l = P_l^T, every B_l^-1 inverts its domain matrix, the applied operator is symmetric
positive definite, and PCG with it converges to the stated residual.
"""
from __future__ import annotations

import numpy as np


def scalar_map(h, level: int) -> np.ndarray:
    """Scalar index in level ``level`` of every fine scalar (composed identity map)."""
    fine = h.levels[0]
    lv = h.levels[level]
    comp = h.composed[level]
    out = np.empty(fine.num_dofs, dtype=np.int64)
    for node in range(fine.num_nodes):
        d = fine.dims[node]
        out[fine.offsets[node]:fine.offsets[node] + d] = lv.offsets[comp[node]] + np.arange(d)
    return out


def level_matrices(h, A_dense):
    """Dense level operators A_l = R_l A R_l^T."""
    mats = []
    n = A_dense.shape[0]
    for l in range(h.num_levels):
        m = scalar_map(h, l)
        R = np.zeros((h.levels[l].num_dofs, n))
        R[m, np.arange(n)] = 1.0
        mats.append(R @ A_dense @ R.T)
    return mats


def domain_scalars(lv, dom):
    return np.concatenate([lv.offsets[v] + np.arange(lv.dims[v]) for v in dom])


def domain_inverses(h, mats):
    invs = []
    for l, lv in enumerate(h.levels):
        invs.append([np.linalg.inv(mats[l][np.ix_(domain_scalars(lv, d), domain_scalars(lv, d))])
                     for d in lv.domains])
    return invs


def level_weights(h, coarsest_weight=2.5):
    return [1.0] * (h.num_levels - 1) + [coarsest_weight]


def apply(h, invs, r, weights):
    r = np.asarray(r, dtype=np.float64)
    z = np.zeros_like(r)
    for l, lv in enumerate(h.levels):
        m = scalar_map(h, l)
        rl = np.bincount(m, weights=r, minlength=lv.num_dofs)
        cl = np.zeros(lv.num_dofs)
        for d, inv in zip(lv.domains, invs[l]):
            s = domain_scalars(lv, d)
            cl[s] = inv @ rl[s]
        z += weights[l] * cl[m]
    return z


def dense_operator(h, invs, weights):
    n = h.levels[0].num_dofs
    return np.column_stack([apply(h, invs, e, weights) for e in np.eye(n)])
