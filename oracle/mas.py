"""Multilevel additive Schwarz apply (FP64 numpy) for a given hierarchy.

Oracle module (test infrastructure only).

Restates the reference CPU path: Galerkin coarse operators by identity aggregation of
the fine blocks (``solver/mas/numeric_assembly.py:64-83``), domain matrices = principal
submatrices over a domain's nodes (``:47-61``), explicit FP64 inverses
(``solver/mas/local_inverse.py:40-78``), and the apply
z = sum_l w_l P_l B_l^-1 R_l r with level weights (1, ..., 1, coarsest)
(``solver/mas/preconditioner.py:26-154``), optionally with the GPU path's topology
weight per domain (``duplicate_domains``: a domain that only repeats a domain of the
previous level, ``solver/mas/cuda_runtime.py:1285-1340``, weighted by
``duplicate_level_weight`` in the collect, ``cuda/preconditioner_specialized.cu:415-440``).

Pinned by invariants (the reference's METIS partition is not reproducible offline):
R_l = P_l^T, every B_l^-1 inverts its domain matrix, the applied operator is symmetric
positive definite, and PCG with it converges to the stated residual; and by closed forms
on one-domain hierarchies (z = w A^-1 r; a second level of one coarse node adds
w R^T (R A R^T)^-1 R r) and a hand-built hierarchy for the duplicate rule.
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


def apply(h, invs, r, weights, scales=None):
    """``scales``: optional per-level lists of a factor per domain (topology weights)."""
    r = np.asarray(r, dtype=np.float64)
    z = np.zeros_like(r)
    for l, lv in enumerate(h.levels):
        m = scalar_map(h, l)
        rl = np.bincount(m, weights=r, minlength=lv.num_dofs)
        cl = np.zeros(lv.num_dofs)
        for k, (d, inv) in enumerate(zip(lv.domains, invs[l])):
            s = domain_scalars(lv, d)
            cl[s] = inv @ rl[s] * (1.0 if scales is None else scales[l][k])
        z += weights[l] * cl[m]
    return z


def duplicate_domains(h):
    """Per level, per domain: the domain repeats one domain of the previous level (every
    node has one child, and the children are exactly that domain's nodes). Plain loops
    over the hierarchy's maps (``solver/mas/cuda_runtime.py:1301-1337``)."""
    out = [[False] * len(h.levels[0].domains)]
    for l in range(1, len(h.levels)):
        prev, cur = h.levels[l - 1], h.levels[l]
        children = [[] for _ in range(cur.num_nodes)]
        for child, parent in enumerate(prev.fine_to_parent):
            children[int(parent)].append(child)
        owner = {}
        for k, d in enumerate(prev.domains):
            for node in d:
                owner[int(node)] = k
        flags = []
        for d in cur.domains:
            kids = []
            ok = True
            for parent in d:
                if len(children[int(parent)]) != 1:
                    ok = False
                    break
                kids.append(children[int(parent)][0])
            if ok:
                src = owner[kids[0]]
                ok = all(owner[c] == src for c in kids) and \
                    set(kids) == {int(v) for v in prev.domains[src]}
            flags.append(ok)
        out.append(flags)
    return out


def dense_operator(h, invs, weights):
    n = h.levels[0].num_dofs
    return np.column_stack([apply(h, invs, e, weights) for e in np.eye(n)])
