"""PSD test, PSD projection and block-Jacobi pseudo-inverse (FP64 numpy).

Oracle module (test infrastructure only).

* ``is_psd``: ``kernel/Hessian/hessianKernelHeader.pyx:79-141`` - a packed
  lower-triangular Schur-complement (LDL^T) sweep. tolerance = 1e-10 * max(1, max|a_ii|);
  a pivot below -tol fails; a pivot within +-tol requires every remaining entry of its
  column to be within tol; otherwise the trailing block is updated with
  fma(-a_rc / pivot, a_tc, a_rt) using the correctly rounded reciprocal of the pivot.
* ``project``: ``hessianKernelHeader.pyx:143-252`` - method 0 returns A, PSD inputs are
  returned unchanged, otherwise eigenvalues below zero become |l| (method 1) or 0
  (method 2). N==1 with method 2 clamps to 1e-6, method 1 takes |a| (``:149-152``).
  Matrices with N<4 are padded to 4x4 with identity before the EVD (``:159-169``),
  which leaves the leading block's projection unchanged.
* ``jacobi_block_inverse``: ``kernel/Solver/diagonalBlockInverseKernel.pyx:51-104`` -
  eigenvalues with |l| < 1e-6 are kept as |l| (not inverted), others become |1/l|.
"""
from __future__ import annotations

import numpy as np


def is_psd(A) -> bool:
    """Restatement of the reference packed Schur PSD test for one dense N x N matrix.

    Reads only the lower triangle (A[row, col] for col <= row), as the reference does.
    """
    A = np.asarray(A, dtype=np.float64)
    n = A.shape[0]
    schur = [[float(A[r, c]) for c in range(r + 1)] for r in range(n)]
    scale = 1.0
    for r in range(n):
        scale = max(scale, abs(schur[r][r]))
    tol = scale * 1.0e-10
    for c in range(n):
        pivot = schur[c][c]
        if pivot < -tol:
            return False
        if abs(pivot) <= tol:
            for r in range(c + 1, n):
                if abs(schur[r][c]) > tol:
                    return False
            continue
        inv = 1.0 / pivot
        for r in range(c + 1, n):
            factor = schur[r][c] * inv
            for t in range(c + 1, r + 1):
                # fma(-factor, schur[t][c], schur[r][t]); numpy has no scalar fma, the
                # product is exact enough for the decision except on the knife edge.
                schur[r][t] = schur[r][t] - factor * schur[t][c]
    return True


def is_psd_batch(A):
    """Vectorised ``is_psd`` over a batch [B,N,N] (same sweep, numpy masks)."""
    S = np.tril(np.asarray(A, dtype=np.float64)).copy()
    B, n, _ = S.shape
    diag = np.abs(np.diagonal(S, axis1=1, axis2=2))
    tol = np.maximum(1.0, diag.max(axis=1)) * 1.0e-10
    ok = np.ones(B, dtype=bool)
    for c in range(n):
        pivot = S[:, c, c]
        ok &= ~(pivot < -tol)
        small = np.abs(pivot) <= tol
        below = np.abs(S[:, c + 1:, c]) > tol[:, None]
        ok &= ~(small & below.any(axis=1))
        upd = ~small
        inv = np.where(upd, 1.0 / np.where(upd, pivot, 1.0), 0.0)
        factor = S[:, c + 1:, c] * inv[:, None]                # [B, n-c-1]
        # trailing lower block rows r>c, cols t in (c, r]
        S[:, c + 1:, c + 1:] -= np.tril(factor[:, :, None] * S[:, c + 1:, c][:, None, :])
    return ok


def project(A, method: int, psd_check: bool = True):
    """Project one symmetric matrix with the reference clamp rules."""
    A = np.asarray(A, dtype=np.float64)
    n = A.shape[0]
    if method == 0 or method < 0:
        return A.copy()
    if n == 1:
        a = A[0, 0]
        return np.array([[abs(a) if method == 1 else (1e-6 if a < 1e-6 else a)]])
    if psd_check and is_psd(A):
        return A.copy()
    w, V = np.linalg.eigh(A)
    w = np.where(w < 0, np.abs(w) if method == 1 else 0.0, w)
    return (V * w) @ V.T


def project_batch(A, method: int, psd_check: bool = True):
    A = np.asarray(A, dtype=np.float64)
    if method <= 0:
        return A.copy()
    n = A.shape[1]
    if n == 1:
        a = A[:, 0, 0]
        v = np.abs(a) if method == 1 else np.where(a < 1e-6, 1e-6, a)
        return v[:, None, None].copy()
    out = A.copy()
    todo = ~is_psd_batch(A) if psd_check else np.ones(A.shape[0], dtype=bool)
    if todo.any():
        w, V = np.linalg.eigh(A[todo])
        w = np.where(w < 0, np.abs(w) if method == 1 else 0.0, w)
        out[todo] = np.einsum("bik,bk,bjk->bij", V, w, V)
    return out


def project_scalar_codegen(a, method: int):
    """Size-1 spd node inside generated element code (``codeGenerator.pyx:820-822``):
    method 0 -> a, 1 -> |a|, 2 -> max(0, a)."""
    if method == 0:
        return a
    return abs(a) if method == 1 else max(0.0, a)


def jacobi_block_inverse(blocks):
    """Pseudo-inverse of each diagonal block [B,n,n] with the reference eigenvalue rule."""
    blocks = np.asarray(blocks, dtype=np.float64)
    w, V = np.linalg.eigh(blocks)
    inv = np.where(np.abs(w) < 1e-6, np.abs(w), np.abs(1.0 / np.where(w == 0, 1.0, w)))
    return np.einsum("bik,bk,bjk->bij", V, inv, V)
