"""Block-Jacobi preconditioned CG with the reference recurrences (FP64 numpy).

Oracle module (test infrastructure only).

``kernel/Solver/solverKernel.pyx:477-687`` (zero-initial-guess path ``:534-549``):
  r = b, c = M^-1 b, delta0 = delta_new = c.b; stop at once if delta_new <= tol*delta0.
  Each iteration: q = A c; den = c.q; alpha = delta_new / den (status -(it+4) when den
  is not finite or <= 0, ``:380-392``); x += alpha c; r -= alpha q; s = M^-1 r;
  delta_new = r.s; beta = delta_new / delta_old; converged when delta_new <= tol*delta0
  (``:411-431``); c = s + beta c. Returns the iteration count (>= 0), -4 when the
  iteration budget is exhausted and -5 for an invalid initial residual (``:600-611``).
The reported true relative residual ||b - A x|| / ||b|| is computed separately.
"""
from __future__ import annotations

import numpy as np


def block_jacobi_apply(inv_blocks, x, block=3):
    nb = inv_blocks.shape[0]
    return np.einsum("bij,bj->bi", inv_blocks, x.reshape(nb, block)).reshape(-1)


def pcg(apply_A, apply_M, b, tol, max_iterations=20000):
    """Returns (status, x, iterations, history of delta_new)."""
    b = np.asarray(b, dtype=np.float64)
    x = np.zeros_like(b)
    r = b.copy()
    c = apply_M(b)
    delta0 = float(c @ b)
    delta_new = delta0
    hist = [delta_new]
    if not np.isfinite(delta0) or delta0 < 0:
        return -5, 0.5 * b, 0, hist
    thresh = tol * delta0
    if delta_new <= thresh:
        return 0, x, 0, hist
    for it in range(1, max_iterations + 1):
        q = apply_A(c)
        den = float(c @ q)
        if not np.isfinite(den) or den <= 0.0:
            return -it - 4, x, it, hist
        alpha = delta_new / den
        x = x + c * alpha
        r = r - q * alpha
        s = apply_M(r)
        delta_old = delta_new
        delta_new = float(r @ s)
        hist.append(delta_new)
        if not np.isfinite(delta_new) or delta_new < 0 or delta_old <= 0:
            return -it - 4, x, it, hist
        beta = delta_new / delta_old
        c = s + c * beta
        if delta_new <= thresh:
            return it, x, it, hist
    return -4, x, max_iterations, hist


def relative_residual(apply_A, x, b):
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(b - apply_A(x)) / (nb if nb > 0 else 1.0))
