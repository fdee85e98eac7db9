"""Mixed-precision solve: FP32 block values in the Krylov loop, FP64 iterative refinement.

No reference counterpart (the reference's MAS has an optional FP32 SpMV with an FP64
audit-and-retry, ``solver/mas/solver.py:643-688``). Here every outer step solves
A32 d = r with the inner PCG (FP32 block values and preconditioner data, FP64 vectors,
dots and scalars), then ``ysp_bsr3_refine`` forms x += d and r = b - A64 x with the FP64
values; the loop stops on the true FP64 relative residual ||b - A64 x|| <= tol ||b||.
"""
from __future__ import annotations

import ctypes as C
import math

from . import _lib, device as dev


class RefinementBuffers:
    def __init__(self):
        self.r = None
        self.d = None

    def ensure(self, n):
        if self.r is None or self.r.numel() != n:
            self.r = dev.empty(n, "float64")
            self.d = dev.empty(n, "float64")


def refine_solve(h, rhs, solution, tol, max_iterations, inner, buffers: RefinementBuffers,
                 max_outer: int = 8, safety: float = 0.7) -> dict:
    """``inner(r, d, tol_in, max_it) -> (rc, iterations)`` solves A32 d = r to the relative
    residual tol_in. Returns the outer statistics; ``solution`` holds x."""
    lib = _lib.load()
    f = h.operator
    n = int(rhs.numel())
    buffers.ensure(n)
    r, d = buffers.r, buffers.d
    s = dev.stream()
    rr, bb = C.c_double(0.0), C.c_double(0.0)

    def step(correction):
        _lib.check(lib.ysp_bsr3_refine(dev.ptr(f.row_ptr), dev.ptr(f.col), dev.ptr(f.v64), 0, f.nnz, n,
                                       dev.ptr(rhs), dev.ptr(solution), dev.ptr(correction),
                                       dev.ptr(r), C.byref(rr), C.byref(bb), s), "ysp_bsr3_refine")

    solution.zero_()
    step(None)                                  # r = b
    bnorm = math.sqrt(bb.value)
    total, outer, rc = 0, 0, 0
    history = []
    while True:
        rel = math.sqrt(rr.value) / bnorm if bnorm > 0 else 0.0
        history.append(rel)
        if not math.isfinite(rel):
            rc = -5
            break
        if rel <= tol or outer >= max_outer or total >= max_iterations:
            break
        tol_in = min(0.5, max(1e-14, safety * tol / rel))
        rc, its = inner(r, d, tol_in, max_iterations - total)
        total += max(0, its)
        outer += 1
        if rc < 0 and rc != -4:                 # breakdown of the inner solve
            break
        step(d)
    converged = history[-1] <= tol
    return {"converged": converged, "outer_iterations": outer, "iterations": total,
            "relative_residual": history[-1], "outer_history": history, "inner_result": rc}
