"""Linear solvers of H x = g (``solver/solver.py``, ``solver/jacobianPCGSolver.pyx``).

``solver("jacobian")`` is the block-Jacobi PCG of ``kernel/Solver/solverKernel.pyx``:
diagonal blocks pseudo-inverted with the |l| < 1e-6 rule
(``diagonalBlockInverseKernel.pyx:69,91``), device-resident recurrence, the reference
stopping rule r^T M^-1 r <= tol * r0^T M^-1 r0 (tolerance is a squared, preconditioned
ratio) and the reference return codes mapped as in ``solverKernel.pyx:800-824``
(iterations -> 0 on success).

Extensions: ``threshold_mode="relative_residual"`` stops on ||r|| <= tol ||b|| instead;
``precision="mixed"`` runs the Krylov loop on FP32 block values and FP32 inverse blocks
inside an FP64 iterative refinement (``refine.py``) and stops on the true FP64 relative
residual.
"""
from __future__ import annotations

import ctypes as C
import time

from . import _lib, device as dev
from .hessian import vector
from .refine import RefinementBuffers, refine_solve

_MODES = {"reference": 0, "relative_residual": 1}
PRECISIONS = ("fp64", "mixed")


class jacobianPCGSolver:
    def __init__(self, threshold_mode: str = "reference", precision: str = "fp64"):
        if threshold_mode not in _MODES:
            raise ValueError(f"threshold_mode must be one of {sorted(_MODES)}")
        if precision not in PRECISIONS:
            raise ValueError(f"precision must be one of {PRECISIONS}")
        self.threshold_mode = threshold_mode
        self.precision = precision
        self._refine = RefinementBuffers()
        self._minv32 = None
        self.__solution = None
        self.__work = None
        self.__scalars = None
        self.__statistics = {}
        self.iterations = 0

    @property
    def solution(self):
        return self.__solution

    @property
    def statistics(self) -> dict:
        return dict(self.__statistics)

    def reset(self):
        self.__solution = None
        self.__work = None
        self.__scalars = None
        self.__statistics = {}

    def _ensure(self, n):
        if self.__solution is None or self.__solution.numel() != n:
            self.__solution = dev.zeros(n, "float64")
            self.__work = dev.zeros(5 * n, "float64")
            self.__scalars = dev.zeros(12288, "float64")   # YSP_PCG3_SCALARS

    def computeSolution(self, active_hessian, right_hand_side, initial_guess=None, tolerance=1e-3,
                        maxIterations=20000, zero_initial_guess=False):
        t0 = time.perf_counter()
        if not isinstance(right_hand_side, vector):
            raise TypeError("jacobianPCGSolver.computeSolution: right_hand_side must be a yasps_amd vector")
        n = right_hand_side.size
        if n == 0:
            self.__statistics = {"solver": "jacobian", "converged": True, "iterations": 0,
                                 "solve_seconds": time.perf_counter() - t0}
            return 0
        self._ensure(n)
        if active_hessian is None:
            self.__solution.zero_()
            self.__statistics = {"solver": "jacobian", "converged": True, "iterations": 0,
                                 "solve_seconds": time.perf_counter() - t0}
            return 0
        if n != active_hessian.rows:
            raise ValueError("jacobianPCGSolver.computeSolution: right_hand_side size must match the Hessian size.")
        lib = _lib.load()
        h = active_hessian
        s = dev.stream()
        counts = [w.numInstances for w in h.wrt]
        sizes = [w.size for w in h.wrt]
        starts = h.diagonal_blocks_start_cpu[:-1]
        t1 = time.perf_counter()
        _lib.check(lib.ysp_invert_diagonal_blocks(
            dev.ptr(h.diagonal_blocks), dev.ptr(h.diagonal_blocks_inverse), _lib.u32_array(starts),
            _lib.u32_array(counts), _lib.u32_array(sizes), len(h.wrt), s), "ysp_invert_diagonal_blocks")
        t2 = time.perf_counter()
        stats = _lib.PcgStats()
        guess = None
        if not zero_initial_guess and initial_guess is not None:
            guess = initial_guess.value if hasattr(initial_guess, "value") else initial_guess
        zero = int(bool(zero_initial_guess) or guess is None)
        if self.precision == "mixed":
            return self._solve_mixed(h, right_hand_side, n, tolerance, maxIterations, t0, t1, t2)
        if h.full is not None:
            f = h.full
            vals, fmt = h.solve_values()
            rc = lib.ysp_pcg3_solve(
                int(maxIterations), float(tolerance), _MODES[self.threshold_mode],
                dev.ptr(f.row_ptr), dev.ptr(f.col), vals, fmt, f.nnz,
                dev.ptr(h.diagonal_blocks_inverse), 0, dev.ptr(right_hand_side.value), n,
                dev.ptr(self.__solution), dev.ptr(guess), zero, dev.ptr(self.__work),
                dev.ptr(self.__scalars), C.byref(stats), s)
        elif h.split is not None:
            f = h.split
            vals, fmt = h.solve_values()
            rc = lib.ysp_pcg_solve_full(
                int(maxIterations), float(tolerance), _MODES[self.threshold_mode],
                dev.ptr(f.row_ptr), dev.ptr(f.col), vals, fmt, f.nnz, 0,
                dev.ptr(h.diagonal_blocks_inverse), _lib.u32_array(starts), _lib.u32_array(counts),
                _lib.u32_array(sizes), _lib.u32_array(h.gradient_segments_start_cpu[:-1]),
                len(h.wrt), dev.ptr(right_hand_side.value), n, dev.ptr(self.__solution),
                dev.ptr(guess), zero, dev.ptr(self.__work), dev.ptr(self.__scalars),
                C.byref(stats), s)
        else:
            a, b = h.static.ctype(), h.dynamic.ctype()
            rc = lib.ysp_pcg_solve(
                int(maxIterations), float(tolerance), _MODES[self.threshold_mode],
                C.byref(a) if a else None, C.byref(b) if b else None,
                dev.ptr(h.diagonal_blocks_inverse), _lib.u32_array(starts), _lib.u32_array(counts),
                _lib.u32_array(sizes), _lib.u32_array(h.gradient_segments_start_cpu[:-1]),
                len(h.wrt), dev.ptr(right_hand_side.value), n, dev.ptr(self.__solution),
                dev.ptr(guess), zero, dev.ptr(self.__work), dev.ptr(self.__scalars),
                C.byref(stats), s)
        _lib.check(rc, "ysp_pcg_solve")
        t3 = time.perf_counter()
        self.iterations = max(0, stats.iterations)
        self.__statistics = {
            "solver": "jacobian", "converged": rc >= 0, "iterations": int(stats.iterations),
            "result": int(rc), "delta0": stats.delta0, "delta": stats.delta,
            "rr0": stats.rr0, "rr": stats.rr,
            "diagonal_inverse_seconds": t2 - t1, "pcg_seconds": t3 - t2,
            "solve_seconds": t3 - t0, "matrix_size": n, "tolerance": float(tolerance),
            "dynamic_block_count": int(sum(h.block_counts_dynamic)),
        }
        if rc == -5:
            self.__solution.copy_(0.5 * right_hand_side.value)
            return rc
        if rc < 0:
            return rc
        return 0

    def _solve_mixed(self, h, rhs, n, tolerance, max_iterations, t0, t1, t2):
        if h.operator is None:
            raise NotImplementedError("precision='mixed' needs 3 x 3 blocks or affine-body "
                                      "segments of multiples of 3")
        lib = _lib.load()
        s = dev.stream()
        f = h.operator
        v32 = h.full_values_f32()
        stats = _lib.PcgStats()
        if h.full is None:
            # affine-body layout: the split operator in FP32, the reference's block-Jacobi
            # (9 x 9 inverse blocks for the affine matrices) kept in FP64
            counts = [w.numInstances for w in h.wrt]
            sizes = [w.size for w in h.wrt]
            starts = h.diagonal_blocks_start_cpu[:-1]

            def inner(r, d, tol_in, max_it):
                rc = lib.ysp_pcg_solve_full(
                    int(max_it), float(tol_in), 1, dev.ptr(f.row_ptr), dev.ptr(f.col), dev.ptr(v32), 1,
                    f.nnz, 0, dev.ptr(h.diagonal_blocks_inverse), _lib.u32_array(starts),
                    _lib.u32_array(counts), _lib.u32_array(sizes),
                    _lib.u32_array(h.gradient_segments_start_cpu[:-1]), len(h.wrt), dev.ptr(r), n,
                    dev.ptr(d), None, 1, dev.ptr(self.__work), dev.ptr(self.__scalars),
                    C.byref(stats), s)
                _lib.check(rc, "ysp_pcg_solve_full")
                return rc, int(stats.iterations)
        else:
            nd = h.diagonal_blocks_inverse.numel()
            if self._minv32 is None or self._minv32.numel() != nd:
                self._minv32 = dev.empty(nd, "float32")
            _lib.check(lib.ysp_cast_f32(dev.ptr(h.diagonal_blocks_inverse), dev.ptr(self._minv32),
                                        nd, s), "ysp_cast_f32")

            def inner(r, d, tol_in, max_it):
                rc = lib.ysp_pcg3_solve(int(max_it), float(tol_in), 1, dev.ptr(f.row_ptr),
                                        dev.ptr(f.col), dev.ptr(v32), 1, f.nnz, dev.ptr(self._minv32),
                                        1, dev.ptr(r), n, dev.ptr(d), None, 1, dev.ptr(self.__work),
                                        dev.ptr(self.__scalars), C.byref(stats), s)
                _lib.check(rc, "ysp_pcg3_solve")
                return rc, int(stats.iterations)

        out = refine_solve(h, rhs.value, self.__solution, float(tolerance), int(max_iterations), inner,
                           self._refine)
        t3 = time.perf_counter()
        self.iterations = out["iterations"]
        self.__statistics = {
            "solver": "jacobian", "precision": "mixed", "converged": out["converged"],
            "iterations": out["iterations"], "outer_iterations": out["outer_iterations"],
            "relative_residual": out["relative_residual"], "outer_history": out["outer_history"],
            "result": 0 if out["converged"] else -4, "diagonal_inverse_seconds": t2 - t1,
            "pcg_seconds": t3 - t2, "solve_seconds": t3 - t0, "matrix_size": n,
            "tolerance": float(tolerance),
        }
        return 0 if out["converged"] else -4


class solver:
    """Solver dispatch (``solver/solver.py:9-68``); MAS by default, as the reference
    (``solver/solver.py:22``)."""

    def __init__(self, solver="mas", **options):
        self.__name = ""
        self.__impl = None
        self.setSolver(solver, **options)

    @property
    def solverName(self) -> str:
        return self.__name

    @property
    def implementation(self):
        return self.__impl

    @property
    def solution(self):
        return self.__impl.solution

    @property
    def statistics(self) -> dict:
        return self.__impl.statistics

    def setSolver(self, solver="mas", **options):
        if not isinstance(solver, str):
            raise TypeError("solver must be either 'jacobian' or 'mas'")
        name = solver.strip().lower()
        if name == "jacobian":
            unexpected = set(options) - {"threshold_mode", "precision"}
            if unexpected:
                raise TypeError(f"jacobianPCGSolver does not accept options: {', '.join(sorted(unexpected))}")
            impl = jacobianPCGSolver(**options)
        elif name == "mas":
            from .mas import masSolver
            impl = masSolver(**options)
        else:
            raise ValueError("solver must be either 'jacobian' or 'mas'")
        if self.__impl is not None:
            self.__impl.reset()
        self.__name = name
        self.__impl = impl
        return self

    def reset(self):
        self.__impl.reset()

    def computeSolution(self, *args, **kwargs):
        return self.__impl.computeSolution(*args, **kwargs)
