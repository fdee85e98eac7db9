"""Energy collection and Newton linear solve (``minimizer/minimizer.pyx``).

Energies are kept per request; once the targets are known the static and dynamic terms
are summed into one Hessian (``computeNumericValue`` ``:323-339``), which is assembled and
solved from a zero initial guess (``__solveLinearSystem`` ``:363-378``). The solution is
H^-1 g (the caller steps x <- x - alpha * solution, ``examples/one_bunny/one_bunny.py:218``).
"""
from __future__ import annotations

from typing import List, Optional

from . import device as dev
from .hessian import check_deferred, hessian
from .solver import solver


class minimizer:
    def __init__(self):
        self.__requests = []             # (energy attribute, targets, method, dynamic)
        self.__wrt = []
        self.__hessian: Optional[hessian] = None
        self.__solver = solver("mas")            # the reference default (minimizer.pyx:101)
        self.__segments: List = []
        self.__ignored = set()

    # ------------------------------------------------------------- configuration
    def addEnergy(self, energy, targets, projection_method, dynamic):
        self.__requests.append((energy, list(targets), int(projection_method), bool(dynamic)))
        self.__hessian = None

    def addWrt(self, wrt):
        for w in wrt:
            if w.isDynamic:
                raise ValueError("minimizer.addWrt: wrt can not contain dynamic attributes.")
            if w.operator != "DATA":
                raise ValueError(f"minimizer.addWrt: {w.fullName} is not a DATA attribute")
        self.__wrt = list(wrt)
        self.__hessian = None

    def ignoreEnergies(self, energies):
        self.__ignored = {e for e in energies}
        self.__hessian = None

    @property
    def wrt(self):
        return list(self.__wrt)

    @property
    def linearSolver(self):
        return self.__solver

    @linearSolver.setter
    def linearSolver(self, value):
        for name in ("computeSolution", "reset", "solution"):
            if not hasattr(type(value), name):
                raise TypeError(f"minimizer.linearSolver requires a solver-compatible object; missing {name}")
        self.__solver.reset()
        self.__solver = value
        self.__segments = []

    @property
    def solverName(self):
        return getattr(self.__solver, "solverName", type(self.__solver).__name__)

    def setSolver(self, solver_name="mas", **options):
        self.linearSolver = solver(solver_name, **options)
        return self.__solver

    # ------------------------------------------------------------- numeric
    def _active(self) -> hessian:
        if not self.__wrt:
            raise ValueError("minimizer: wrt is not initialized. Please call addMinimizeTarget first.")
        if self.__hessian is None:
            static, dynamic = [], []
            for energy, _targets, method, dyn in self.__requests:
                if energy in self.__ignored:
                    continue
                (dynamic if dyn else static).append((energy.expression, method))
            self.__hessian = hessian(self.__wrt, static, dynamic)
        return self.__hessian

    @property
    def activeHessian(self) -> Optional[hessian]:
        return self.__hessian

    def computeNumericValue(self, check: bool = True) -> hessian:
        h = self._active()
        h.compute(check=check)
        return h

    @property
    def gradient(self):
        if self.__hessian is None or not self.__hessian._setup:
            return dev.zeros(0, "float64")
        return self.__hessian.gradient.value

    @property
    def diagonal(self):
        """``minimizer.pyx:162-165``: the active Hessian's scalar diagonal."""
        if self.__hessian is None or not self.__hessian._setup:
            return dev.zeros(0, "float64")
        return self.__hessian.diagonal

    @property
    def gradientSegments(self):
        g = self.gradient
        out = []
        if self.__hessian is None:
            return out
        st = self.__hessian.gradient_segments_start_cpu
        for i in range(len(st) - 1):
            out.append(g[st[i]:st[i + 1]])
        return out

    @property
    def solutionSegments(self):
        return self.__segments

    def computeHessianAndGradient(self, tolerance=1e-3, maxIterations=20000):
        h = self.computeNumericValue(check=False)
        rc = self.__solver.computeSolution(h, h.gradient, None, tolerance=tolerance,
                                           maxIterations=maxIterations, zero_initial_guess=True)
        check_deferred()                  # element flags of the assembly, after the solve
        sol = self.__solver.solution
        st = h.gradient_segments_start_cpu
        self.__segments = [sol[st[i]:st[i + 1]] for i in range(len(st) - 1)]
        return rc

    def computeSolution(self, tolerance=1e-3, maxIterations=20000):
        rc = self.computeHessianAndGradient(tolerance=tolerance, maxIterations=maxIterations)
        if rc < 0:
            print("Warning: solver did not converge. Returning the best solution found.")
        return self.__segments

    def computeTotalEnergy(self) -> float:
        total = 0.0
        for energy, _t, _m, dyn in self.__requests:
            if energy in self.__ignored or energy.numInstances == 0:
                continue
            total += float(energy.compute().value.sum().item())
        return total
