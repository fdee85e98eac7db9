"""Differentiable material inverse: static equilibrium of a tet cube under gravity with a
per-tet parameter theta_e = log E_e, and its adjoint.

Equilibrium: minimise Pi(x) = sum_e vol_e psi_SNH(F_e; mu(theta_e), lam(theta_e))
+ sum_v m_v g y_v over the free vertices (bottom face fixed) by Newton with a
backtracking line search, each step solving H d = grad Pi with the PCG of ``solver``
(MAS by default, as the reference's minimizer; or block-Jacobi).

Adjoint: for a loss L(x*(theta)), H(x*) lambda = dL/dx (the same PCG, FP64 or the mixed
FP32/FP64 refinement) and dL/dtheta_e = -lambda_e . dg_e/dtheta_e, evaluated per tet by
``ysp_snh_theta_adjoint``. ``EquilibriumFunction`` exposes theta -> x* as a
``torch.autograd.Function`` whose backward is this adjoint.

The tets carry Young's modulus and Poisson's ratio as per-tet constants and the SNH
coefficients are expressions of them (``lame_parameters`` applied to the attributes), so
``vjp_material`` gives dL/dE and dL/dnu per tet through the generalised mixed
Jacobian d g / d (E, nu) of ``yasps_amd.differentiator`` (the material example's
``young_jacobian`` / ``poisson_jacobian`` products,
``examples/bunny_material_optimization/bunny_material_optimization.py:616-617,1123-1126``).
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib, device as dev
from .hessian import vector
from .scenes import cube_scene, lame_parameters


class MaterialProblem:
    def __init__(self, n: int = 8, young: float = 1e5, poisson: float = 0.4,
                 density: float = 1000.0, gravity: float = 9.8, precision: str = "fp64",
                 projection_method: int = 2, solver: str = "mas", **solver_options):
        from . import scene, stable_neo_hookean
        cs = cube_scene(n, young=young, poisson=poisson, density=density, dt=1.0, noise=0.0)
        self.cs = cs
        self.poisson = poisson
        self.young0 = young
        nf = cs.n_fixed
        self.n_free = cs.rest.shape[0] - nf
        s0 = scene("material")
        dt = s0.addConstant("dt", rows=1, cols=1)
        dt.updateValue([1.0])
        m = s0.addMesh("cube")
        fv = m.addPrimitive("fixed_vertices", numInstances=nf)
        mv = m.addPrimitive("moving_vertices", numInstances=self.n_free)
        m.addPrimitiveUnion("vertices", [fv, mv])
        tets = m.addPrimitive("tets", numInstances=cs.tets.shape[0])
        fv.addAttribute("position", rows=3, cols=1).updateValue(cs.rest[:nf])
        fv.addConstant("rest_position", rows=3, cols=1).updateValue(cs.rest[:nf])
        mv.addConstant("rest_position", rows=3, cols=1).updateValue(cs.rest[nf:])
        self.x_att = mv.addAttribute("position", rows=3, cols=1)
        self.x_att.updateValue(cs.rest[nf:])
        m.vertices.addAttribute("position")
        m.vertices.addAttribute("rest_position")
        self.young_att = tets.addConstant("young", rows=1, cols=1)
        self.poisson_att = tets.addConstant("poisson", rows=1, cols=1)
        self.mu_expr, self.lam_expr = lame_parameters(self.young_att, self.poisson_att)
        tet2v = tets.addConnectivity("tet2v", m.vertices, cs.tets, 4)
        tp = tets.addAttribute("positions", through=tet2v, source=m.vertices["position"])
        trp = tets.addAttribute("rest_positions", through=tet2v, source=m.vertices["rest_position"])
        self.snh = stable_neo_hookean(trp, tp, self.mu_expr, self.lam_expr, dt)
        snh_att = tets.addAttribute("stable_neo_hookean", computed_attribute=self.snh)
        s0.addEnergy(snh_att, projection_method=projection_method)
        s0.addMinimizeTarget([self.x_att])
        s0.minimizer.setSolver(solver, threshold_mode="relative_residual", precision=precision,
                               **solver_options)
        self.scene = s0
        self.snh_att = snh_att
        # gravity potential sum_v m_v g y_v: constant gradient m_v g e_y
        f = np.zeros((self.n_free, 3))
        f[:, 1] = cs.mass[nf:] * gravity
        self.f_host = f.reshape(-1)
        self.f = dev.to_device(self.f_host, "float64")
        top = np.flatnonzero(np.isclose(cs.rest[nf:, 2], 1.0))
        self.top = top
        self.theta = None

    # ------------------------------------------------------------ helpers
    @property
    def num_tets(self) -> int:
        return self.cs.tets.shape[0]

    def set_material(self, young, poisson):
        """Per-tet Young's modulus and Poisson's ratio (arrays or scalars)."""
        T = self.num_tets
        self.young = np.broadcast_to(np.asarray(young, dtype=np.float64), (T,)).copy()
        self.nu = np.broadcast_to(np.asarray(poisson, dtype=np.float64), (T,)).copy()
        self.young_att.updateValue(self.young)
        self.poisson_att.updateValue(self.nu)

    def set_theta(self, theta: np.ndarray):
        theta = np.asarray(theta, dtype=np.float64)
        self.set_material(np.exp(theta), self.poisson)
        self.theta = theta.copy()

    def _set_x(self, x):
        self.x_att.updateValue(x)

    def potential(self) -> float:
        e = dev.to_host(self.snh_att.compute().value).sum()
        x = dev.to_host(self.x_att.value)
        return float(e + self.f_host @ x)

    def _residual(self):
        hs = self.scene.minimizer.computeNumericValue()
        r = hs.gradient.value + self.f
        return hs, r

    def _solve(self, hs, rhs, tol):
        slv = self.scene.minimizer.linearSolver
        rc = slv.computeSolution(hs, vector(hs.rows, rhs), None, tolerance=tol, zero_initial_guess=True)
        if rc != 0:
            raise RuntimeError(f"material problem: PCG failed ({rc})")
        return slv.solution.clone()

    # ------------------------------------------------------------ forward
    def equilibrium(self, theta, x0=None, tol: float = 1e-12, max_newton: int = 60,
                    pcg_tol: float = 1e-11):
        """Static equilibrium x* (free vertex positions, flat) for theta."""
        self.set_theta(theta)
        return self._newton(x0, tol, max_newton, pcg_tol)

    def equilibrium_material(self, young, poisson, x0=None, tol: float = 1e-12,
                             max_newton: int = 60, pcg_tol: float = 1e-11):
        """Static equilibrium x* for per-tet Young's modulus and Poisson's ratio."""
        self.set_material(young, poisson)
        self.theta = None
        return self._newton(x0, tol, max_newton, pcg_tol)

    def _newton(self, x0, tol, max_newton, pcg_tol):
        x = (self.cs.rest[self.cs.n_fixed:].reshape(-1) if x0 is None else np.asarray(x0)).copy()
        self._set_x(x)
        fnorm = float(np.linalg.norm(self.f_host))
        self.newton_iterations = 0
        history = []
        hs = self.scene.minimizer._active()
        saved = list(hs.terms)
        try:
            for it in range(max_newton):
                # projected Hessians far from equilibrium, the exact one near it (quadratic
                # convergence; the stable equilibrium's Hessian is positive definite)
                rel = history[-1] / fnorm if history else 1.0
                method = None if rel > 1e-4 else 0
                hs.terms = saved if method is None else [(e, method) for e, _ in saved]
                hs, r = self._residual()
                rn = float(dev.torch().linalg.norm(r))
                history.append(rn)
                if rn <= tol * fnorm:
                    break
                # at the rounding floor the residual stops falling: accept it there
                if len(history) >= 3 and rn <= 1e-11 * fnorm and rn >= 0.5 * history[-3]:
                    break
                d = dev.to_host(self._solve(hs, r, pcg_tol))
                alpha = 1.0
                if rn > 1e-6 * fnorm:          # backtracking on the potential
                    pi0 = self.potential()
                    slope = -float(dev.to_host(r) @ d)
                    while True:
                        self._set_x(x - alpha * d)
                        if self.potential() <= pi0 + 1e-4 * alpha * slope or alpha < 1e-8:
                            break
                        alpha *= 0.5
                x = x - alpha * d
                self._set_x(x)
                self.newton_iterations = it + 1
            else:
                raise RuntimeError("material problem: Newton did not converge")
        finally:
            hs.terms = saved
        self.residual_history = history
        return x

    def loss(self, x, target):
        xt = x.reshape(-1, 3)[self.top]
        return float(((xt - target) ** 2).sum())

    def loss_grad_x(self, x, target):
        g = np.zeros_like(x.reshape(-1, 3))
        g[self.top] = 2.0 * (x.reshape(-1, 3)[self.top] - target)
        return g.reshape(-1)

    # ------------------------------------------------------------ adjoint
    def adjoint(self, x, xbar, pcg_tol: float = 1e-12, projection_method: int = 0):
        """lambda with H(x) lambda = xbar at the equilibrium x (H unprojected by default:
        the adjoint of the exact equilibrium); returns (lambda, the assembled Hessian)."""
        self._set_x(x)
        mn = self.scene.minimizer
        hs = mn._active()
        saved = list(hs.terms)
        hs.terms = [(e, projection_method) for e, _ in saved]
        try:
            hs = mn.computeNumericValue()
            lam = self._solve(hs, dev.to_device(np.asarray(xbar, dtype=np.float64), "float64"), pcg_tol)
        finally:
            hs.terms = saved
        return lam, hs

    def vjp_material(self, x, xbar, pcg_tol: float = 1e-12, projection_method: int = 0):
        """(dL/dE, dL/dnu) per tet for dL/dx = xbar at the equilibrium x:
        -J^T lambda with J = d g / d (E, nu) from ``differentiator().diff2`` (the chain
        through the SNH coefficients' expressions)."""
        from .differentiator import differentiator
        lam, _ = self.adjoint(x, xbar, pcg_tol, projection_method)
        if getattr(self, "_material_jacobian", None) is None:
            self._material_jacobian = differentiator().diff2(
                [self.snh_att], [self.x_att], [self.young_att, self.poisson_att])
        J = self._material_jacobian.compute()
        g = -dev.to_host(J.spmv(lam, transpose=True).value)
        T = self.num_tets
        return g[:T], g[T:]

    def vjp(self, x, xbar, pcg_tol: float = 1e-12, projection_method: int = 0):
        """dL/dtheta for dL/dx = xbar at the equilibrium x (H unprojected by default: the
        adjoint of the exact equilibrium)."""
        from .expr import values as expr_values
        lam, hs = self.adjoint(x, xbar, pcg_tol, projection_method)
        out = dev.empty(self.num_tets, "float64")
        ti = hs.static_index[0]
        sa = hs._scatter_args(ti, hs.static)
        from .energies import union_view
        e = self.snh
        dm_inv, vol = e.rest_data()
        u = union_view(e.pos.source)
        T = self.num_tets
        mu = expr_values(self.mu_expr, T, 1).reshape(-1)
        la = expr_values(self.lam_expr, T, 1).reshape(-1)
        _lib.check(_lib.load().ysp_snh_theta_adjoint(
            C.byref(u), dev.ptr(e.conn.value), T, dev.ptr(dm_inv), dev.ptr(vol),
            dev.ptr(mu), 1, dev.ptr(la), 1, 1.0, C.byref(sa), dev.ptr(lam), dev.ptr(out),
            dev.stream()), "ysp_snh_theta_adjoint")
        _lib.check(_lib.load().ysp_check_deferred(dev.stream()), "ysp_snh_theta_adjoint")
        return dev.to_host(out)


def equilibrium_function():
    """``torch.autograd.Function`` theta -> x* with the adjoint backward (built lazily so
    that importing this module does not import torch)."""
    import torch

    class EquilibriumFunction(torch.autograd.Function):
        @staticmethod
        def forward(ctx, theta, problem):
            x = problem.equilibrium(theta.detach().cpu().numpy())
            ctx.problem = problem
            ctx.x = x
            return torch.as_tensor(x, dtype=theta.dtype, device=theta.device)

        @staticmethod
        def backward(ctx, xbar):
            g = ctx.problem.vjp(ctx.x, xbar.detach().cpu().numpy())
            return torch.as_tensor(g, dtype=xbar.dtype, device=xbar.device), None

    return EquilibriumFunction
