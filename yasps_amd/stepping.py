"""Implicit time stepping over the solve path, as the reference's example driver runs it
(examples/one_bunny/one_bunny.py:173-295): last_position <- position; Newton iterations,
each minimizeEnergy (H d = g on the current positions) followed by a backtracking line
search on the total energy (computeTotalEnergy): x - step d with step halved up to
``max_halvings`` times until the energy does not increase. The step is always taken; the
loop then ends when the largest gradient entry of the iterate the direction was computed
at fell below ``gradient_tol`` (``one_bunny.py:176-283``: max_grad is read before the
line search and tested after it). Then velocity = (x - last_position) / dt.

The first trial step is ``first_step`` (1.0 by default) or, when ``step_bound`` is given,
``step_bound(x, d)``: the example's continuous-collision bound
(``ccd.compute_largest_step_size(0.8, ...)``, ``one_bunny.py:205-206``). ``on_positions(x)``
runs whenever the target positions change (the start of the step and every trial of the
line search), before any energy is evaluated there: the example's collision detection and
update of the contact primitives (``one_bunny.py:218-234``). ``ContactSets`` provides both
callbacks over ``yasps_amd.ccd.CCD``."""
from __future__ import annotations

from . import device as dev

_WIDTHS = {"pp": 2, "pe": 3, "pt": 4, "ee": 4}


class ContactSets:
    """The example's contact loop (``examples/one_bunny/one_bunny.py:196-234``) for a vertex
    union [fixed, moving]: ``step_bound`` sweeps the Newton direction (the fixed vertices do
    not move) and returns ``compute_largest_step_size(slackness, ...)``; ``refresh`` runs
    the proximity detection at the trial positions and loads the pp / pe / pt / ee pairs
    into the dynamic primitives (``updateNumInstances`` + ``updateConnectivity``).
    ``primitives`` maps "pp", "pe", "pt", "ee" to (primitive, connectivity)."""

    def __init__(self, ccd, fixed_positions, primitives: dict, dhat: float,
                 slackness: float = 0.8):
        torch = dev.torch()
        self.ccd = ccd
        fp = fixed_positions if hasattr(fixed_positions, "is_cuda") else \
            dev.to_device(fixed_positions, "float64")
        self.fixed = fp.reshape(-1)
        self.primitives = primitives
        self.dhat = float(dhat)
        self.slackness = float(slackness)
        self.counts = [(0, 0, 0, 0)]
        self._zeros = torch.zeros_like(self.fixed)

    def positions(self, moving):
        return dev.torch().cat([self.fixed, moving.reshape(-1)])

    def step_bound(self, x_moving, direction) -> float:
        torch = dev.torch()
        x = self.positions(x_moving)
        d = torch.cat([self._zeros, direction.reshape(-1)])
        self.ccd.ccd(x, self.dhat, d, 1.0)
        return self.ccd.compute_largest_step_size(self.slackness, x, d)

    def refresh(self, x_moving) -> None:
        counts = self.ccd.cd(self.positions(x_moving), self.dhat)
        self.counts.append(tuple(counts))
        for (name, w), n in zip(_WIDTHS.items(), counts):
            prim, conn = self.primitives[name]
            prim.updateNumInstances(n)
            if n:
                conn.updateConnectivity(getattr(self.ccd, name)[:w * n])


def newton_time_step(s0, moving, dt: float, tolerance: float = 1e-6, gradient_tol: float = 1e-1,
                     max_newton: int = 50, max_halvings: int = 8, first_step: float = 1.0,
                     max_iterations: int = 20000, step_bound=None, on_positions=None) -> dict:
    """One time step on scene ``s0`` whose minimise target is ``moving["position"]`` (a
    primitive that also holds ``last_position`` and ``velocity``). Returns the Newton
    history (energies, largest gradient entries, steps taken, PCG iterations)."""
    torch = dev.torch()
    pos = moving["position"]
    last = pos.value.clone()
    moving["last_position"].updateValue(last)
    if on_positions is not None:
        on_positions(pos.value)
    energies, grads, steps, pcg = [], [], [], []
    converged = False
    for _ in range(max_newton):
        (d,) = s0.minimizeEnergy(tolerance=tolerance, maxIterations=max_iterations)
        g = s0.gradient
        gmax = float(torch.max(torch.abs(g.value if hasattr(g, "value") else g)).item())
        e0 = s0.minimizer.computeTotalEnergy()
        energies.append(e0)
        grads.append(gmax)
        pcg.append(int(s0.minimizer.linearSolver.statistics.get("iterations", 0)))
        x0 = pos.value.clone()
        direction = d.reshape(x0.shape)
        step = first_step if step_bound is None else float(step_bound(x0, direction))
        trial = None
        for _h in range(max_halvings):
            pos.updateValue(x0 - step * direction)
            if on_positions is not None:
                on_positions(pos.value)
            trial = s0.minimizer.computeTotalEnergy()
            if trial <= e0:
                break
            step *= 0.5
        else:
            pos.updateValue(x0)
            if on_positions is not None:
                on_positions(pos.value)
            raise RuntimeError(
                f"newton_time_step: the line search found no decrease (Newton iteration "
                f"{len(energies)}, max|g| {gmax:.3e}, energy {e0!r}, last trial {trial!r} at "
                f"step {2 * step:.3e})")
        steps.append(step)
        if gmax < gradient_tol:
            converged = True
            break
    moving["velocity"].updateValue((pos.value - last) / dt)
    return {"converged": converged, "energies": energies, "max_gradient": grads,
            "steps": steps, "pcg_iterations": pcg, "newton_iterations": len(steps)}
