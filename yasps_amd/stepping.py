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
(``ccd.compute_largest_step_size(0.8, ...)``, ``one_bunny.py:205-206``)."""
from __future__ import annotations

from . import device as dev


def newton_time_step(s0, moving, dt: float, tolerance: float = 1e-6, gradient_tol: float = 1e-1,
                     max_newton: int = 50, max_halvings: int = 8, first_step: float = 1.0,
                     max_iterations: int = 20000, step_bound=None) -> dict:
    """One time step on scene ``s0`` whose minimise target is ``moving["position"]`` (a
    primitive that also holds ``last_position`` and ``velocity``). Returns the Newton
    history (energies, largest gradient entries, steps taken, PCG iterations)."""
    torch = dev.torch()
    pos = moving["position"]
    last = pos.value.clone()
    moving["last_position"].updateValue(last)
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
        for _h in range(max_halvings):
            pos.updateValue(x0 - step * direction)
            if s0.minimizer.computeTotalEnergy() <= e0:
                break
            step *= 0.5
        else:
            pos.updateValue(x0)
            raise RuntimeError("newton_time_step: the line search found no decrease")
        steps.append(step)
        if gmax < gradient_tol:
            converged = True
            break
    moving["velocity"].updateValue((pos.value - last) / dt)
    return {"converged": converged, "energies": energies, "max_gradient": grads,
            "steps": steps, "pcg_iterations": pcg, "newton_iterations": len(steps)}
