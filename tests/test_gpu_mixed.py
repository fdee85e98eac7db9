"""Mixed-precision solve (FP32 block values and preconditioner data in the Krylov loop,
FP64 iterative refinement): the true FP64 relative residual ||g - H x|| / ||g||, computed
by the oracle from its own FP64 Hessian, reaches the requested tolerance, and the error
against the FP64 solution is reported and small."""
import numpy as np
import pytest

from yasps_amd.scenes import cube_scene, build_yasps_scene
from oracle import newton, pcg as opcg

pytestmark = pytest.mark.gpu


def _oracle(cs):
    st = newton.fast_structure(cs)
    asm = newton.evaluate(cs, st)
    A = newton.bsr_matrix(st, asm.values)
    return A, asm.gradient


@pytest.mark.parametrize("solver,options", [
    ("jacobian", {}),
    ("mas", {"max_levels": 3, "coarsest_level_weight": 1.0}),
])
@pytest.mark.parametrize("tol", [1e-6, 1e-10])
def test_mixed_refinement_reaches_fp64_residual(solver, options, tol):
    from yasps_amd import device as dev
    print(f"mixed {solver} n=8 tol={tol}", flush=True)
    cs = cube_scene(8, seed=0)
    s0, _ = build_yasps_scene(cs, solver_name=solver, precision="mixed", **options)
    x = dev.to_host(s0.minimizeEnergy(tolerance=tol)[0]).copy()
    st = s0.minimizer.linearSolver.statistics
    assert st["converged"] and st["precision"] == "mixed"
    A, b = _oracle(cs)
    rel = opcg.relative_residual(lambda v: A @ v, x, b)
    assert rel <= 1.02 * tol, (rel, st["outer_history"])
    assert abs(rel - st["relative_residual"]) <= 1e-3 * tol + 1e-15
    xs = np.linalg.solve(A.toarray(), b)
    err = np.linalg.norm(x - xs) / np.linalg.norm(xs)
    assert err < (1e-3 if tol >= 1e-6 else 1e-7), err
    if tol < 1e-8:
        # beyond what an FP32 operator resolves: refinement takes more than one step
        assert st["outer_iterations"] >= 2


def test_mixed_matches_fp64_path_iterations():
    """The first refinement step of the mixed path is the FP64 recurrence on FP32 values:
    its iteration count stays close to the FP64 path at the same tolerance."""
    from yasps_amd import device as dev
    cs = cube_scene(8, seed=1)
    s64, _ = build_yasps_scene(cs, threshold_mode="relative_residual")
    x64 = dev.to_host(s64.minimizeEnergy(tolerance=1e-6)[0]).copy()
    it64 = s64.minimizer.linearSolver.statistics["iterations"]
    s32, _ = build_yasps_scene(cs, precision="mixed")
    x32 = dev.to_host(s32.minimizeEnergy(tolerance=1e-6)[0]).copy()
    st = s32.minimizer.linearSolver.statistics
    assert st["iterations"] <= it64 + 20
    assert np.linalg.norm(x32 - x64) / np.linalg.norm(x64) < 1e-4
