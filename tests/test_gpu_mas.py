"""Device MAS preconditioner vs the MAS oracle on the tet cube: domain inverses, the apply
z = sum_l w_l P_l B_l^-1 R_l r (FP32 inverse bank and corrections, FP64 in/out: 1e-4
norm-wise), and MAS-PCG reaching the requested residual with both stopping rules."""
import numpy as np
import pytest

from yasps_amd.scenes import cube_scene, build_yasps_scene
from oracle import newton, mas as omas, pcg as opcg

pytestmark = pytest.mark.gpu

APPLY_TOL = 1e-4      # FP32 inverses / corrections: norm-wise relative bar of the apply


def _mas_scene(n, seed=0, **options):
    from yasps_amd import device as dev
    cs = cube_scene(n, seed=seed)
    s0, _ = build_yasps_scene(cs, solver_name="mas", **options)
    hs = s0.minimizer.computeNumericValue()
    dev.torch().cuda.synchronize()
    return cs, s0, hs


def _oracle_matrix(cs):
    st = newton.fast_structure(cs)
    asm = newton.evaluate(cs, st)
    return newton.bsr_matrix(st, asm.values), asm.gradient


@pytest.mark.parametrize("levels,coarsest", [(4, 2.5), (3, 1.0)])
def test_mas_apply_and_inverses_match_oracle(levels, coarsest):
    from yasps_amd import device as dev
    from yasps_amd.mas import DOMAIN_SCALARS
    print(f"mas apply n=5 levels={levels}", flush=True)
    cs, s0, hs = _mas_scene(5, seed=1, max_levels=levels, coarsest_level_weight=coarsest)
    slv = s0.minimizer.linearSolver.implementation
    slv._ensure(hs, hs.rows)
    m = slv.device_tables
    m.update(hs.full)
    dev.torch().cuda.synchronize()
    assert int(m.status.item()) == 0
    h = slv.hierarchy
    assert h.num_levels >= 2
    A, _ = _oracle_matrix(cs)
    mats = omas.level_matrices(h, A.toarray())
    invs = omas.domain_inverses(h, mats)
    D = DOMAIN_SCALARS
    inv_dev = dev.to_host(m.inverses).reshape(-1, D, D)      # column-major per domain
    k = 0
    for l, lv in enumerate(h.levels):
        for d, ref in zip(lv.domains, invs[l]):
            P = ref.shape[0]
            got = inv_dev[k].T[:P, :P].astype(np.float64)
            assert np.abs(got - ref).max() <= 1e-5 * np.abs(ref).max(), (l, k)
            k += 1
    w = np.asarray(omas.level_weights(h, coarsest))
    rng = np.random.default_rng(11)
    for _ in range(3):
        r = rng.standard_normal(hs.rows)
        rt = dev.to_device(r, "float64")
        z = dev.empty(hs.rows, "float64")
        m.apply(rt, z)
        got = dev.to_host(z)
        ref = omas.apply(h, invs, r, w)
        assert np.linalg.norm(got - ref) <= APPLY_TOL * np.linalg.norm(ref)


def test_mas_pcg_relative_residual():
    from yasps_amd import device as dev
    print("mas pcg n=8 relative residual 1e-8", flush=True)
    cs = cube_scene(8, seed=0)
    s0, _ = build_yasps_scene(cs, solver_name="mas", threshold_mode="relative_residual",
                              max_levels=4)
    x = dev.to_host(s0.minimizeEnergy(tolerance=1e-8)[0]).copy()
    stats = s0.minimizer.linearSolver.statistics
    assert stats["converged"] and stats["result"] == 0 and stats["inverse_status"] == 0
    A, b = _oracle_matrix(cs)
    rel = opcg.relative_residual(lambda v: A @ v, x, b)
    assert rel < 1.5e-8
    xs = np.linalg.solve(A.toarray(), b)
    assert np.linalg.norm(x - xs) / np.linalg.norm(xs) < 1e-6
    # second solve on the same structure reuses the hierarchy and tables
    s0.minimizeEnergy(tolerance=1e-8)
    stats2 = s0.minimizer.linearSolver.statistics
    assert stats2["hierarchy_build_count"] == 1 and not stats2["layout_rebuilt"]
    assert stats2["iterations"] == stats["iterations"]


def test_mas_pcg_reference_rule_matches_oracle_iterations():
    from yasps_amd import device as dev
    print("mas pcg n=4 reference rule 1e-14", flush=True)
    cs = cube_scene(4, seed=2)
    s0, _ = build_yasps_scene(cs, solver_name="mas", max_levels=3)
    tol = 1e-14
    x = dev.to_host(s0.minimizeEnergy(tolerance=tol)[0]).copy()
    stats = s0.minimizer.linearSolver.statistics
    assert stats["converged"]
    h = s0.minimizer.linearSolver.implementation.hierarchy
    A, b = _oracle_matrix(cs)
    invs = omas.domain_inverses(h, omas.level_matrices(h, A.toarray()))
    w = omas.level_weights(h)
    status, xo, it, _ = opcg.pcg(lambda v: A @ v, lambda v: omas.apply(h, invs, v, w), b, tol)
    assert status >= 0
    assert abs(stats["iterations"] - it) <= 2
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) < 1e-6


def test_mas_pcg_initial_guess():
    from yasps_amd import device as dev
    from yasps_amd.hessian import vector
    print("mas pcg n=6 initial guess", flush=True)
    cs, s0, hs = _mas_scene(6, seed=3, threshold_mode="relative_residual")
    slv = s0.minimizer.linearSolver
    assert slv.computeSolution(hs, hs.gradient, None, tolerance=1e-10, zero_initial_guess=True) == 0
    x_ref = dev.to_host(slv.solution).copy()
    it0 = slv.statistics["iterations"]
    guess = vector(hs.rows).updateValue(0.9 * x_ref)
    assert slv.computeSolution(hs, hs.gradient, guess, tolerance=1e-10) == 0
    x = dev.to_host(slv.solution)
    assert slv.statistics["iterations"] < it0
    assert np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref) < 1e-8
