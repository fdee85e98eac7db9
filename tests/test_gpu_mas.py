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


@pytest.mark.parametrize("levels,coarsest,target,dofs", [(4, 2.5, 8, 32), (3, 1.0, 8, 32),
                                                         (6, 2.5, 3, 32), (4, 2.5, 16, 48)])
def test_mas_apply_and_inverses_match_oracle(levels, coarsest, target, dofs):
    """(6 levels, 3 nodes per partition: levels 4 and 5 restrict from level 3 in a launch
    of their own; 48 scalars per domain: the 48 x 48 inversion and local solve.)"""
    from yasps_amd import device as dev
    print(f"mas apply n=5 levels={levels} target={target} domain dofs={dofs}", flush=True)
    cs, s0, hs = _mas_scene(5, seed=1, max_levels=levels, coarsest_level_weight=coarsest,
                            target_nodes_per_partition=target, max_domain_dofs=dofs)
    slv = s0.minimizer.linearSolver.implementation
    slv._ensure(hs, hs.rows)
    m = slv.device_tables
    m.update(hs.full)
    dev.torch().cuda.synchronize()
    assert int(m.status.item()) == 0
    h = slv.hierarchy
    assert h.num_levels >= (5 if levels == 6 else 2)
    A, _ = _oracle_matrix(cs)
    mats = omas.level_matrices(h, A.toarray())
    invs = omas.domain_inverses(h, mats)
    D = m.layout.domain_scalars
    assert D == (48 if dofs == 48 else 32)
    if dofs == 48:
        assert max(3 * d.size for lv in h.levels for d in lv.domains) > 32
    inv_dev = dev.to_host(m.inverses).reshape(-1, D, D)      # column-major per domain
    k = 0
    for l, lv in enumerate(h.levels):
        for d, ref in zip(lv.domains, invs[l]):
            P = ref.shape[0]
            got = inv_dev[k].T[:P, :P].astype(np.float64)
            assert np.abs(got - ref).max() <= 1e-5 * np.abs(ref).max(), (l, k)
            k += 1
    w = np.asarray(omas.level_weights(h, coarsest))
    dup = omas.duplicate_domains(h)
    scales = [[slv.duplicate_level_weight if f else 1.0 for f in lvl] for lvl in dup]
    rng = np.random.default_rng(11)
    for _ in range(3):
        r = rng.standard_normal(hs.rows)
        rt = dev.to_device(r, "float64")
        z = dev.empty(hs.rows, "float64")
        m.apply(rt, z)
        got = dev.to_host(z)
        ref = omas.apply(h, invs, r, w, scales)
        assert np.linalg.norm(got - ref) <= APPLY_TOL * np.linalg.norm(ref)


@pytest.mark.parametrize("dofs", [32, 48])
def test_mas_bf16_coarse_bank(dofs):
    """Coarse-level inverses stored BF16, level 0 FP32: every coarse inverse within BF16
    rounding (2^-8 of the domain's largest entry) of the oracle's, level 0 as before, the
    apply within 1e-2 norm-wise, and the mixed MAS-PCG still reaches the requested true
    FP64 residual (32- and 48-scalar domains)."""
    from yasps_amd import device as dev
    print(f"mas bf16 coarse bank n=5 domain dofs={dofs}", flush=True)
    cs, s0, hs = _mas_scene(5, seed=1, max_levels=4, coarsest_level_weight=2.5, coarse_bank="bf16",
                            max_domain_dofs=dofs, target_nodes_per_partition=dofs // 4)
    slv = s0.minimizer.linearSolver.implementation
    slv._ensure(hs, hs.rows)
    m = slv.device_tables
    m.update(hs.full)
    dev.torch().cuda.synchronize()
    assert int(m.status.item()) == 0
    h = slv.hierarchy
    A, _ = _oracle_matrix(cs)
    invs = omas.domain_inverses(h, omas.level_matrices(h, A.toarray()))
    D = m.layout.domain_scalars
    assert D == dofs
    n0 = len(h.levels[0].domains)
    assert m.bf16_first == n0 and m.inverses_bf16 is not None
    f32 = dev.to_host(m.inverses).reshape(-1, D, D)
    u16 = dev.to_host(m.inverses_bf16, "uint16").astype(np.uint32) << 16
    b16 = u16.view(np.float32).reshape(-1, D, D)
    k = 0
    for l, lv in enumerate(h.levels):
        for ref in invs[l]:
            P = ref.shape[0]
            if k < n0:
                got, tol = f32[k].T[:P, :P].astype(np.float64), 1e-5
            else:
                got, tol = b16[k - n0].T[:P, :P].astype(np.float64), 2.0 ** -8
            assert np.abs(got - ref).max() <= tol * np.abs(ref).max(), (l, k)
            k += 1
    assert k == n0 + b16.shape[0]
    w = np.asarray(omas.level_weights(h, 2.5))
    r = np.random.default_rng(12).standard_normal(hs.rows)
    z = dev.empty(hs.rows, "float64")
    m.apply(dev.to_device(r, "float64"), z)
    ref = omas.apply(h, invs, r, w)
    assert np.linalg.norm(dev.to_host(z) - ref) <= 1e-2 * np.linalg.norm(ref)
    cs8 = cube_scene(8, seed=0)
    s8, _ = build_yasps_scene(cs8, solver_name="mas", threshold_mode="relative_residual",
                              max_levels=4, precision="mixed", coarse_bank="bf16",
                              max_domain_dofs=dofs)
    x = dev.to_host(s8.minimizeEnergy(tolerance=1e-8)[0]).copy()
    assert s8.minimizer.linearSolver.statistics["converged"]
    A8, b8 = _oracle_matrix(cs8)
    assert opcg.relative_residual(lambda v: A8 @ v, x, b8) < 1.5e-8


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


@pytest.mark.parametrize("levels", [1, 2])
def test_mas_single_domain_closed_form(levels):
    """n=1 cube, one level-0 domain: the device apply equals w A^-1 r (one level) or
    A^-1 r + w R^T (R A R^T)^-1 R r (a second level of one coarse node, R = [I I I I]),
    within the FP32 bank's rounding."""
    from yasps_amd import device as dev
    print(f"mas closed form levels={levels}", flush=True)
    cs, s0, hs = _mas_scene(1, seed=2, max_levels=levels, coarsest_level_weight=2.5)
    slv = s0.minimizer.linearSolver.implementation
    slv._ensure(hs, hs.rows)
    m = slv.device_tables
    m.update(hs.full)
    assert slv.hierarchy.num_levels == levels
    A, _ = _oracle_matrix(cs)
    Ad = A.toarray()
    r = np.random.default_rng(3).standard_normal(hs.rows)
    if levels == 1:
        expect = 2.5 * np.linalg.solve(Ad, r)
    else:
        R = np.tile(np.eye(3), (1, hs.rows // 3))
        expect = np.linalg.solve(Ad, r) + 2.5 * R.T @ np.linalg.solve(R @ Ad @ R.T, R @ r)
    z = dev.empty(hs.rows, "float64")
    m.apply(dev.to_device(r, "float64"), z)
    got = dev.to_host(z)
    assert np.linalg.norm(got - expect) <= 1e-5 * np.linalg.norm(expect)
