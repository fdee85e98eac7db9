"""Dynamic terms (hessian.pyx:846-880 re-index, static + dynamic block storage summed by
the SpMV, matrix.pyx:45-85): a point-point barrier over a dynamic contact primitive on
the tet cube, re-indexed on every assembly. Gradient and Hessian against the dense
oracle (cube SNH + inertia from oracle/newton.py, barrier elements from
oracle/element_energies.py, PSD-projected per element), before and after the pair set
changes size, and the block-Jacobi PCG reaching the requested residual on the sum."""
import numpy as np
import pytest

from yasps_amd.scenes import cube_scene, build_yasps_scene
from oracle import element_energies as ee, newton, projection

pytestmark = pytest.mark.gpu

REL = 1e-12
DHAT, KAPPA = 0.15, 1.0       # squared-distance barrier: grid neighbours (d^2 ~ 0.11) are active


def _rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300)


def _pairs(cs, count, seed):
    """Vertex pairs closer than the barrier's reach (squared distance < DHAT)."""
    rng = np.random.default_rng(seed)
    x = cs.position
    nv = x.shape[0]
    out = []
    while len(out) < count:
        a, b = rng.integers(0, nv, 2)
        if a != b and ((x[a] - x[b]) ** 2).sum() < DHAT and (cs.n_fixed <= a or cs.n_fixed <= b):
            out.append((a, b))
    return np.array(out, dtype=np.uint32)


def _oracle(cs, pairs):
    """Dense gradient and projected Hessian over the moving DoFs."""
    return _oracle_with(cs, pairs, DHAT, KAPPA)


def _oracle_with(cs, pairs, dhat, kappa):
    """Dense gradient and projected Hessian over the moving DoFs, barrier (dhat, kappa)."""
    st = newton.fast_structure(cs)
    asm = newton.evaluate(cs, st)
    A = newton.bsr_matrix(st, asm.values).toarray()
    g = asm.gradient.copy()
    X = cs.position[pairs].reshape(len(pairs), 6)
    _, ge, He = ee.evaluate("point_point", X, np.zeros((len(pairs), 0)), np.array([dhat, kappa]))
    He = projection.project_batch(He, 2)
    nf = cs.n_fixed
    for e, (a, b) in enumerate(pairs):
        loc, glo = [], []
        for k, v in enumerate((a, b)):
            if v >= nf:
                loc += [3 * k + i for i in range(3)]
                glo += [3 * (v - nf) + i for i in range(3)]
        loc, glo = np.array(loc), np.array(glo)
        g[glo] += ge[e][loc]
        A[np.ix_(glo, glo)] += He[e][np.ix_(loc, loc)]
    return g, A


def test_dynamic_contact_term_reindexed_and_solved():
    from yasps_amd import device as dev
    from yasps_amd.energies import point_point
    print("dynamic point-point term n=3", flush=True)
    cs = cube_scene(3, seed=3)
    s0, h = build_yasps_scene(cs, threshold_mode="relative_residual")
    m = h["mesh"]
    p1 = _pairs(cs, 40, seed=1)
    contacts = m.addPrimitive("contacts", numInstances=len(p1), isDynamic=True)
    c2v = contacts.addConnectivity("c2v", m.vertices, p1, 2)
    cp = contacts.addAttribute("positions", through=c2v, source=m.vertices["position"])
    pp = contacts.addAttribute("pp", computed_attribute=point_point(cp, DHAT, KAPPA))
    s0.addEnergy(pp, projection_method=2, dynamic_instances=True)
    for step, pairs in enumerate((p1, _pairs(cs, 57, seed=2), _pairs(cs, 23, seed=3))):
        print(f"  pair set {step}: {len(pairs)} pairs", flush=True)
        if step:
            contacts.updateNumInstances(len(pairs))
            c2v.updateConnectivity(pairs)
        hs = s0.minimizer.computeNumericValue()
        dev.torch().cuda.synchronize()
        g, A = _oracle(cs, pairs)
        assert hs.dynamic.num_blocks > 0
        assert _rel(dev.to_host(hs.gradient.value), g) < REL
        assert _rel(hs.dense(), A) < REL
        x = dev.to_host(s0.minimizeEnergy(tolerance=1e-10)[0])
        assert np.linalg.norm(g - A @ x) <= 1.5e-10 * np.linalg.norm(g)


def test_mas_with_dynamic_term_solves():
    """MAS PCG on the cube plus the dynamic point-point term: blocks present in both parts
    are summed into the domain matrices; the solve reaches the requested residual."""
    from yasps_amd import device as dev
    from yasps_amd.energies import point_point
    print("mas with a dynamic term", flush=True)
    cs = cube_scene(3, seed=3)
    s0, h = build_yasps_scene(cs, solver_name="mas", threshold_mode="relative_residual")
    m = h["mesh"]
    pairs = _pairs(cs, 40, seed=1)
    contacts = m.addPrimitive("contacts", numInstances=len(pairs), isDynamic=True)
    c2v = contacts.addConnectivity("c2v", m.vertices, pairs, 2)
    cp = contacts.addAttribute("positions", through=c2v, source=m.vertices["position"])
    pp = contacts.addAttribute("pp", computed_attribute=point_point(cp, DHAT, KAPPA))
    s0.addEnergy(pp, projection_method=2, dynamic_instances=True)
    x = dev.to_host(s0.minimizeEnergy(tolerance=1e-10)[0])
    g, A = _oracle(cs, pairs)
    st = s0.minimizer.linearSolver.statistics
    assert st["converged"] and st["inverse_status"] == 0
    assert np.linalg.norm(g - A @ x) <= 1.5e-10 * np.linalg.norm(g)


@pytest.mark.parametrize("solver", ["jacobian", "mas"])
def test_long_rows_in_the_product(solver):
    """A hub vertex paired with 200 others gives a block row of more than 128 entries,
    which the 3x3 PCG product hands to a workgroup of its own: the solve still reaches
    the requested residual of the dense oracle's system (block-Jacobi and MAS)."""
    from yasps_amd import device as dev
    from yasps_amd.energies import point_point
    print(f"long rows, {solver}", flush=True)
    cs = cube_scene(6, seed=5)
    s0, h = build_yasps_scene(cs, solver_name=solver, threshold_mode="relative_residual")
    m = h["mesh"]
    nf, nv = cs.n_fixed, cs.rest.shape[0]
    hub = nv - 1
    others = np.random.default_rng(4).choice(np.arange(nf, nv - 1), 200, replace=False)
    pairs = np.stack([np.full(200, hub), others], 1).astype(np.uint32)
    contacts = m.addPrimitive("contacts", numInstances=len(pairs), isDynamic=True)
    c2v = contacts.addConnectivity("c2v", m.vertices, pairs, 2)
    cp = contacts.addAttribute("positions", through=c2v, source=m.vertices["position"])
    pp = contacts.addAttribute("pp", computed_attribute=point_point(cp, 2.0, 1e-3))
    s0.addEnergy(pp, projection_method=2, dynamic_instances=True)
    x = dev.to_host(s0.minimizeEnergy(tolerance=1e-10)[0])
    hs = s0.minimizer.activeHessian
    rp = dev.to_host(hs.full.row_ptr, "uint32").astype(np.int64)
    assert np.diff(rp).max() > 128
    g, A = _oracle_with(cs, pairs, 2.0, 1e-3)
    assert s0.minimizer.linearSolver.statistics["converged"]
    assert np.linalg.norm(g - A @ x) <= 1.5e-10 * np.linalg.norm(g)


def test_dynamic_term_emptied_and_refilled():
    """The dynamic pair set shrinks to zero instances and grows back (the reference's
    sequence with an empty active set, test_hessian_sparsity_permutation_gpu.py:256-271):
    with no pairs the Hessian is the static part alone; the solve runs on it; refilled, the
    dense oracle matches again."""
    from yasps_amd import device as dev
    from yasps_amd.energies import point_point
    print("dynamic term emptied and refilled", flush=True)
    cs = cube_scene(3, seed=3)
    s0, h = build_yasps_scene(cs, threshold_mode="relative_residual")
    m = h["mesh"]
    p1 = _pairs(cs, 40, seed=1)
    contacts = m.addPrimitive("contacts", numInstances=len(p1), isDynamic=True)
    c2v = contacts.addConnectivity("c2v", m.vertices, p1, 2)
    cp = contacts.addAttribute("positions", through=c2v, source=m.vertices["position"])
    pp = contacts.addAttribute("pp", computed_attribute=point_point(cp, DHAT, KAPPA))
    s0.addEnergy(pp, projection_method=2, dynamic_instances=True)
    empty = np.zeros((0, 2), dtype=np.uint32)
    for step, pairs in enumerate((p1, empty, _pairs(cs, 31, seed=5), empty)):
        print(f"  pair set {step}: {len(pairs)} pairs", flush=True)
        contacts.updateNumInstances(len(pairs))
        if len(pairs):
            c2v.updateConnectivity(pairs)
        for solver in ("jacobian", "mas"):
            s0.minimizer.setSolver(solver, threshold_mode="relative_residual")
            hs = s0.minimizer.computeNumericValue()
            dev.torch().cuda.synchronize()
            g, A = _oracle(cs, pairs)
            if len(pairs) == 0:
                assert hs.dynamic.num_blocks == 0 and sum(hs.block_counts_dynamic) == 0
            assert _rel(dev.to_host(hs.gradient.value), g) < REL
            assert _rel(hs.dense(), A) < REL
            x = dev.to_host(s0.minimizeEnergy(tolerance=1e-10)[0])
            assert s0.minimizer.linearSolver.statistics["converged"], solver
            assert np.linalg.norm(g - A @ x) <= 1.5e-10 * np.linalg.norm(g), solver
