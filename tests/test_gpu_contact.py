"""Contact scene at test size: one affine and one free voxelised ball
resting on a cloth with a pinned border, SNH on the balls, StVK membrane and dihedral
bending on the cloth, inertia, affine orthogonality, and dynamic IPC point-triangle /
edge-edge barrier pairs through the [fixed, affine, moving] vertex union. Gradient and
PSD-projected Hessian over [moving positions, A, t] against the dense chain-rule oracle
(oracle/contact.py, element terms from tests/golden/contact_small.npz) within 1e-12 of
the largest entry, before and after the pair sets change (dynamic re-index), and the
generic block-Jacobi PCG reaching the requested residual on the assembled system."""
from pathlib import Path

import numpy as np
import pytest

from oracle import contact
from yasps_amd.scenes import contact_scene, build_contact_scene

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).with_name("golden") / "contact_small.npz"
SCENE = dict(bodies_x=2, bodies_z=1, cells=4, radius=0.12, cloth_m=9, max_pairs=10 ** 9, seed=0)
REL = 1e-12


def golden_terms():
    z = np.load(GOLDEN)
    out = {}
    for kind in ("stvk_membrane", "bending", "point_triangle", "edge_edge"):
        g, packed = z[f"{kind}_g"], z[f"{kind}_H"]
        n = g.shape[1]
        iu = np.triu_indices(n)
        H = np.zeros((g.shape[0], n, n))
        H[:, iu[0], iu[1]] = packed
        H[:, iu[1], iu[0]] = packed
        out[kind] = (g, H)
    return out, z["checksum"]


def checksum(cs):
    parts = [cs.positions(), cs.rest, cs.tets, cs.triangles, cs.hinges, cs.pt, cs.ee]
    return np.array([float(np.asarray(p, dtype=np.float64).sum()) for p in parts])


def _rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300)


def test_contact_scene_matches_oracle_and_solves():
    from yasps_amd import device as dev
    cs = contact_scene(**SCENE)
    terms, chk = golden_terms()
    np.testing.assert_allclose(checksum(cs), chk, rtol=1e-13)
    s0, h = build_contact_scene(cs, threshold_mode="relative_residual")
    rng = np.random.default_rng(5)
    npt, nee = cs.pt.shape[0], cs.ee.shape[0]
    sets = [(np.arange(npt), np.arange(nee)),
            (np.sort(rng.choice(npt, npt // 2, replace=False)), np.arange(nee)),
            (np.arange(npt), np.sort(rng.choice(nee, nee // 3, replace=False)))]
    for step, (pr, er) in enumerate(sets):
        print(f"contact pair set {step}: {pr.size} point-triangle, {er.size} edge-edge", flush=True)
        if step:
            for key, rows, pairs in (("pt", pr, cs.pt), ("ee", er, cs.ee)):
                h[key + "_primitive"].updateNumInstances(rows.size)
                h[key + "_connectivity"].updateConnectivity(pairs[rows])
        hs = s0.minimizer.computeNumericValue()
        dev.torch().cuda.synchronize()
        g, H = contact.assemble(cs, 2, terms, pr, er)
        assert hs.dynamic.num_blocks > 0
        assert _rel(dev.to_host(hs.gradient.value), g) < REL
        assert _rel(hs.dense(), H) < REL
        # the solve operator: every block split into mirrored 3x3 sub-blocks
        assert hs.full is None and hs.split is not None
        from yasps_amd.hessian import vector
        v = np.random.default_rng(step).standard_normal(H.shape[0])
        y = dev.to_host(hs.spmv(vector(H.shape[0], dev.to_device(v, "float64"))).value)
        assert np.abs(y - H @ v).max() <= 1e-12 * np.abs(H @ v).max()
        sol = s0.minimizeEnergy(tolerance=1e-10)
        x = np.concatenate([dev.to_host(v) for v in sol])
        assert [v.numel() for v in sol] == [3 * cs.n_free, 9 * cs.n_bodies, 3 * cs.n_bodies]
        assert np.linalg.norm(H @ x - g) <= 2e-10 * np.linalg.norm(g)


def test_split_operator_long_rows_match_short_rows():
    """ysp_pcg_solve_full with a long-row threshold of 4 (nearly every block row handled
    by a workgroup of its own) follows the default threshold's iterates (five iterations
    agree to rounding) and reaches the requested residual of the assembled system."""
    import ctypes as C
    from yasps_amd import _lib, device as dev
    cs = contact_scene(**SCENE)
    s0, _ = build_contact_scene(cs, threshold_mode="relative_residual")
    s0.minimizeEnergy(tolerance=1e-10)
    hs = s0.minimizer.activeHessian
    f = hs.split
    n = hs.rows
    g, H = contact.assemble(cs, 2, golden_terms()[0])
    lib = _lib.load()

    def solve(threshold, max_it, tol):
        x = dev.zeros(n, "float64")
        work = dev.zeros(5 * n, "float64")
        scalars = dev.zeros(12288, "float64")
        st = _lib.PcgStats()
        rc = lib.ysp_pcg_solve_full(
            max_it, tol, 1, dev.ptr(f.row_ptr), dev.ptr(f.col), dev.ptr(f.v64), 0, f.nnz, threshold,
            dev.ptr(hs.diagonal_blocks_inverse), _lib.u32_array(hs.diagonal_blocks_start_cpu[:-1]),
            _lib.u32_array([w.numInstances for w in hs.wrt]), _lib.u32_array([w.size for w in hs.wrt]),
            _lib.u32_array(hs.gradient_segments_start_cpu[:-1]), len(hs.wrt),
            dev.ptr(hs.gradient.value), n, dev.ptr(x), None, 1, dev.ptr(work), dev.ptr(scalars),
            C.byref(st), dev.stream())
        dev.torch().cuda.synchronize()
        return rc, dev.to_host(x)

    early = []
    for threshold in (0, 4):
        print(f"long-row threshold {threshold}", flush=True)
        rc, x5 = solve(threshold, 5, 1e-30)
        assert rc == -4
        early.append(x5)
        rc, x = solve(threshold, 20000, 1e-10)
        assert rc >= 0
        assert np.linalg.norm(H @ x - g) <= 2e-10 * np.linalg.norm(g)
    assert np.linalg.norm(early[0] - early[1]) <= 1e-11 * np.linalg.norm(early[0])


def test_mas_on_split_operator_solves_contact_scene():
    """MAS PCG on the split operator (an affine body's 12 DoFs as four 3-DoF nodes; 32-DoF
    leaf domains, 4 levels; static and dynamic blocks at one position summed before the
    domain matrices) reaches the requested residual of the assembled system."""
    from yasps_amd import device as dev
    print("mas on the split operator", flush=True)
    cs = contact_scene(**SCENE)
    s0, _ = build_contact_scene(cs, solver_name="mas", threshold_mode="relative_residual",
                                max_domain_dofs=32, max_levels=4)
    sol = s0.minimizeEnergy(tolerance=1e-10)
    x = np.concatenate([dev.to_host(v) for v in sol])
    g, H = contact.assemble(cs, 2, golden_terms()[0])
    st = s0.minimizer.linearSolver.statistics
    assert st["converged"] and st["iterations"] > 0 and st["inverse_status"] == 0
    assert np.linalg.norm(H @ x - g) <= 2e-10 * np.linalg.norm(g)
