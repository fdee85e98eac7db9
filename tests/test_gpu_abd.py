"""Affine bodies through JOIN/UNION on the GPU vs the chain-rule oracle (oracle/abd.py):
tet cube whose top layers follow two affine bodies (x = t + A r), free vertices below,
fixed bottom. Gradient and PSD-projected Hessian over [free positions, A, t] within
1e-12 of the largest entry, the 9x9 / 9x3 / 3x9 / 3x3 block categories, and the generic
block-Jacobi PCG solving H x = g."""
import numpy as np
import pytest

from oracle import abd
from yasps_amd.scenes import abd_cube_scene, build_abd_scene

pytestmark = pytest.mark.gpu
REL = 1e-12


def _rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("n,layers,bodies,method", [(2, 1, 2, 2), (3, 2, 2, 2), (3, 2, 3, 0),
                                                    (3, 2, 2, 1)])
def test_abd_gradient_hessian_match_oracle(n, layers, bodies, method):
    from yasps_amd import device as dev
    print(f"abd n={n} layers={layers} bodies={bodies} method={method}", flush=True)
    sc = abd_cube_scene(n=n, abd_layers=layers, bodies=bodies, seed=n + bodies)
    s0, hd = build_abd_scene(sc, projection_method=method)
    hs = s0.minimizer.computeNumericValue()
    dev.torch().cuda.synchronize()
    g, H = abd.assemble(sc, method=method)
    assert _rel(dev.to_host(hs.gradient.value), g) < REL
    assert _rel(hs.dense(), H) < REL
    dims = set(zip(hs.static.dims[0::2], hs.static.dims[1::2]))
    assert (9, 9) in dims and (3, 3) in dims and hs.full is None


def test_abd_solve_matches_dense():
    from yasps_amd import device as dev
    sc = abd_cube_scene(n=3, abd_layers=2, bodies=2, seed=7)
    s0, _ = build_abd_scene(sc, projection_method=2, threshold_mode="relative_residual")
    sol = s0.minimizeEnergy(tolerance=1e-10)
    x = np.concatenate([dev.to_host(v) for v in sol])
    g, H = abd.assemble(sc, method=2)
    assert np.linalg.norm(H @ x - g) <= 2e-10 * np.linalg.norm(g)
    xs = np.linalg.solve(H, g)
    assert np.linalg.norm(x - xs) <= 1e-6 * np.linalg.norm(xs)
    assert [v.numel() for v in sol] == [3 * sc.n_free, 9 * sc.n_bodies, 3 * sc.n_bodies]


@pytest.mark.parametrize("solver,options", [("jacobian", {}),
                                            ("mas", {"max_levels": 3, "coarsest_level_weight": 1.0})])
def test_abd_mixed_precision_reaches_fp64_residual(solver, options):
    """Mixed precision on the split operator of an affine-body layout (FP32 sub-block
    values in the Krylov loop, FP64 refinement): the true residual on the oracle's FP64
    Hessian reaches the tolerance and the solution matches the dense solve."""
    from yasps_amd import device as dev
    print(f"abd mixed {solver}", flush=True)
    sc = abd_cube_scene(n=3, abd_layers=2, bodies=2, seed=7)
    s0, _ = build_abd_scene(sc, projection_method=2, solver_name=solver,
                            threshold_mode="relative_residual", precision="mixed", **options)
    tol = 1e-9
    sol = s0.minimizeEnergy(tolerance=tol)
    st = s0.minimizer.linearSolver.statistics
    assert st["converged"] and st["precision"] == "mixed"
    x = np.concatenate([dev.to_host(v) for v in sol])
    g, H = abd.assemble(sc, method=2)
    rel = np.linalg.norm(H @ x - g) / np.linalg.norm(g)
    assert rel <= 1.02 * tol, (rel, st["outer_history"])
    xs = np.linalg.solve(H, g)
    assert np.linalg.norm(x - xs) <= 1e-5 * np.linalg.norm(xs)
    assert st["outer_iterations"] >= 2          # beyond FP32 resolution


def test_abd_energy_values():
    from yasps_amd import device as dev
    from oracle import energies
    sc = abd_cube_scene(n=2, abd_layers=1, bodies=2, seed=2)
    s0, hd = build_abd_scene(sc)
    cs = sc.cube
    x = sc.positions()
    dinv, vol = energies.snh_rest_data(cs.rest[cs.tets])
    e_ref = energies.snh_energy(x[cs.tets], dinv, vol, cs.mu, cs.lam, cs.dt)
    e = dev.to_host(hd["snh"].compute().value)
    assert _rel(e, e_ref) < REL
    nf, na = sc.n_fixed, sc.n_abd
    ia = energies.inertia_energy(x[nf:nf + na], cs.last_position[nf:nf + na], cs.velocity[nf:nf + na],
                                 cs.mass[nf:nf + na], cs.dt)
    assert _rel(dev.to_host(hd["inertia_abd"].compute().value), ia) < REL


def test_abd_index_pipeline_bit_exact():
    """JOIN inside UNION (affine vertices carry their body's A and t slots): index
    buffers, permutations, coordinates and the global block layout against the oracle."""
    from yasps_amd import device as dev
    from oracle import indices as ix
    sc = abd_cube_scene(n=3, abd_layers=2, bodies=2, seed=5)
    s0, hd = build_abd_scene(sc)
    hs = s0.minimizer.computeNumericValue()
    cs = sc.cube
    nf, na, nm, nb = sc.n_fixed, sc.n_abd, sc.n_free, sc.n_bodies
    body = sc.body_of.astype(np.uint32).reshape(-1, 1)
    affine = ix.Join(body, [ix.Leaf(9, 1), ix.Leaf(3, 2)])
    union = ix.Union(counts=[nf, na, nm], members=[[ix.Leaf(3, None)], [affine], [ix.Leaf(3, 0)]])
    wrt_start = [0, 3 * nm, 3 * nm + 9 * nb]
    progs = [(ix.Concat([ix.Join(cs.tets.astype(np.uint32), [union])]), cs.tets.shape[0]),
             (ix.Concat([ix.Leaf(9, 1)]), nb),
             (ix.Concat([ix.Join(body, [ix.Leaf(9, 1), ix.Leaf(3, 2)])]), na),
             (ix.Concat([ix.Leaf(3, 0)]), nm)]
    terms = [ix.term_indices(p, cnt, wrt_start) for p, cnt in progs]
    layout = ix.build_layout(terms)
    for ti, ot in zip(hs.static_index, terms):
        K = ti.K
        assert K == ot.K
        np.testing.assert_array_equal(dev.to_host(ti.indices, "uint32").reshape(-1, K), ot.indices)
        np.testing.assert_array_equal(dev.to_host(ti.sizes, "uint16").reshape(-1, K), ot.sizes)
        np.testing.assert_array_equal(dev.to_host(ti.permutations).reshape(-1, K),
                                      ot.compression.permutations)
        np.testing.assert_array_equal(dev.to_host(ti.coord_outer, "uint32"),
                                      ot.compression.coordinates_outer)
    np.testing.assert_array_equal(dev.to_host(hs.static.positions, "uint32"), layout.block_positions)
    assert hs.static.starts == layout.blocks_start_indices
    assert hs.static.counts == layout.block_counts
    assert hs.static.dims == layout.block_dimensions
    for ti, lk in zip(hs.static_index, layout.lookups):
        np.testing.assert_array_equal(dev.to_host(ti.lookups, "uint32"), lk)
