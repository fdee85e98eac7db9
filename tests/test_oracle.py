"""CPU checks that pin the oracle (no GPU): golden vectors, finite differences, invariants,
brute force on tiny inputs and the reference's special cases."""
import json
from pathlib import Path

import numpy as np
import pytest

from oracle import energies, projection, indices as ix, assembly, pcg, newton
from yasps_amd.scenes import cube_scene, cube_tet_mesh

GOLDEN = Path(__file__).with_name("golden") / "snh_cube6.json"


# --------------------------------------------------------------------------- energies
@pytest.fixture(scope="module")
def golden():
    return json.loads(GOLDEN.read_text())


def _golden_inputs(g):
    x = np.asarray(g["tet_positions"]).reshape(6, 4, 3)
    # the fixture's IB is (TB^T)^-1 = Dm^-1, row-major
    dinv = np.asarray(g["IB"]).reshape(6, 3, 3)
    vol = np.asarray(g["vols"])
    return x, dinv, vol


def test_snh_matches_sympy_golden(golden):
    x, dinv, vol = _golden_inputs(golden)
    mu, lam = golden["mu"], golden["lam"]
    e = energies.snh_energy(x, dinv, vol, mu, lam, 1.0)
    g, H = energies.snh_gradient_hessian(x, dinv, vol, mu, lam, 1.0)
    ge = np.asarray(golden["energy"])
    gg = np.asarray(golden["gradient"])
    gH = np.asarray(golden["hessian"]).reshape(6, 12, 12)
    assert np.abs(e - ge).max() <= 1e-12 * np.abs(ge).max()
    assert np.abs(g - gg).max() <= 1e-12 * np.abs(gg).max()
    assert np.abs(H - gH).max() <= 1e-12 * np.abs(gH).max()
    # the closed-form block evaluation used by the CPU baseline (no projection here)
    g2, H2 = energies.snh_gradient_hessian_fast(x, dinv, vol, mu, lam, 1.0, method=0)
    assert np.abs(g2 - gg).max() <= 1e-12 * np.abs(gg).max()
    assert np.abs(H2 - gH).max() <= 1e-12 * np.abs(gH).max()


def test_snh_assembled_cube_matches_golden(golden):
    x, dinv, vol = _golden_inputs(golden)
    g, H = energies.snh_gradient_hessian(x, dinv, vol, golden["mu"], golden["lam"], 1.0)
    gs = np.zeros(24)
    Hs = np.zeros((24, 24))
    for t, idx in enumerate(golden["indices"]):
        dofs = (3 * np.asarray(idx)[:, None] + np.arange(3)).ravel()
        np.add.at(gs, dofs, g[t])
        np.add.at(Hs, (dofs[:, None], dofs[None, :]), H[t])
    ref = np.asarray(golden["hessian_sum"]).reshape(24, 24)
    assert np.abs(Hs - ref).max() <= 1e-12 * np.abs(ref).max()
    assert np.abs(gs - np.asarray(golden["gradient_sum"])).max() <= 1e-12 * np.abs(gs).max()


def _random_tets(seed, n=20, noise=0.3):
    rng = np.random.default_rng(seed)
    rest = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float)[None] + rng.normal(0, 0.1, (n, 4, 3))
    x = rest * rng.uniform(0.7, 1.3) + rng.normal(0, noise * 0.3, (n, 4, 3))
    dinv, vol = energies.snh_rest_data(rest)
    return x, dinv, vol


def test_snh_central_differences():
    x, dinv, vol = _random_tets(1)
    mu, lam, dt = 3e4, 8e4, 1e-2
    g, H = energies.snh_gradient_hessian(x, dinv, vol, mu, lam, dt)
    h = 1e-6
    for k in range(12):
        xp = x.reshape(-1, 12).copy()
        xm = xp.copy()
        xp[:, k] += h
        xm[:, k] -= h
        fd = (energies.snh_energy(xp.reshape(-1, 4, 3), dinv, vol, mu, lam, dt)
              - energies.snh_energy(xm.reshape(-1, 4, 3), dinv, vol, mu, lam, dt)) / (2 * h)
        assert np.abs(fd - g[:, k]).max() <= 1e-6 * np.abs(g).max()
        gp, _ = energies.snh_gradient_hessian(xp.reshape(-1, 4, 3), dinv, vol, mu, lam, dt)
        gm, _ = energies.snh_gradient_hessian(xm.reshape(-1, 4, 3), dinv, vol, mu, lam, dt)
        assert np.abs((gp - gm) / (2 * h) - H[:, :, k]).max() <= 1e-6 * np.abs(H).max()


def test_snh_rigid_motion_invariants():
    x, dinv, vol = _random_tets(2)
    mu, lam = 1e4, 4e4
    g, H = energies.snh_gradient_hessian(x, dinv, vol, mu, lam, 1.0)
    # translation: gradient sums to zero, H annihilates translations
    assert np.abs(g.reshape(-1, 4, 3).sum(1)).max() <= 1e-10 * np.abs(g).max()
    t = np.tile(np.eye(3), (4, 1))
    assert np.abs(H @ t).max() <= 1e-10 * np.abs(H).max()
    # rotation: energy invariant, gradient rotates
    th = 0.7
    R = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
    e0 = energies.snh_energy(x, dinv, vol, mu, lam, 1.0)
    e1 = energies.snh_energy(x @ R.T, dinv, vol, mu, lam, 1.0)
    assert np.abs(e0 - e1).max() <= 1e-10 * np.abs(e0).max()
    g1, _ = energies.snh_gradient_hessian(x @ R.T, dinv, vol, mu, lam, 1.0)
    assert np.abs(g1.reshape(-1, 4, 3) - g.reshape(-1, 4, 3) @ R.T).max() <= 1e-10 * np.abs(g).max()


def test_snh_rest_state_is_stress_free():
    """At rest (F = I) P = mu (1 - 1/4) I + lam (1 - alpha) I = 0 with alpha = 1 + 3/4 mu/lam."""
    x, dinv, vol = _random_tets(3, noise=0.0)
    rest = np.linalg.inv(dinv)      # columns are rest edges
    xr = np.concatenate([np.zeros((len(rest), 1, 3)), np.transpose(rest, (0, 2, 1))], axis=1)
    g, _ = energies.snh_gradient_hessian(xr, dinv, vol, 1e4, 4e4, 1.0)
    assert np.abs(g).max() < 1e-9


def test_inertia_closed_form():
    rng = np.random.default_rng(4)
    x, xp, v = rng.normal(size=(3, 10, 3))
    m = rng.uniform(0.1, 2, 10)
    dt = 0.01
    e = energies.inertia_energy(x, xp, v, m, dt)
    g, H = energies.inertia_gradient_hessian(x, xp, v, m, dt)
    target = xp + v * dt - np.array([0, 9.8 * dt * dt, 0])
    np.testing.assert_allclose(e, 0.5 * m * ((x - target) ** 2).sum(1), rtol=1e-14)
    np.testing.assert_allclose(g, m[:, None] * (x - target), rtol=1e-14)
    np.testing.assert_allclose(H, m[:, None, None] * np.eye(3), rtol=1e-14)


# --------------------------------------------------------------------------- projection
def test_projection_clamp_rules_and_fast_path():
    rng = np.random.default_rng(5)
    Q, _ = np.linalg.qr(rng.normal(size=(12, 12)))
    lam = np.linspace(-2, 3, 12)
    A = (Q * lam) @ Q.T
    P2 = projection.project(A, 2)
    P1 = projection.project(A, 1)
    np.testing.assert_allclose(P2, (Q * np.maximum(lam, 0)) @ Q.T, atol=1e-12)
    np.testing.assert_allclose(P1, (Q * np.abs(lam)) @ Q.T, atol=1e-12)
    np.testing.assert_array_equal(projection.project(A, 0), A)
    B = (Q * np.abs(lam)) @ Q.T
    assert projection.is_psd(B)
    np.testing.assert_array_equal(projection.project(B, 2), B)   # PSD fast path: unchanged


def test_projection_scalar_and_small_cases():
    assert projection.project(np.array([[-3.0]]), 1)[0, 0] == 3.0
    assert projection.project(np.array([[-3.0]]), 2)[0, 0] == 1e-6
    assert projection.project(np.array([[5e-7]]), 2)[0, 0] == 1e-6
    assert projection.project(np.array([[2.0]]), 2)[0, 0] == 2.0
    assert projection.project_scalar_codegen(-2.0, 2) == 0.0
    assert projection.project_scalar_codegen(-2.0, 1) == 2.0
    A = np.array([[1.0, 2.0], [2.0, 1.0]])      # eigenvalues 3, -1
    np.testing.assert_allclose(projection.project(A, 2), 1.5 * np.ones((2, 2)), atol=1e-14)


@pytest.mark.parametrize("seed", range(5))
def test_psd_test_agrees_with_eigenvalues_away_from_the_tolerance(seed):
    rng = np.random.default_rng(10 + seed)
    for _ in range(40):
        n = rng.integers(2, 13)
        Q, _ = np.linalg.qr(rng.normal(size=(n, n)))
        lam = rng.uniform(0.1, 5.0, n)
        if rng.uniform() < 0.5:
            lam[rng.integers(n)] = -rng.uniform(1e-3, 1.0)
        if rng.uniform() < 0.3:
            lam[rng.integers(n)] = 0.0                            # singular PSD
        A = (Q * lam) @ Q.T
        A = 0.5 * (A + A.T)
        assert projection.is_psd(A) == bool(lam.min() >= 0.0)
    batch = rng.normal(size=(50, 6, 6))
    batch = batch @ np.transpose(batch, (0, 2, 1)) - 0.5 * np.eye(6)
    np.testing.assert_array_equal(projection.is_psd_batch(batch),
                                  [projection.is_psd(b) for b in batch])


def test_jacobi_block_inverse_rule():
    rng = np.random.default_rng(6)
    Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    lam = np.array([4.0, -2.0, 5e-7])
    B = (Q * lam) @ Q.T
    inv = projection.jacobi_block_inverse(B[None])[0]
    expect = (Q * np.array([0.25, 0.5, 5e-7])) @ Q.T
    np.testing.assert_allclose(inv, expect, atol=1e-14)


# --------------------------------------------------------------------------- indices
def _dense_reference(idx_rows, local_H, local_g, n):
    H = np.zeros((n, n))
    g = np.zeros(n)
    for dofs, Hl, gl in zip(idx_rows, local_H, local_g):
        dofs = np.asarray(dofs)
        keep = dofs >= 0
        d = dofs[keep]
        np.add.at(H, (d[:, None], d[None, :]), Hl[np.ix_(keep, keep)])
        np.add.at(g, d, gl[keep])
    return H, g


def _join_union_case(conn, n_fixed, n_moving, rng):
    union = ix.Union(counts=[n_fixed, n_moving], members=[[ix.Leaf(3, None)], [ix.Leaf(3, 0)]])
    prog = ix.Concat([ix.Join(np.asarray(conn, np.uint32), [union])])
    N, A = np.asarray(conn).shape
    term = ix.term_indices(prog, N, [0, 3 * n_moving])
    layout = ix.build_layout([term])
    n = 3 * n_moving
    local_H = []
    for _ in range(N):
        M = rng.normal(size=(3 * A, 3 * A))
        local_H.append(M + M.T)
    local_H = np.asarray(local_H)
    local_g = rng.normal(size=(N, 3 * A))
    grad = np.zeros(n)
    vals = np.zeros(layout.total_size)
    gs, ds = assembly.diagonal_layout([3], [n_moving])
    diag = np.zeros(ds[-1])
    assembly.scatter_term(local_g, local_H, term, layout.lookups[0], grad, vals, diag, gs, ds)
    dof_rows = []
    for row in np.asarray(conn):
        d = []
        for v in row:
            d += [-1] * 3 if v < n_fixed else list(3 * (v - n_fixed) + np.arange(3))
        dof_rows.append(d)
    Hd, gd = _dense_reference(dof_rows, local_H, local_g, n)
    return term, layout, grad, vals, diag, Hd, gd, n


def test_scatter_with_repeated_and_fixed_vertices_matches_dense():
    """Repeated vertices inside one element (as in the reference sparsity test's
    [3,1,5,0], [1,1,2,0], [0,2,0,2], [5,5,5,5]) and fixed (non-target) vertices."""
    rng = np.random.default_rng(7)
    conn = [[3, 1, 5, 0], [1, 1, 2, 0], [0, 2, 0, 2], [5, 5, 5, 5], [6, 7, 3, 2], [7, 6, 5, 4]]
    term, layout, grad, vals, diag, Hd, gd, n = _join_union_case(conn, 2, 6, rng)
    D = assembly.dense_from_blocks(layout, vals, n)
    np.testing.assert_allclose(D, Hd, atol=1e-12)
    np.testing.assert_allclose(grad, gd, atol=1e-12)
    np.testing.assert_allclose(diag.reshape(-1, 3, 3),
                               np.array([Hd[3 * v:3 * v + 3, 3 * v:3 * v + 3] for v in range(n // 3)]),
                               atol=1e-12)
    x = rng.normal(size=n)
    np.testing.assert_allclose(assembly.spmv(layout, vals, x), Hd @ x, atol=1e-10)


def test_local_compression_hand_checked():
    """Element [1, 1, 2, 0] on union [fixed(1 vertex), moving]: slots (idx, size) =
    (2,3)(0,0) (2,3)(0,0) (5,3)(0,0) (1,3)(0,0); the repeat gets -perm of the first."""
    conn = np.array([[1, 1, 2, 0]], np.uint32)
    union = ix.Union(counts=[1, 4], members=[[ix.Leaf(3, None)], [ix.Leaf(3, 0)]])
    prog = ix.Concat([ix.Join(conn, [union])])
    idx, sz = ix.get_indices(prog, 1, [0, 12])
    np.testing.assert_array_equal(idx[0], [2, 0, 2, 0, 5, 0, 1, 0])
    np.testing.assert_array_equal(sz[0], [3, 0, 3, 0, 3, 0, 3, 0])
    comp = ix.compress_indices(idx, sz)
    np.testing.assert_array_equal(comp.permutations[0], [1, 0, -1, 0, 4, 0, 7, 0])
    assert comp.gradient_sizes[0] == 9 and comp.unique_counts[0] == 2
    np.testing.assert_array_equal(comp.coordinates_outer, [0, 3])
    coords, dims = ix.compute_coordinates(idx, comp.permutations, sz, comp.coordinates_outer)
    np.testing.assert_array_equal(coords, [0, 0, 0, 3, 3, 3])
    np.testing.assert_array_equal(dims, [3, 3, 3, 3, 3, 3])


def test_union_padding_sizes_follow_the_largest_child():
    """A union of a 3-vector and a 5-vector child pads the 3-vector to 5."""
    conn = np.array([[0, 1], [1, 2]], np.uint32)
    union = ix.Union(counts=[2, 1], members=[[ix.Leaf(3, 0)], [ix.Leaf(5, 1)]])
    prog = ix.Concat([ix.Join(conn, [union])])
    assert prog.sizes() == (10, 4)
    idx, sz = ix.get_indices(prog, 2, [0, 6])
    np.testing.assert_array_equal(idx, [[2, 0, 5, 0], [5, 0, 8, 0]])
    np.testing.assert_array_equal(sz, [[3, 2, 3, 2], [3, 2, 5, 0]])


def test_global_dedupe_orders_by_h_w_row_col():
    coords = [np.array([6, 9, 0, 0, 3, 6, 0, 3], np.uint32), np.array([0, 0, 2, 8], np.uint32)]
    dims = [np.array([3, 3, 3, 3, 3, 3, 3, 3], np.uint16), np.array([3, 3, 2, 2], np.uint16)]
    lay = ix.unique_coordinates(coords, dims)
    # (2,2,2,8) sorts first, then the 3x3 blocks by (row, col); (0,0) appears twice
    np.testing.assert_array_equal(lay.block_positions, [2, 8, 0, 0, 0, 3, 3, 6, 6, 9])
    assert lay.block_dimensions == [2, 2, 3, 3] and lay.block_counts == [1, 4]
    assert lay.blocks_start_indices == [0, 4, 40]
    np.testing.assert_array_equal(lay.lookups[0], [31, 4, 22, 13])
    np.testing.assert_array_equal(lay.lookups[1], [4, 0])


def test_bytecode_encoding_matches_oracle_indices():
    from yasps_amd.program import PLeaf, PJoin, PUnion, PConcat, encode, interpret

    class Conn:
        def __init__(self, a):
            self.a = np.asarray(a, np.uint32)

    class U:
        def __init__(self, c):
            self.children_counts = c
    c1 = Conn([[0, 3, 1, 4], [2, 2, 3, 0]])
    u = U([1, 4])
    prog = PConcat([PJoin(c1, 4, [PUnion(u, [[PLeaf(3, None)], [PLeaf(3, 0)]])]), PLeaf(2, 1)])
    conns, unions = [], []
    code = encode(prog, conns, unions)
    got = interpret(code, [c.a for c in conns], [v.children_counts for v in unions], [0, 12], 2)
    oprog = ix.Concat([ix.Join(c1.a, [ix.Union([1, 4], [[ix.Leaf(3, None)], [ix.Leaf(3, 0)]])]),
                       ix.Leaf(2, 1)])
    want = ix.get_indices(oprog, 2, [0, 12])
    np.testing.assert_array_equal(got[0], want[0])
    np.testing.assert_array_equal(got[1], want[1])


# --------------------------------------------------------------------------- cube and PCG
def test_cube_scene_counts():
    pos, tets = cube_tet_mesh(8)
    assert pos.shape == (729, 3) and tets.shape == (3072, 4)
    p = pos[tets]
    vol = np.linalg.det(np.stack([p[:, 1] - p[:, 0], p[:, 2] - p[:, 0], p[:, 3] - p[:, 0]], 2)) / 6
    assert (vol > 0).all() and np.isclose(vol.sum(), 1.0)
    cs = cube_scene(8)
    assert cs.n_fixed == 81 and cs.n_dof == 1944
    st = newton.fast_structure(cs)
    assert st.layout.num_unique == 4335


def test_fast_structure_matches_general_pipeline():
    cs = cube_scene(3, seed=1)
    _, _, layout = newton.general_structure(cs)
    st = newton.fast_structure(cs)
    np.testing.assert_array_equal(layout.block_positions, st.layout.block_positions)
    assert layout.blocks_start_indices == st.layout.blocks_start_indices


def test_pcg_matches_direct_solve_and_status_codes():
    cs = cube_scene(3, seed=2)
    st = newton.fast_structure(cs)
    asm = newton.evaluate(cs, st)
    A = newton.bsr_matrix(st, asm.values)
    inv = projection.jacobi_block_inverse(asm.diagonal_blocks.reshape(-1, 3, 3))
    M = lambda v: pcg.block_jacobi_apply(inv, v)
    status, x, it, _ = pcg.pcg(lambda v: A @ v, M, asm.gradient, 1e-24)
    assert status == it > 0
    xs = np.linalg.solve(A.toarray(), asm.gradient)
    assert np.linalg.norm(x - xs) / np.linalg.norm(xs) < 1e-10
    assert pcg.pcg(lambda v: A @ v, M, asm.gradient, 1e-24, max_iterations=2)[0] == -4
    assert pcg.pcg(lambda v: A @ v, M, np.full(A.shape[0], np.nan), 1e-8)[0] == -5
    neg = -np.eye(A.shape[0])
    assert pcg.pcg(lambda v: neg @ v, lambda v: v, asm.gradient, 1e-8)[0] == -5   # r.z < 0
    assert pcg.pcg(lambda v: neg @ v, lambda v: -v, asm.gradient, 1e-8)[0] == -5


def test_pcg_flags_nonpositive_curvature():
    A = np.diag([1.0, -3.0, 1.0])     # p^T A p = -1 at iteration 1
    status, *_ = pcg.pcg(lambda v: A @ v, lambda v: v, np.array([1.0, 1.0, 1.0]), 1e-12)
    assert status == -5   # -(iteration 1 + 4)
    A = np.diag([1.0, -1.0, 2.0])     # p^T A p = 2 > 0 first, negative curvature at iteration 2
    status, *_ = pcg.pcg(lambda v: A @ v, lambda v: v, np.array([1.0, 1.0, 1.0]), 1e-12)
    assert status == -6
