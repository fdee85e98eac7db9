"""HIP path vs the CPU oracle on the tet-cube Newton step (SNH + inertia, fixed/moving
union): index buffers and block layout bit-exact, FP64 gradient / Hessian / diagonal
blocks within 1e-12 of the largest entry, SpMV against the dense product, and the PCG
reaching the requested residual."""
import numpy as np
import pytest

from yasps_amd.scenes import cube_scene, build_yasps_scene
from oracle import newton, energies, projection, assembly, pcg as opcg

pytestmark = pytest.mark.gpu

REL = 1e-12  # FP64 parity bar, relative to the largest entry of the compared array


def _rel_err(a, b):
    scale = max(np.abs(b).max(), 1e-300)
    return np.abs(np.asarray(a) - np.asarray(b)).max() / scale


def _gpu(n, seed=0, noise=0.02, method=2):
    from yasps_amd import device as dev
    cs = cube_scene(n, seed=seed, noise=noise)
    s0, h = build_yasps_scene(cs, projection_method=method)
    hs = s0.minimizer.computeNumericValue()
    dev.torch().cuda.synchronize()
    return cs, s0, hs


def _oracle_general(cs, method=2):
    t_tet, t_in, layout = newton.general_structure(cs)
    dinv, vol = energies.snh_rest_data(cs.rest[cs.tets])
    g_t, H_t = energies.snh_gradient_hessian(cs.position[cs.tets], dinv, vol, cs.mu, cs.lam, cs.dt)
    H_t = projection.project_batch(H_t, method)
    gs, ds = assembly.diagonal_layout([3], [cs.n_moving])
    grad = np.zeros(cs.n_dof)
    vals = np.zeros(layout.total_size)
    diag = np.zeros(ds[-1])
    assembly.scatter_term(g_t, H_t, t_tet, layout.lookups[0], grad, vals, diag, gs, ds)
    nf = cs.n_fixed
    gi, Hi = energies.inertia_gradient_hessian(cs.position[nf:], cs.last_position[nf:],
                                               cs.velocity[nf:], cs.mass[nf:], cs.dt)
    assembly.scatter_term(gi, Hi, t_in, layout.lookups[1], grad, vals, diag, gs, ds)
    return t_tet, t_in, layout, grad, vals, diag


@pytest.mark.parametrize("n", [2, 3])
def test_index_pipeline_bit_exact(n):
    from yasps_amd import device as dev
    cs, s0, hs = _gpu(n)
    t_tet, t_in, layout, *_ = _oracle_general(cs)
    for ti, ot in zip(hs.static_index, (t_tet, t_in)):
        K = ti.K
        assert K == ot.K
        np.testing.assert_array_equal(dev.to_host(ti.indices, "uint32").reshape(-1, K), ot.indices)
        np.testing.assert_array_equal(dev.to_host(ti.sizes, "uint16").reshape(-1, K), ot.sizes)
        np.testing.assert_array_equal(dev.to_host(ti.permutations).reshape(-1, K),
                                      ot.compression.permutations)
        np.testing.assert_array_equal(dev.to_host(ti.coord_outer, "uint32"),
                                      ot.compression.coordinates_outer)
        # sizes are sorted in place by the grouping, as in the reference
        np.testing.assert_array_equal(dev.to_host(ti.total_sizes, "uint16"),
                                      np.sort(ot.compression.gradient_sizes, kind="stable"))
        np.testing.assert_array_equal(dev.to_host(ti.grouped, "uint32"), ot.compression.grouped_inner)
        nu = int(dev.to_host(ti.num_unique, "uint16")[0])
        np.testing.assert_array_equal(dev.to_host(ti.unique_sizes, "uint16")[:nu],
                                      ot.compression.unique_sizes)
        np.testing.assert_array_equal(dev.to_host(ti.grouped_outer, "uint32")[:nu + 1],
                                      ot.compression.grouped_outer)
        np.testing.assert_array_equal(dev.to_host(ti.coords, "uint32")[:len(ot.coords)], ot.coords)
        np.testing.assert_array_equal(dev.to_host(ti.dims, "uint16")[:len(ot.dims)], ot.dims)
    np.testing.assert_array_equal(dev.to_host(hs.static.positions, "uint32"), layout.block_positions)
    assert hs.static.starts == layout.blocks_start_indices
    assert hs.static.counts == layout.block_counts
    assert hs.static.dims == layout.block_dimensions
    for ti, lk in zip(hs.static_index, layout.lookups):
        np.testing.assert_array_equal(dev.to_host(ti.lookups, "uint32"), lk)


@pytest.mark.parametrize("method", [1, 2])
@pytest.mark.parametrize("n,seed", [(2, 0), (3, 1), (3, 5)])
def test_gradient_hessian_parity(n, seed, method):
    """Projection method 1 (|lambda|, the reference's addEnergy default) and 2 (max)."""
    from yasps_amd import device as dev
    print(f"cube parity n={n} seed={seed} method={method}", flush=True)
    cs, s0, hs = _gpu(n, seed, method=method)
    *_, grad, vals, diag = _oracle_general(cs, method)
    assert _rel_err(dev.to_host(hs.gradient.value), grad) < REL
    assert _rel_err(dev.to_host(hs.blocks_flattened), vals) < REL
    assert _rel_err(dev.to_host(hs.diagonal_blocks), diag) < REL
    # scene.diagonal (scene.pyx:186-188): the scalar diagonal of the assembled Hessian
    d = dev.to_host(s0.diagonal)
    assert _rel_err(d, np.diagonal(diag.reshape(-1, 3, 3), axis1=1, axis2=2).reshape(-1)) < REL


@pytest.mark.parametrize("method", [1, 2])
def test_large_noise_exercises_projection(method):
    """With 10x the noise most tets are indefinite: the projected blocks still match, for
    |lambda| (method 1) and max(lambda, 0) (method 2), which differ on every such tet."""
    from yasps_amd import device as dev
    print(f"cube large noise method={method}", flush=True)
    cs, s0, hs = _gpu(3, seed=3, noise=0.2, method=method)
    dinv, vol = energies.snh_rest_data(cs.rest[cs.tets])
    _, H_t = energies.snh_gradient_hessian(cs.position[cs.tets], dinv, vol, cs.mu, cs.lam, cs.dt)
    assert (~projection.is_psd_batch(H_t)).sum() > len(H_t) // 4
    *_, grad, vals, diag = _oracle_general(cs, method)
    *_, vals_other, _ = _oracle_general(cs, 3 - method)
    assert _rel_err(vals_other, vals) > 1e-3          # the two rules do differ here
    assert _rel_err(dev.to_host(hs.blocks_flattened), vals) < REL
    assert _rel_err(dev.to_host(hs.gradient.value), grad) < REL
    assert _rel_err(dev.to_host(hs.diagonal_blocks), diag) < REL


def test_cube_n8_fast_structure_parity():
    """The n=8 cube (3072 tets, 1944 unknowns, 4335 blocks)."""
    from yasps_amd import device as dev
    cs, s0, hs = _gpu(8)
    st = newton.fast_structure(cs)
    asm = newton.evaluate(cs, st)
    assert hs.static.num_blocks == 4335
    np.testing.assert_array_equal(dev.to_host(hs.static.positions, "uint32"), st.layout.block_positions)
    assert _rel_err(dev.to_host(hs.blocks_flattened), asm.values) < REL
    assert _rel_err(dev.to_host(hs.gradient.value), asm.gradient) < REL


def test_spmv_matches_dense():
    from yasps_amd import device as dev
    from yasps_amd.hessian import vector
    cs, s0, hs = _gpu(3)
    D = hs.dense()
    x = np.linspace(-0.8, 0.9, hs.cols)
    v = vector(hs.cols).updateValue(x)
    for tr in (False, True):
        y = hs.spmv(v, transpose=tr).get()
        assert _rel_err(y, D @ x) < REL


def test_energy_values_match_oracle():
    from yasps_amd import device as dev
    cs = cube_scene(3, seed=2)
    s0, h = build_yasps_scene(cs)
    e_gpu = dev.to_host(h["snh"].compute().value)
    dinv, vol = energies.snh_rest_data(cs.rest[cs.tets])
    e_ref = energies.snh_energy(cs.position[cs.tets], dinv, vol, cs.mu, cs.lam, cs.dt)
    assert _rel_err(e_gpu, e_ref) < REL
    nf = cs.n_fixed
    ei = dev.to_host(h["inertia"].compute().value)
    ei_ref = energies.inertia_energy(cs.position[nf:], cs.last_position[nf:], cs.velocity[nf:],
                                     cs.mass[nf:], cs.dt)
    assert _rel_err(ei, ei_ref) < REL


@pytest.mark.parametrize("n", [3, 8])
def test_jacobi_pcg_reference_rule(n):
    """Reference stopping rule (squared preconditioned ratio) and iteration count."""
    from yasps_amd import device as dev
    cs = cube_scene(n, seed=0)
    s0, h = build_yasps_scene(cs)
    tol = 1e-16
    sol = s0.minimizeEnergy(tolerance=tol)
    x = dev.to_host(sol[0])
    stats = s0.minimizer.linearSolver.statistics
    assert stats["converged"]
    st = newton.fast_structure(cs)
    asm = newton.evaluate(cs, st)
    A = newton.bsr_matrix(st, asm.values)
    inv = projection.jacobi_block_inverse(asm.diagonal_blocks.reshape(-1, 3, 3))
    status, xo, it, _ = opcg.pcg(lambda v: A @ v, lambda v: opcg.block_jacobi_apply(inv, v),
                                 asm.gradient, tol)
    assert status >= 0
    assert abs(stats["iterations"] - it) <= 1
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) < 1e-7
    rel = opcg.relative_residual(lambda v: A @ v, x, asm.gradient)
    assert rel < 1e-7


def test_relative_residual_mode():
    from yasps_amd import device as dev
    cs = cube_scene(8, seed=0)
    s0, h = build_yasps_scene(cs, threshold_mode="relative_residual")
    sol = s0.minimizeEnergy(tolerance=1e-8)
    x = dev.to_host(sol[0])
    st = newton.fast_structure(cs)
    asm = newton.evaluate(cs, st)
    A = newton.bsr_matrix(st, asm.values)
    rel = opcg.relative_residual(lambda v: A @ v, x, asm.gradient)
    assert rel < 1.5e-8
    xs = np.linalg.solve(A.toarray(), asm.gradient)
    assert np.linalg.norm(x - xs) / np.linalg.norm(xs) < 1e-6


def test_full_bsr_matches_upper_storage():
    """The mirrored full block-CSR applies the same operator as the upper storage."""
    from yasps_amd import device as dev
    from yasps_amd.hessian import vector
    cs, s0, hs = _gpu(4, seed=2)
    assert hs.full is not None
    D = hs.dense()
    x = np.cos(np.arange(hs.cols) * 0.37)
    v = vector(hs.cols).updateValue(x)
    y_full = hs.spmv(v).get()
    y_upper = hs.spmv(v, transpose=True).get()     # generic upper-storage kernel
    assert _rel_err(y_full, D @ x) < REL
    assert _rel_err(y_upper, D @ x) < REL
    rp = dev.to_host(hs.full.row_ptr, "uint32")
    assert rp[-1] == hs.full.nnz == 2 * hs.static.num_blocks - hs.rows // 3


def test_generic_pcg_path_matches_fast_path():
    """ysp_pcg_solve (upper storage, any block size) and ysp_pcg3_solve agree."""
    from yasps_amd import device as dev
    cs = cube_scene(6, seed=4)
    s0, h = build_yasps_scene(cs)
    x_fast = dev.to_host(s0.minimizeEnergy(tolerance=1e-16)[0]).copy()
    it_fast = s0.minimizer.linearSolver.statistics["iterations"]
    hs = s0.minimizer.activeHessian
    hs.full = None
    slv = s0.minimizer.linearSolver
    rc = slv.computeSolution(hs, hs.gradient, None, tolerance=1e-16, zero_initial_guess=True)
    assert rc == 0
    x_gen = dev.to_host(slv.solution)
    assert abs(slv.statistics["iterations"] - it_fast) <= 1
    assert np.linalg.norm(x_gen - x_fast) / np.linalg.norm(x_fast) < 1e-9


def test_staged_and_atomic_assembly_agree_and_staged_is_deterministic():
    from yasps_amd import device as dev
    cs = cube_scene(5, seed=7, noise=0.1)
    out = {}
    for mode in (True, False):
        s0, h = build_yasps_scene(cs)
        hs = s0.minimizer._active()
        hs.use_staging = mode
        hs.compute()
        assert hs.staged == mode
        out[mode] = (dev.to_host(hs.gradient.value).copy(), dev.to_host(hs.blocks_flattened).copy(),
                     dev.to_host(hs.diagonal_blocks).copy())
        if mode:
            hs.compute()
            again = (dev.to_host(hs.gradient.value), dev.to_host(hs.blocks_flattened),
                     dev.to_host(hs.diagonal_blocks))
            for a, b in zip(out[mode], again):
                np.testing.assert_array_equal(a, b)   # bitwise reproducible
    for a, b in zip(out[True], out[False]):
        assert _rel_err(a, b) < REL
    st = newton.fast_structure(cs)
    asm = newton.evaluate(cs, st)
    assert _rel_err(out[True][1], asm.values) < REL
    assert _rel_err(out[True][0], asm.gradient) < REL
