"""Edge cases of the solve path against the oracle: the Jacobi PCG status codes
(solverKernel.pyx:600-686, 800-824: 0 converged, -4 iteration budget, -5 invalid initial
residual, -(it+4) breakdown) for a zero right-hand side, an exhausted budget, a
non-finite right-hand side and an indefinite operator (unprojected SNH, dt = 1, large
noise), and the SNH fast kernel's rejection of an element that repeats a vertex."""
import numpy as np
import pytest

from yasps_amd.scenes import cube_scene, build_yasps_scene
from oracle import newton, pcg as opcg, projection

pytestmark = pytest.mark.gpu


def _gpu_solve(s0, rhs, tol, max_it):
    from yasps_amd import device as dev
    from yasps_amd.hessian import vector
    hs = s0.minimizer.computeNumericValue()
    slv = s0.minimizer.linearSolver
    rc = slv.computeSolution(hs, vector(hs.rows, dev.to_device(rhs, "float64")), None,
                             tolerance=tol, maxIterations=max_it, zero_initial_guess=True)
    dev.torch().cuda.synchronize()
    return rc, dev.to_host(slv.solution).copy()


def _oracle(cs, method):
    st = newton.fast_structure(cs)
    asm = newton.evaluate(cs, st, method=method)
    A = newton.bsr_matrix(st, asm.values)
    inv = projection.jacobi_block_inverse(asm.diagonal_blocks.reshape(-1, 3, 3))
    return (lambda v: A @ v), (lambda v: opcg.block_jacobi_apply(inv, v)), asm.gradient


def _expected(status):
    """Python-level return of the reference's Jacobi solve: iteration counts map to 0."""
    return 0 if status >= 0 else status


def test_zero_budget_and_nonfinite_rhs_status():
    cs = cube_scene(3, seed=4)
    s0, _ = build_yasps_scene(cs)
    A, M, g = _oracle(cs, 2)
    print("zero right-hand side", flush=True)
    rc, x = _gpu_solve(s0, np.zeros_like(g), 1e-12, 100)
    status, xo, _, _ = opcg.pcg(A, M, np.zeros_like(g), 1e-12, 100)
    assert rc == _expected(status) == 0 and not x.any()
    print("iteration budget", flush=True)
    rc, x = _gpu_solve(s0, g, 1e-30, 3)
    status, xo, it, _ = opcg.pcg(A, M, g, 1e-30, 3)
    assert status == -4 and rc == -4
    assert np.linalg.norm(x - xo) <= 1e-10 * np.linalg.norm(xo)
    print("non-finite right-hand side", flush=True)
    bad = g.copy()
    bad[5] = np.nan
    rc, _ = _gpu_solve(s0, bad, 1e-12, 100)
    status, _, _, _ = opcg.pcg(A, M, bad, 1e-12, 100)
    assert status == -5 and rc == -5


@pytest.mark.parametrize("seed", [0, 1])
def test_breakdown_on_indefinite_operator_matches_oracle(seed):
    """Unprojected SNH Hessians (projection method 0) with dt = 1 and 30% noise are
    indefinite: p.Ap <= 0 stops the iteration with -(it + 4) at the oracle's iteration."""
    print(f"indefinite operator seed={seed}", flush=True)
    cs = cube_scene(3, seed=seed, dt=1.0, noise=0.3)
    s0, _ = build_yasps_scene(cs, projection_method=0)
    A, M, g = _oracle(cs, 0)
    status, xo, it, _ = opcg.pcg(A, M, g, 1e-16, 200)
    assert status < -4                          # the case does break down
    rc, x = _gpu_solve(s0, g, 1e-16, 200)
    assert rc == status
    assert np.linalg.norm(x - xo) <= 1e-8 * np.linalg.norm(xo)


def test_snh_element_repeating_a_vertex_is_rejected():
    """The SNH fast kernel handles elements whose four vertices are distinct; a tet that
    repeats a vertex (zero volume) raises instead of scattering a degenerate Hessian."""
    print("repeated vertex", flush=True)
    cs = cube_scene(2, seed=0)
    cs.tets = cs.tets.copy()
    e = int(np.flatnonzero((cs.tets[:, 2] >= cs.n_fixed) & (cs.tets[:, 3] >= cs.n_fixed))[0])
    cs.tets[e, 3] = cs.tets[e, 2]             # a moving (target) vertex twice
    s0, _ = build_yasps_scene(cs)
    with pytest.raises(RuntimeError, match="repeated vertex"):
        s0.minimizer.computeNumericValue()


@pytest.mark.parametrize("size", [1, 2, 3, 4, 6, 9, 12])
def test_jacobi_pseudo_inverse_entrywise(size):
    """ysp_invert_diagonal_blocks against oracle.projection.jacobi_block_inverse entry by
    entry, on blocks whose spectra hold negative eigenvalues, eigenvalues with |l| < 1e-6
    (kept as |l|, not inverted: diagonalBlockInverseKernel.pyx:69,91) and ordinary ones."""
    import ctypes as C
    from yasps_amd import _lib, device as dev
    print(f"jacobi pseudo-inverse size {size}", flush=True)
    rng = np.random.default_rng(100 + size)
    spectra = []
    for kind in range(6):
        if kind == 0:
            w = rng.uniform(0.5, 4.0, size)                       # SPD
        elif kind == 1:
            w = rng.uniform(-4.0, -0.5, size)                     # negative definite
        elif kind == 2:
            w = rng.uniform(0.5, 4.0, size)
            w[0] = 3e-7                                           # |l| < 1e-6: kept as l
        elif kind == 3:
            w = rng.uniform(-3.0, 3.0, size)
            w[-1] = -4e-7                                         # kept as |l|
        elif kind == 4:
            w = np.full(size, 2.5e-7)                             # every eigenvalue tiny
        else:
            w = rng.uniform(-3.0, 3.0, size)                      # indefinite
        spectra.append(w)
    blocks = []
    for w in spectra:
        V, _ = np.linalg.qr(rng.standard_normal((size, size)))
        B = (V * w) @ V.T
        blocks.append(0.5 * (B + B.T))
    blocks = np.asarray(blocks)
    ref = projection.jacobi_block_inverse(blocks)
    count = len(blocks)
    src = dev.to_device(blocks.reshape(-1), "float64")
    out = dev.zeros(blocks.size, "float64")
    lib = _lib.load()
    rc = lib.ysp_invert_diagonal_blocks(dev.ptr(src), dev.ptr(out), _lib.u32_array([0]),
                                        _lib.u32_array([count]), _lib.u32_array([size]), 1,
                                        dev.stream())
    assert rc == 0
    got = dev.to_host(out).reshape(count, size, size)
    for b in range(count):
        scale = np.abs(ref[b]).max()
        assert np.abs(got[b] - ref[b]).max() <= 1e-9 * scale, (b, spectra[b])
