"""Element scatter with repeated vertices inside an element, on the GPU: the scenarios of
the reference's tests/differentiator/test_hessian_sparsity_permutation_gpu.py:115-197
(connectivity rows [3,1,5,0], [1,1,2,0], [0,2,0,2], [5,5,5,5] followed by 67 ring
elements, a quadratic form over a linear map of each element's coordinates, projection
on the inner Hessian), with 3-vector positions (the reference uses 2-vectors; the map's
third column per vertex is this test's own). Expected gradient, Hessian, diagonal blocks
and products are built densely with np.add.at as the reference test does; block-Jacobi
and MAS then solve the system with the 0.5 I regulariser of the reference's method-1
case against numpy.linalg.solve."""
import numpy as np
import pytest

from oracle import projection

pytestmark = pytest.mark.gpu

REL = 3e-11        # the reference test's rtol / atol on these O(1) quantities
NV = 67


def _inner_matrix(size):
    """test_hessian_sparsity_permutation_gpu.py:41-45."""
    rng = np.random.default_rng(209 + size)
    basis, _ = np.linalg.qr(rng.normal(size=(size, size)))
    return (basis * np.linspace(-1.7, 3.1, size)) @ basis.T


def _map(pattern):
    """The reference's 2-vector maps (5 x 8 or 4 x 8, :118-126) with a third column per
    vertex added: J3[:, 3a + i] = J2[:, 2a + i] (i < 2), J3[:, 3a + 2] = 0.37 J2[:, 2a + 1]."""
    if pattern == "connected":
        j2 = np.arange(32, dtype=np.float64).reshape(4, 8) / 41.0 + 0.13
    else:
        j2 = np.array([[0.3, 0, 0, -0.7, 0, 0, 1.1, 0], [0, 0.6, 0, 0, 0.8, 0, 0, -0.4],
                       [0, 0, 1.2, 0, 0, -0.9, 0, 0], [-0.8, 0, 0, 0.4, 0, 0, 0.5, 0],
                       [0, 1.1, 0, 0, -0.2, 0, 0, 0.7]])
        if pattern == "zero_axes":
            j2[4, :] = 0.0
            j2[:, 7] = 0.0
    m = j2.shape[0]
    j3 = np.zeros((m, 12))
    for a in range(4):
        j3[:, 3 * a] = j2[:, 2 * a]
        j3[:, 3 * a + 1] = j2[:, 2 * a + 1]
        j3[:, 3 * a + 2] = 0.37 * j2[:, 2 * a + 1]
    offset = np.zeros(m)
    if pattern == "zero_axes":
        offset[4] = 0.15
    return j3, offset


def _connectivity():
    return np.array([[3, 1, 5, 0], [1, 1, 2, 0], [0, 2, 0, 2], [5, 5, 5, 5]]
                    + [[i, (i + 1) % NV, (i + 9) % NV, (i + 3) % NV] for i in range(NV)],
                    dtype=np.uint32)


def _expected(values, conn, J, Q, lin, off, method):
    n = values.size
    H = np.zeros((n, n))
    g = np.zeros(n)
    Qp = projection.project(Q, method) if method > 0 else Q
    local_h = J.T @ Qp @ J
    for c in conn:
        idx = (3 * c[:, None].astype(np.int64) + np.arange(3)).ravel()
        y = J @ values[c].ravel() + off
        np.add.at(H, (idx[:, None], idx[None, :]), local_h)
        np.add.at(g, idx, J.T @ (Q @ y + lin))
    return g, H


def _scene(values, conn, J, Q, lin, off, method, regulariser):
    from yasps_amd import scene
    from yasps_amd.energies import quadratic_form, inertia
    s0 = scene("sparsity")
    m = s0.addMesh("mesh")
    v = m.addPrimitive("vertices", numInstances=NV)
    pos = v.addAttribute("position", rows=3, cols=1)
    pos.updateValue(values)
    maps = m.addPrimitive("maps", numInstances=len(conn))
    to_v = maps.addConnectivity("to_vertices", v, conn, 4)
    lp = maps.addAttribute("local_position", through=to_v, source=pos)
    e = maps.addAttribute("energy", computed_attribute=quadratic_form(lp, J, Q, lin, off))
    s0.addEnergy(e, [pos], projection_method=method)
    if regulariser:
        # 0.25 |x|^2 (Hessian 0.5 I), as the reference's method-1 case adds
        v.addConstant("zero", rows=3, cols=1).updateValue(np.zeros((NV, 3)))
        v.addConstant("mass", rows=1, cols=1).updateValue(np.full(NV, 0.5))
        ine = v.addAttribute("inertia", computed_attribute=inertia(
            v["zero"], v["zero"], 0.0, pos, v["mass"]))
        s0.addEnergy(ine, [pos], projection_method=-1)
    s0.addMinimizeTarget([pos])
    return s0


def _rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1.0)


@pytest.mark.parametrize("pattern,method", [("interlaced", 0), ("interlaced", 1), ("interlaced", 2),
                                            ("connected", 1), ("zero_axes", 1)])
def test_repeated_vertices_and_sparsity_permutation(pattern, method):
    from yasps_amd import device as dev
    from yasps_amd.hessian import vector
    print(f"sparsity {pattern} method {method}", flush=True)
    values = np.sin(np.arange(3 * NV) * 0.17).reshape(-1, 3)
    conn = _connectivity()
    J, off = _map(pattern)
    Q = _inner_matrix(J.shape[0])
    lin = np.linspace(-0.3, 0.4, J.shape[0])
    g, H = _expected(values, conn, J, Q, lin, off, method)
    reg = method == 1
    if reg:
        H += 0.5 * np.eye(H.shape[0])
        g += 0.5 * values.ravel()
    s0 = _scene(values, conn, J, Q, lin, off, method, reg)
    hs = s0.minimizer.computeNumericValue()
    assert _rel(dev.to_host(hs.gradient.value), g) < REL
    assert _rel(hs.dense(), H) < REL
    diag = np.concatenate([H[3 * i:3 * i + 3, 3 * i:3 * i + 3].ravel() for i in range(NV)])
    assert _rel(dev.to_host(hs.diagonal_blocks), diag) < REL
    assert _rel(dev.to_host(s0.diagonal), np.diagonal(H)) < REL
    direction = np.linspace(-0.8, 0.9, hs.cols)
    for tr in (False, True):
        y = hs.spmv(vector(hs.cols).updateValue(direction), transpose=tr).get()
        assert _rel(y, H @ direction) < REL
    if not reg:
        return
    rhs = np.cos(np.arange(hs.rows) * 0.23)
    xs = np.linalg.solve(H, rhs)
    for name in ("jacobian", "mas"):
        s0.minimizer.setSolver(name)
        slv = s0.minimizer.linearSolver
        rc = slv.computeSolution(hs, vector(hs.rows).updateValue(rhs), None, tolerance=1e-18,
                                 maxIterations=1000, zero_initial_guess=True)
        assert rc == 0, name
        x = dev.to_host(slv.solution)
        np.testing.assert_allclose(x, xs, rtol=2e-8, atol=2e-8)
        assert np.linalg.norm(H @ x - rhs) / np.linalg.norm(rhs) < 2e-8, name


def test_energy_values_of_the_quadratic_form():
    from yasps_amd import device as dev
    values = np.sin(np.arange(3 * NV) * 0.17).reshape(-1, 3)
    conn = _connectivity()
    J, off = _map("interlaced")
    Q = _inner_matrix(J.shape[0])
    lin = np.linspace(-0.3, 0.4, J.shape[0])
    s0 = _scene(values, conn, J, Q, lin, off, 1, False)
    att = s0.meshes["mesh"].maps["energy"]
    e = dev.to_host(att.compute().value)
    y = np.einsum("ij,ej->ei", J, values[conn].reshape(len(conn), 12)) + off
    ref = 0.5 * np.einsum("ei,ij,ej->e", y, Q, y) + y @ lin
    assert _rel(e, ref) < REL
