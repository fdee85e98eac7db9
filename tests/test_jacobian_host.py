"""Mixed second derivatives on the CPU: the oracle (oracle/jacobian.py) pinned by central
differences of the gradients w.r.t. each parameter, by homogeneity and by the chain rule
through the reference example's Lamé and velocity expressions
(examples/bunny_material_optimization/bunny_material_optimization.py:340,355-359); and
the elementwise expressions of yasps_amd.expr (values and derivatives) against the same
closed forms, on CPU tensors."""
import numpy as np
import pytest

from oracle import element_energies as ee
from oracle import energies as oe
from oracle import jacobian as oj


def _tets(T, seed):
    rng = np.random.default_rng(seed)
    ref = np.array([[0.0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]])
    rest = ref[None] + 0.1 * rng.standard_normal((T, 4, 3))
    x = rest + 0.15 * rng.standard_normal((T, 4, 3))
    return rest, x, rng


def test_snh_parameter_gradients_match_central_differences():
    rest, x, rng = _tets(24, 1)
    dm_inv, vol = oe.snh_rest_data(rest)
    mu = rng.uniform(1e3, 1e5, 24)
    lam = rng.uniform(1e3, 1e5, 24)
    dt = 0.01
    g_mu, g_lam, g_dt = oj.snh_param_gradients(x, dm_inv, vol, mu, lam, dt)

    def grad(mu_, lam_, dt_):
        return oe.snh_gradient_hessian(x, dm_inv, vol, mu_, lam_, dt_)[0]

    for got, fd in (
            (g_mu, lambda h: (grad(mu + h * mu, lam, dt) - grad(mu - h * mu, lam, dt)) / (2 * h * mu[:, None])),
            (g_lam, lambda h: (grad(mu, lam + h * lam, dt) - grad(mu, lam - h * lam, dt)) / (2 * h * lam[:, None])),
            (g_dt, lambda h: (grad(mu, lam, dt * (1 + h)) - grad(mu, lam, dt * (1 - h))) / (2 * h * dt))):
        ref = fd(1e-5)
        np.testing.assert_allclose(got, ref, rtol=1e-6, atol=1e-8 * np.abs(ref).max())


def test_snh_parameter_gradients_are_homogeneous():
    """E is homogeneous of degree 1 in (mu, lam) (alpha depends on mu / lam only) and of
    degree 2 in dt: mu dg/dmu + lam dg/dlam = g and dt dg/ddt = 2 g."""
    rest, x, rng = _tets(16, 2)
    dm_inv, vol = oe.snh_rest_data(rest)
    mu, lam, dt = rng.uniform(1e3, 1e5, 16), rng.uniform(1e3, 1e5, 16), 0.02
    g = oe.snh_gradient_hessian(x, dm_inv, vol, mu, lam, dt)[0]
    g_mu, g_lam, g_dt = oj.snh_param_gradients(x, dm_inv, vol, mu, lam, dt)
    np.testing.assert_allclose(mu[:, None] * g_mu + lam[:, None] * g_lam, g, rtol=1e-10,
                               atol=1e-12 * np.abs(g).max())
    np.testing.assert_allclose(dt * g_dt, 2 * g, rtol=1e-12, atol=1e-14 * np.abs(g).max())


def test_lame_derivatives_match_central_differences():
    rng = np.random.default_rng(3)
    y = rng.uniform(1e4, 1e6, 50)
    nu = rng.uniform(0.05, 0.45, 50)
    d = oj.lame_derivatives(y, nu)
    h = 1e-6
    mu_y = (oj.lame(y * (1 + h), nu)[0] - oj.lame(y * (1 - h), nu)[0]) / (2 * h * y)
    mu_n = (oj.lame(y, nu + h)[0] - oj.lame(y, nu - h)[0]) / (2 * h)
    la_y = (oj.lame(y * (1 + h), nu)[1] - oj.lame(y * (1 - h), nu)[1]) / (2 * h * y)
    la_n = (oj.lame(y, nu + h)[1] - oj.lame(y, nu - h)[1]) / (2 * h)
    for got, ref in zip(d, (mu_y, mu_n, la_y, la_n)):
        np.testing.assert_allclose(got, ref, rtol=1e-7)
    # mu, lam are linear in Y: Y d/dY = identity
    mu, lam = oj.lame(y, nu)
    np.testing.assert_allclose(y * d[0], mu, rtol=1e-14)
    np.testing.assert_allclose(y * d[2], lam, rtol=1e-14)


def test_inertia_parameter_gradients_match_central_differences():
    rng = np.random.default_rng(4)
    V = 7
    x, xp, v = (rng.standard_normal((V, 3)) for _ in range(3))
    m = rng.uniform(0.1, 2.0, V)
    dt = 0.01
    d = oj.inertia_param_gradients(x, xp, v, m, dt)

    def grad(x_, xp_, v_, m_, dt_):
        return oe.inertia_gradient_hessian(x_, xp_, v_, m_, dt_)[0]

    h = 1e-6
    np.testing.assert_allclose(d["mass"], (grad(x, xp, v, m + h, dt) - grad(x, xp, v, m - h, dt)) / (2 * h),
                               rtol=1e-8, atol=1e-10)
    np.testing.assert_allclose(d["dt"], (grad(x, xp, v, m, dt + h * dt) - grad(x, xp, v, m, dt - h * dt))
                               / (2 * h * dt), rtol=1e-6, atol=1e-9)
    for key, which in (("x_prev", 1), ("velocity", 2)):
        for k in range(3):
            e = np.zeros(3)
            e[k] = h
            args_p = [x, xp, v]
            args_m = [x, xp, v]
            args_p[which] = args_p[which] + e
            args_m[which] = args_m[which] - e
            fd = (grad(*args_p, m, dt) - grad(*args_m, m, dt)) / (2 * h)
            np.testing.assert_allclose(d[key][:, :, k], fd, rtol=1e-7, atol=1e-10)
    # g is linear in m
    np.testing.assert_allclose(m[:, None] * d["mass"], grad(x, xp, v, m, dt), rtol=1e-14)


def test_inertia_chain_through_the_example_velocity():
    """v = (x_prev - x_prev2) / dt: d g / d x_prev = -2 m I, d g / d x_prev2 = m I."""
    rng = np.random.default_rng(5)
    V = 5
    x, xp, xp2 = (rng.standard_normal((V, 3)) for _ in range(3))
    m = rng.uniform(0.1, 2.0, V)
    dt = 0.01

    def grad(xp_, xp2_):
        return oe.inertia_gradient_hessian(x, xp_, (xp_ - xp2_) / dt, m, dt)[0]

    h = 1e-6
    for k in range(3):
        e = np.zeros(3)
        e[k] = h
        np.testing.assert_allclose((grad(xp + e, xp2) - grad(xp - e, xp2)) / (2 * h),
                                   -2 * m[:, None] * np.eye(3)[k], rtol=1e-7, atol=1e-9)
        np.testing.assert_allclose((grad(xp, xp2 + e) - grad(xp, xp2 - e)) / (2 * h),
                                   m[:, None] * np.eye(3)[k], rtol=1e-7, atol=1e-9)


def _kind_sample(kind, rng, n=6):
    arity, width, npar, nsc = ee.SHAPES[kind]
    if kind == "affine_orthogonality":
        X = (np.eye(3) + 0.2 * rng.standard_normal((n, 3, 3))).reshape(n, 9)
        return X, None, np.array([3.0])
    base = {2: [[0, 0, 0], [0.03, 0.01, 0.0]],
            3: [[0, 0, 0.02], [0, 0, 0], [1, 0, 0]],
            4: [[0.3, 0.3, 0.02], [0, 0, 0], [1, 0, 0], [0, 1, 0]]}[arity]
    X = np.array(base, dtype=np.float64).reshape(1, -1) + 0.002 * rng.standard_normal((n, arity * 3))
    if kind in ("point_point", "point_edge", "point_triangle", "edge_edge"):
        if kind == "edge_edge":
            X = np.array([[0, 0, 0], [1, 0.1, 0], [0.2, -0.5, 0.02], [0.3, 0.5, 0.03]]).reshape(1, -1) \
                + 0.002 * rng.standard_normal((n, 12))
        return X, None, np.array([1e-3, 2.0])
    if kind == "bending":
        X = np.array([[0, 0, 0], [1, 0, 0], [0.5, 0.8, 0.1], [0.5, -0.8, 0.2]]).reshape(1, -1) \
            + 0.01 * rng.standard_normal((n, 12))
        q = X.reshape(n, 4, 3)
        P = ee.bending_rest(q[:, 0], q[:, 1], q[:, 2], q[:, 3]) + [0.1, 0.0]
        return X, P, np.array([5.0])
    # membranes: rest and current shapes perturbed independently (shear and stretch)
    tri = np.array([[0, 0, 0], [1, 0, 0], [0, 0, 1]]).reshape(1, -1)
    R = tri + 0.05 * rng.standard_normal((n, 9))
    X = tri + 0.15 * rng.standard_normal((n, 9))
    P = ee.triangle_rest(R[:, 0:3], R[:, 3:6], R[:, 6:9], 0.01)
    if kind == "baraff_witkin":
        return X, P, np.array([100.0, 20.0, 0.01, 0.0, 0.0])
    return X, P, np.array([50.0, 30.0, 0.01])


@pytest.mark.parametrize("kind", ee.KINDS)
def test_generated_scene_gradients_match_central_differences(kind):
    rng = np.random.default_rng(6)
    X, P, s = _kind_sample(kind, rng)
    G = oj.element_scene_gradients(kind, X, P, s)
    _, nsc = ee.SHAPES[kind][2], ee.SHAPES[kind][3]
    P_ = np.zeros((X.shape[0], 1)) if P is None else P
    for j in range(nsc):
        if kind == "baraff_witkin" and j >= 3:
            continue                       # branch coefficients chosen by the element
        h = 1e-6 * max(abs(s[j]), 1e-3)
        sp, sm = s.copy(), s.copy()
        sp[j] += h
        sm[j] -= h
        gp = oj.element_gradients(kind, X, P_, sp)
        gm = oj.element_gradients(kind, X, P_, sm)
        fd = (gp - gm) / (2 * h)
        np.testing.assert_allclose(G[:, :, j], fd, rtol=2e-6, atol=1e-7 * max(np.abs(fd).max(), 1e-300))


LINEAR = {"point_point": [1], "point_edge": [1], "point_triangle": [1], "edge_edge": [1],
          "bending": [0], "baraff_witkin": [0, 1], "stvk_membrane": [0, 1],
          "affine_orthogonality": [0]}


@pytest.mark.parametrize("kind", ee.KINDS)
def test_generated_scene_gradients_are_homogeneous_in_stiffness(kind):
    """g is linear in the stiffness parameters (kappa, bending stiffness, (ks, kh), (mu,
    lam), k): their d g / d s weighted by the parameters add up to g."""
    rng = np.random.default_rng(7)
    X, P, s = _kind_sample(kind, rng)
    G = oj.element_scene_gradients(kind, X, P, s)
    P_ = np.zeros((X.shape[0], 1)) if P is None else P
    g = oj.element_gradients(kind, X, P_, s)
    acc = sum(s[j] * G[:, :, j] for j in LINEAR[kind])
    np.testing.assert_allclose(acc, g, rtol=1e-9, atol=1e-12 * np.abs(g).max())


def test_dense_jacobian_sums_shared_blocks():
    rows = np.array([[0, 1, 2, -1, -1, -1], [0, 1, 2, 3, 4, 5]])
    B = np.arange(2 * 6 * 2, dtype=float).reshape(2, 6, 2)
    J = oj.dense_jacobian(rows, B, [0, 0], 6, 2)
    ref = np.zeros((6, 2))
    ref[:3] += B[0, :3]
    ref += B[1]
    np.testing.assert_array_equal(J, ref)


# ------------------------------------------------------------------ expressions
def _attrs():
    from yasps_amd.core import scene
    sc = scene("jac")
    m = sc.addMesh("m")
    tets = m.addPrimitive("tets", numInstances=6)
    verts = m.addPrimitive("verts", numInstances=6)
    return sc, tets, verts


def test_lame_expressions_values_and_derivatives():
    import torch
    sc, tets, _ = _attrs()
    young = tets.addConstant("young")
    poisson = tets.addConstant("poisson")
    rng = np.random.default_rng(8)
    vals = {young: rng.uniform(1e4, 1e6, 6), poisson: rng.uniform(0.1, 0.45, 6)}

    def value_of(a):
        return torch.tensor(vals[a], dtype=torch.float64)

    mu = young / (2.0 * (1.0 + poisson))
    lam = young * poisson / ((1.0 + poisson) * (1.0 - 2.0 * poisson))
    rmu, rlam = oj.lame(vals[young], vals[poisson])
    np.testing.assert_allclose(mu.evaluate(6, value_of, "cpu").numpy()[:, 0], rmu, rtol=1e-15)
    np.testing.assert_allclose(lam.evaluate(6, value_of, "cpu").numpy()[:, 0], rlam, rtol=1e-15)
    d = oj.lame_derivatives(vals[young], vals[poisson])
    for e, th, ref in ((mu, young, d[0]), (mu, poisson, d[1]), (lam, young, d[2]), (lam, poisson, d[3])):
        got = e.partial(th, 6, value_of, "cpu").numpy()
        assert got.shape == (6, 1, 1)
        np.testing.assert_allclose(got[:, 0, 0], ref, rtol=1e-13)


def test_velocity_expression_derivatives():
    import torch
    sc, _, verts = _attrs()
    last = verts.addConstant("last", rows=3)
    last2 = verts.addConstant("last2", rows=3)
    dt = sc.addConstant("dt")
    rng = np.random.default_rng(9)
    vals = {last: rng.standard_normal(18), last2: rng.standard_normal(18), dt: np.array([0.02])}

    def value_of(a):
        return torch.tensor(vals[a], dtype=torch.float64)

    v = (last - last2) / dt
    assert v.rows == 3 and v.instances() == 6 and v.owner() is verts
    ref = (vals[last] - vals[last2]).reshape(6, 3) / 0.02
    np.testing.assert_allclose(v.evaluate(6, value_of, "cpu").numpy(), ref, rtol=1e-15)
    np.testing.assert_allclose(v.partial(last, 6, value_of, "cpu").numpy(),
                               np.broadcast_to(np.eye(3) / 0.02, (6, 3, 3)), rtol=1e-15)
    np.testing.assert_allclose(v.partial(last2, 6, value_of, "cpu").numpy(),
                               np.broadcast_to(-np.eye(3) / 0.02, (6, 3, 3)), rtol=1e-15)
    np.testing.assert_allclose(v.partial(dt, 6, value_of, "cpu").numpy()[:, :, 0], -ref / 0.02,
                               rtol=1e-14)
    assert v.partial(sc.addConstant("other"), 6, value_of, "cpu") is None


def test_affine_map_syntax_still_builds_matrix_vector():
    from yasps_amd.affine import MatrixVector
    from yasps_amd.expr import Expr
    sc, tets, verts = _attrs()
    A = verts.addConstant("A", rows=3, cols=3)
    r = verts.addConstant("r", rows=3)
    assert isinstance(A * r, MatrixVector)
    assert isinstance(r * 2.0, Expr) and isinstance(2.0 * r, Expr)
    with pytest.raises(ValueError):
        _ = r + tets.addConstant("s")      # 3-vector + scalar
