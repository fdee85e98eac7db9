"""Material adjoint: dL/dtheta from the adjoint solve and the per-tet dg/dtheta kernel
against central finite differences of the loss (equilibrium re-solved, h = 1e-6), the
torch.autograd.Function backward equal to it, and the mixed-precision adjoint close to
the FP64 one; equilibrium and adjoint solved with MAS (the default) and block-Jacobi."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=["mas", "jacobian"])
def setup(request):
    from yasps_amd.adjoint import MaterialProblem
    print(f"adjoint setup, {request.param}", flush=True)
    prob = MaterialProblem(n=3, solver=request.param)
    rng = np.random.default_rng(0)
    theta = np.log(1e5) + 0.3 * rng.standard_normal(prob.num_tets)
    theta_t = np.log(1e5) + 0.3 * np.random.default_rng(1).standard_normal(prob.num_tets)
    xt = prob.equilibrium(theta_t)
    target = xt.reshape(-1, 3)[prob.top].copy()
    x = prob.equilibrium(theta, tol=1e-13)
    return prob, theta, target, x


def test_adjoint_matches_finite_differences(setup):
    prob, theta, target, x = setup
    print("adjoint vs FD n=3", flush=True)
    g = prob.vjp(x, prob.loss_grad_x(x, target))
    idx = np.random.default_rng(3).choice(prob.num_tets, 8, replace=False)
    h = 1e-6
    for i in idx:
        tp, tm = theta.copy(), theta.copy()
        tp[i] += h
        tm[i] -= h
        lp = prob.loss(prob.equilibrium(tp, x0=x, tol=1e-13), target)
        lm = prob.loss(prob.equilibrium(tm, x0=x, tol=1e-13), target)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-4 * np.abs(g).max(), (i, fd, g[i])


def test_material_gradients_through_mixed_jacobians(setup):
    """dL/dE and dL/dnu per tet from the generalised mixed Jacobian d g / d (E, nu)
    (diff2 through the SNH coefficients' expressions): E dL/dE equals the log-E adjoint of
    the dedicated kernel (two independent evaluations of the same derivative), and dL/dnu
    matches central differences of the loss (equilibrium re-solved per perturbed tet)."""
    prob, theta, target, x = setup
    print("material gradients via diff2", flush=True)
    prob.set_theta(theta)               # earlier tests leave a perturbed theta set
    xbar = prob.loss_grad_x(x, target)
    g_theta = prob.vjp(x, xbar)
    gE, gnu = prob.vjp_material(x, xbar)
    E = np.exp(theta)
    assert np.abs(E * gE - g_theta).max() <= 1e-10 * np.abs(g_theta).max()
    nu0 = np.full(prob.num_tets, prob.poisson)
    h = 1e-6
    try:
        for i in np.random.default_rng(5).choice(prob.num_tets, 6, replace=False):
            npl, nmi = nu0.copy(), nu0.copy()
            npl[i] += h
            nmi[i] -= h
            lp = prob.loss(prob.equilibrium_material(E, npl, x0=x, tol=1e-13), target)
            lm = prob.loss(prob.equilibrium_material(E, nmi, x0=x, tol=1e-13), target)
            fd = (lp - lm) / (2 * h)
            assert abs(fd - gnu[i]) <= 1e-4 * np.abs(gnu).max(), (i, fd, gnu[i])
    finally:
        prob.set_theta(theta)


def test_autograd_function_backward_is_the_adjoint(setup):
    import torch
    from yasps_amd.adjoint import equilibrium_function
    prob, theta, target, x = setup
    EF = equilibrium_function()
    th = torch.tensor(theta, dtype=torch.float64, requires_grad=True)
    xs = EF.apply(th, prob)
    tt = torch.tensor(target, dtype=torch.float64)
    loss = ((xs.view(-1, 3)[torch.as_tensor(prob.top)] - tt) ** 2).sum()
    loss.backward()
    ref = prob.vjp(prob.equilibrium(theta), prob.loss_grad_x(prob.equilibrium(theta), target))
    assert np.allclose(th.grad.numpy(), ref, rtol=1e-9, atol=1e-12 * np.abs(ref).max())


def test_mixed_precision_adjoint_close_to_fp64(setup):
    from yasps_amd.adjoint import MaterialProblem
    prob, theta, target, x = setup
    g64 = prob.vjp(x, prob.loss_grad_x(x, target))
    pm = MaterialProblem(n=3, precision="mixed", solver=prob.scene.minimizer.solverName)
    xm = pm.equilibrium(theta)
    assert np.linalg.norm(xm - x) <= 1e-9 * np.linalg.norm(x)
    g32 = pm.vjp(xm, pm.loss_grad_x(xm, target))
    assert np.linalg.norm(g32 - g64) <= 1e-6 * np.linalg.norm(g64)
