"""The drop-in Python surface on the host (no kernel launches): addEnergy's reference
argument order and errors (scene/scene.pyx:155-168, minimizer/minimizer.pyx:185-189),
addMinimizeTarget's checks (scene.pyx:190-198), the reference's default solver (MAS,
solver/solver.py:22, minimizer.pyx:101,133) and the supported target sizes."""
import numpy as np
import pytest


def _scene():
    from yasps_amd import scene, inertia
    s0 = scene("api")
    m = s0.addMesh("m")
    v = m.addPrimitive("vertices", numInstances=4)
    x = v.addAttribute("position", rows=3, cols=1)
    x.updateValue(np.zeros((4, 3)))
    v.addConstant("last", rows=3, cols=1)
    v.addConstant("vel", rows=3, cols=1)
    e = v.addAttribute("inertia", computed_attribute=inertia(v["last"], v["vel"], 0.01, x, 1.0))
    return s0, v, x, e


def test_add_energy_reference_argument_order():
    s0, v, x, e = _scene()
    # (e, targets, projection_method, save_intermediate, gradient_only, dynamic_instances)
    s0.addEnergy(e, [x], 1, True, False, True)
    req = s0.minimizer._minimizer__requests[-1]
    assert req[2] == 1 and req[3] is True          # the sixth argument is dynamic_instances
    s0.addEnergy(e, [x], projection_method=2, save_intermediate=True)
    assert s0.minimizer._minimizer__requests[-1][3] is False


def test_add_energy_errors():
    s0, v, x, e = _scene()
    with pytest.raises(NotImplementedError, match="gradient_only"):
        s0.addEnergy(e, [x], gradient_only=True)
    with pytest.raises(TypeError, match="grouped_add"):
        s0.addEnergy(e, [x], grouped_add=1)
    with pytest.raises(TypeError, match="lto"):
        s0.addEnergy(e, [x], lto="yes")
    with pytest.raises(ValueError, match="projection_method"):
        s0.addEnergy(e, [x], projection_method=3)


def test_add_minimize_target_checks():
    s0, v, x, e = _scene()
    y = v.addAttribute("other", rows=3, cols=1)
    s0.addEnergy(e, [x])
    with pytest.raises(ValueError, match="missing"):
        s0.addMinimizeTarget([y])
    with pytest.raises(ValueError, match="duplicate"):
        s0.addMinimizeTarget([x, x])
    s0.addMinimizeTarget([x])
    assert s0.minimizer.wrt == [x]


def test_one_scalar_targets_are_rejected():
    s0, v, x, e = _scene()
    w = v.addAttribute("weight", rows=1, cols=1)
    with pytest.raises(NotImplementedError, match="3-vectors"):
        s0.addMinimizeTarget([w])


def test_mas_is_the_default_solver():
    from yasps_amd import solver
    s0, *_ = _scene()
    assert s0.minimizer.solverName == "mas"
    assert solver().solverName == "mas"
    s0.minimizer.setSolver()
    assert s0.minimizer.solverName == "mas"
    s0.minimizer.setSolver("jacobian")
    assert s0.minimizer.solverName == "jacobian"


def test_quadratic_form_uses_the_symmetric_part_of_q():
    """The inner matrix may carry rounding asymmetry (B diag(d) B^T); the energy's Hessian
    is its symmetric part. A matrix that is not symmetric to rounding is refused."""
    import numpy as np
    from yasps_amd.energies import quadratic_form, QUADRATIC_FORM_ROWS
    rng = np.random.default_rng(214)
    B, _ = np.linalg.qr(rng.normal(size=(5, 5)))
    Q = (B * np.linspace(-1.7, 3.1, 5)) @ B.T
    J = rng.normal(size=(5, 12))
    e = quadratic_form(None, J, Q)
    R = QUADRATIC_FORM_ROWS
    Qp = np.asarray(e.scene_params[12 * R:12 * R + R * R]).reshape(R, R)
    assert np.array_equal(Qp, Qp.T)
    np.testing.assert_allclose(Qp[:5, :5], 0.5 * (Q + Q.T), rtol=0, atol=0)
    with pytest.raises(ValueError, match="symmetric"):
        quadratic_form(None, J, Q + 1e-6 * np.triu(np.ones((5, 5)), 1))
