"""The solve loop reading the upper block storage through each entry's source word
(``YSP_VALUES_UPPER``) against the mirrored component-major copy (``YSP_VALUES_F64``).
The products read the same doubles and sum them with the same fused multiply-add chain,
so the SpMV and the refinement residual are equal bit for bit. The solves take the same
iteration counts and both reach the requested true residual; their dot products add
workgroup partials into shared slots with FP64 atomics, whose order (and so the last bits)
varies between any two solves, and the stiff contact scene amplifies those bits (its
solutions agree to about 1e-8). Covered: 3x3 layouts with a dynamic term and with a hub
row longer than 128 entries (32-bit source words, the dynamic-part flag, the long-row
kernel), the split layout of an affine-body scene and the contact scene (64-bit source
words, 9x9 / 9x3 / 3x9 sub-blocks, dynamic barrier pairs), with block-Jacobi and MAS."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _cube_with_hub(solver):
    from yasps_amd.energies import point_point
    from yasps_amd.scenes import cube_scene, build_yasps_scene
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
    return s0


def _abd(solver):
    from yasps_amd.scenes import abd_cube_scene, build_abd_scene
    opts = {"max_levels": 3, "coarsest_level_weight": 1.0} if solver == "mas" else {}
    sc = abd_cube_scene(n=3, abd_layers=2, bodies=2, seed=7)
    s0, _ = build_abd_scene(sc, projection_method=2, solver_name=solver,
                            threshold_mode="relative_residual", **opts)
    return s0


def _contact(solver):
    from yasps_amd.scenes import contact_scene, build_contact_scene
    cs = contact_scene(bodies_x=2, bodies_z=1, cells=4, radius=0.12, cloth_m=9,
                       max_pairs=10 ** 9, seed=0)
    opts = {"max_levels": 3, "max_domain_dofs": 32} if solver == "mas" else {}
    s0, _ = build_contact_scene(cs, solver_name=solver, threshold_mode="relative_residual", **opts)
    return s0


SCENES = {"cube_hub": _cube_with_hub, "abd": _abd, "contact": _contact}


TOL = 1e-9


def _solve(s0, source):
    """Solution, iterations, convergence and the true relative residual ||g - H x|| / ||g||
    on the FP64 operator (mirrored product)."""
    from yasps_amd import device as dev
    from yasps_amd.hessian import vector
    hs = s0.minimizer.activeHessian
    if hs is not None:
        hs.value_source = source
    s0.minimizer.computeNumericValue()
    hs = s0.minimizer.activeHessian
    hs.value_source = source
    sol = s0.minimizeEnergy(tolerance=TOL)
    x = dev.torch().cat([v.reshape(-1) for v in sol])
    st = s0.minimizer.linearSolver.statistics
    g = hs.gradient.value
    r = hs.spmv(vector(hs.rows, x.clone())).value - g
    rel = float(dev.torch().linalg.norm(r) / dev.torch().linalg.norm(g))
    return dev.to_host(x), int(st["iterations"]), bool(st["converged"]), rel


@pytest.mark.parametrize("solver", ["jacobian", "mas"])
@pytest.mark.parametrize("scene", list(SCENES))
def test_upper_values_solve_equals_mirrored(scene, solver):
    print(f"value source {scene} {solver}", flush=True)
    s0 = SCENES[scene](solver)
    xm, im, cm, rm = _solve(s0, "mirrored")
    xu, iu, cu, ru = _solve(s0, "upper")
    assert cm and cu
    assert im == iu, (im, iu)
    assert rm <= 1.05 * TOL and ru <= 1.05 * TOL, (rm, ru)
    assert np.linalg.norm(xu - xm) <= 1e-6 * np.linalg.norm(xm)


def test_upper_values_spmv_and_residual_equal_mirrored():
    from yasps_amd import _lib, device as dev
    print("value source spmv / refine", flush=True)
    for build in (_cube_with_hub, _abd, _contact):
        s0 = build("jacobian")
        hs = s0.minimizer.computeNumericValue()
        f = hs.operator
        n = hs.rows
        lib = _lib.load()
        x = dev.to_device(np.random.default_rng(1).standard_normal(n), "float64")
        b = dev.to_device(np.random.default_rng(2).standard_normal(n), "float64")
        out = {}
        for source in ("mirrored", "upper"):
            hs.value_source = source
            vals, fmt = hs.solve_values()
            y = dev.zeros(n, "float64")
            _lib.check(lib.ysp_bsr3_spmv(dev.ptr(f.row_ptr), dev.ptr(f.col), vals, fmt, f.nnz, n,
                                         dev.ptr(x), dev.ptr(y), dev.stream()), "ysp_bsr3_spmv")
            r = dev.zeros(n, "float64")
            xs = x.clone()
            rr, bb = C.c_double(0.0), C.c_double(0.0)
            _lib.check(lib.ysp_bsr3_refine(dev.ptr(f.row_ptr), dev.ptr(f.col), vals, fmt, f.nnz, n,
                                           dev.ptr(b), dev.ptr(xs), None, dev.ptr(r), C.byref(rr),
                                           C.byref(bb), dev.stream()), "ysp_bsr3_refine")
            out[source] = (dev.to_host(y), dev.to_host(r), rr.value)
        np.testing.assert_array_equal(out["upper"][0], out["mirrored"][0])
        np.testing.assert_array_equal(out["upper"][1], out["mirrored"][1])
        assert abs(out["upper"][2] - out["mirrored"][2]) <= 1e-12 * out["mirrored"][2]
        assert fmt == _lib.VALUES_UPPER
