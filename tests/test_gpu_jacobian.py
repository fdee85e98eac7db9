"""Mixed second derivatives on the GPU (yasps_amd.differentiator over csrc/jacobian.hip)
against the oracle (oracle/jacobian.py), on the structure of the reference's material
example (examples/bunny_material_optimization/bunny_material_optimization.py:333-470,
588-634): a tet cube with per-tet Young's modulus and Poisson's ratio (Lamé coefficients
as expressions of them), per-vertex mass, previous positions and the velocity expression
(last - last_last) / dt, SNH with projection 1 and inertia with -1.

Checked, each to 1e-12 of the largest entry (FP64):
* J^T a and J b of d g / d (young, poisson, dt) of SNH and d g / d (mass, last_position,
  last_last_position, dt) of the inertia against the oracle's dense Jacobians;
* sum_t Y_t J_young[:, t] = g_elastic and sum_v m_v J_mass[:, v] = g_inertia
  (homogeneity, against the assembled gradients);
* the Hessian from diff2 with equal targets (sum of two terms) against the scene's own
  Hessian, and the PCG solve with it;
* d g / d (scene parameters) of the generated kinds (IPC barriers, bending, Baraff-Witkin,
  StVK) against the oracle on the golden elements, with fixed vertices;
* targets the build does not differentiate raise NotImplementedError."""
from pathlib import Path

import numpy as np
import pytest

from oracle import energies as oe
from oracle import jacobian as oj

pytestmark = pytest.mark.gpu

REL = 1e-12
GOLDEN = Path(__file__).with_name("golden") / "element_energies.npz"


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _material_scene(n=3, seed=0):
    from yasps_amd import scene, energies as en
    from yasps_amd.scenes import cube_scene
    cs = cube_scene(n, noise=0.05, seed=seed)
    rng = np.random.default_rng(seed + 10)
    nf, V, T = cs.n_fixed, cs.rest.shape[0], cs.tets.shape[0]
    h = 1.0 / n
    s0 = scene("material")
    dt = s0.addConstant("dt").updateValue([cs.dt])
    m = s0.addMesh("m")
    fv = m.addPrimitive("fixed", numInstances=nf)
    mv = m.addPrimitive("moving", numInstances=V - nf)
    m.addPrimitiveUnion("vertices", [fv, mv])
    fv.addAttribute("position", rows=3, cols=1).updateValue(cs.position[:nf])
    pos = mv.addAttribute("position", rows=3, cols=1).updateValue(cs.position[nf:])
    fv.addConstant("rest_position", rows=3, cols=1).updateValue(cs.rest[:nf])
    mv.addConstant("rest_position", rows=3, cols=1).updateValue(cs.rest[nf:])
    m.vertices.addAttribute("position")
    m.vertices.addAttribute("rest_position")
    vals = {
        "last": cs.rest[nf:] + 0.02 * h * rng.standard_normal((V - nf, 3)),
        "last2": cs.rest[nf:] + 0.02 * h * rng.standard_normal((V - nf, 3)),
        "mass": cs.mass[nf:] * rng.uniform(0.5, 1.5, V - nf),
        "young": 1e5 * np.exp(0.3 * rng.standard_normal(T)),
        "poisson": rng.uniform(0.25, 0.45, T),
    }
    last = mv.addConstant("last_position", rows=3, cols=1).updateValue(vals["last"])
    last2 = mv.addConstant("last_last_position", rows=3, cols=1).updateValue(vals["last2"])
    mass = mv.addConstant("mass").updateValue(vals["mass"])
    velocity = (last - last2) / dt
    tets = m.addPrimitive("tets", numInstances=T)
    t2v = tets.addConnectivity("t2v", m.vertices, cs.tets, 4)
    tp = tets.addAttribute("positions", through=t2v, source=m.vertices["position"])
    trp = tets.addAttribute("rest_positions", through=t2v, source=m.vertices["rest_position"])
    young = tets.addConstant("young").updateValue(vals["young"])
    poisson = tets.addConstant("poisson").updateValue(vals["poisson"])
    mu = young / (2.0 * (1.0 + poisson))
    lam = young * poisson / ((1.0 + poisson) * (1.0 - 2.0 * poisson))
    elastic = tets.addAttribute("elastic_energy",
                                computed_attribute=en.stable_neo_hookean(trp, tp, mu, lam, dt))
    inertia = mv.addAttribute("inertia_energy",
                              computed_attribute=en.inertia(last, velocity, dt, pos, mass))
    s0.addEnergy(elastic, projection_method=1)
    s0.addEnergy(inertia, projection_method=-1)
    s0.addMinimizeTarget([pos])
    att = dict(pos=pos, last=last, last2=last2, mass=mass, young=young, poisson=poisson, dt=dt,
               elastic=elastic, inertia=inertia)
    return s0, cs, att, vals


def _oracle(cs, vals):
    """Dense oracle Jacobians (rows: the moving vertices' 3 n_m scalars)."""
    nf, T = cs.n_fixed, cs.tets.shape[0]
    nm = cs.rest.shape[0] - nf
    dt = cs.dt
    rows_t = np.where(cs.tets[:, :, None] >= nf, 3 * (cs.tets[:, :, None] - nf) + np.arange(3), -1)
    rows_t = rows_t.reshape(T, 12)
    dm_inv, vol = oe.snh_rest_data(cs.rest[cs.tets])
    mu, lam = oj.lame(vals["young"], vals["poisson"])
    g_mu, g_lam, g_dt = oj.snh_param_gradients(cs.position[cs.tets], dm_inv, vol, mu, lam, dt)
    dmy, dmn, dly, dln = oj.lame_derivatives(vals["young"], vals["poisson"])
    R = 3 * nm
    J = {}
    J["young"] = oj.dense_jacobian(rows_t, (g_mu * dmy[:, None] + g_lam * dly[:, None])[:, :, None],
                                   np.arange(T), R, T)
    J["poisson"] = oj.dense_jacobian(rows_t, (g_mu * dmn[:, None] + g_lam * dln[:, None])[:, :, None],
                                     np.arange(T), R, T)
    J["dt_snh"] = oj.dense_jacobian(rows_t, g_dt[:, :, None], np.zeros(T, int), R, 1)
    x = cs.position[nf:]
    v = (vals["last"] - vals["last2"]) / dt
    d = oj.inertia_param_gradients(x, vals["last"], v, vals["mass"], dt)
    rows_v = 3 * np.arange(nm)[:, None] + np.arange(3)
    m = vals["mass"]
    J["mass"] = oj.dense_jacobian(rows_v, d["mass"][:, :, None], np.arange(nm), R, nm)
    J["last"] = oj.dense_jacobian(rows_v, d["x_prev"] + d["velocity"] / dt, 3 * np.arange(nm), R, 3 * nm)
    J["last2"] = oj.dense_jacobian(rows_v, -d["velocity"] / dt, 3 * np.arange(nm), R, 3 * nm)
    # d v / d dt = -v / dt
    J["dt_inertia"] = oj.dense_jacobian(rows_v, (d["dt"] + m[:, None] * v)[:, :, None],
                                        np.zeros(nm, int), R, 1)
    # closed forms of the chain through the velocity
    np.testing.assert_allclose(d["x_prev"] + d["velocity"] / dt, -2 * m[:, None, None] * np.eye(3))
    return J


def _products(J, dense, rng):
    from yasps_amd import device as dev
    a = rng.standard_normal(dense.shape[0])
    b = rng.standard_normal(dense.shape[1])
    jt = dev.to_host(J.spmv(dev.to_device(a, "float64"), transpose=True).value)
    jb = dev.to_host(J.spmv(dev.to_device(b, "float64")).value)
    return _rel(jt, dense.T @ a), _rel(jb, dense @ b)


def test_material_jacobians_match_oracle():
    from yasps_amd.differentiator import differentiator
    print("jacobian material scene", flush=True)
    s0, cs, A, vals = _material_scene()
    Jo = _oracle(cs, vals)
    rng = np.random.default_rng(1)
    d = differentiator()
    cases = [("young", [A["elastic"]], A["young"]), ("poisson", [A["elastic"]], A["poisson"]),
             ("dt_snh", [A["elastic"]], A["dt"]), ("mass", [A["inertia"]], A["mass"]),
             ("last", [A["inertia"]], A["last"]), ("last2", [A["inertia"]], A["last2"]),
             ("dt_inertia", [A["inertia"]], A["dt"])]
    for name, src, theta in cases:
        J = d.diff2(src, [A["pos"]], [theta]).compute()
        assert (J.rows, J.cols) == Jo[name].shape, name
        et, eb = _products(J, Jo[name], rng)
        print(f"  {name}: J^T a {et:.2e}, J b {eb:.2e}", flush=True)
        assert et < REL and eb < REL, name
    # both energies w.r.t. [young, dt]: columns concatenated, terms summed
    J = d.diff2([A["elastic"], A["inertia"]], [A["pos"]], [A["young"], A["dt"]]).compute()
    dense = np.concatenate([Jo["young"], Jo["dt_snh"] + Jo["dt_inertia"]], axis=1)
    et, eb = _products(J, dense, rng)
    assert et < REL and eb < REL
    assert J.block_dimensions == [[3, 1]] and J.block_counts[0] > 0


def test_homogeneity_recovers_the_assembled_gradients():
    from yasps_amd import device as dev
    from yasps_amd.differentiator import differentiator
    print("jacobian homogeneity", flush=True)
    s0, cs, A, vals = _material_scene(seed=2)
    d = differentiator()
    for src, theta, key in (([A["elastic"]], A["young"], "young"), ([A["inertia"]], A["mass"], "mass")):
        H = d.diff2(src, [A["pos"]], [A["pos"]], projection_method=1).compute()
        g = dev.to_host(H.gradient.value)
        J = d.diff2(src, [A["pos"]], [theta]).compute()
        got = dev.to_host(J.spmv(dev.to_device(vals[key], "float64")).value)
        assert _rel(got, g) < REL, key


def test_hessian_terms_sum_and_solve_like_the_scene():
    from yasps_amd import device as dev
    from yasps_amd.differentiator import differentiator
    from yasps_amd.hessian import vector
    from yasps_amd.solver import solver
    print("jacobian hessian sum", flush=True)
    s0, cs, A, vals = _material_scene(seed=3)
    d = differentiator()
    H = d.diff2([A["elastic"]], [A["pos"]], [A["pos"]], projection_method=1) + \
        d.diff2([A["inertia"]], [A["pos"]], [A["pos"]], projection_method=-1)
    H.compute()
    hs = s0.minimizer.computeNumericValue()
    rng = np.random.default_rng(4)
    v = rng.standard_normal(hs.rows)
    a = dev.to_host(H.spmv(vector(hs.rows, dev.to_device(v, "float64"))).value)
    b = dev.to_host(hs.spmv(vector(hs.rows, dev.to_device(v, "float64"))).value)
    assert _rel(a, b) < REL
    assert _rel(dev.to_host(H.gradient.value), dev.to_host(hs.gradient.value)) < REL
    sol = solver("jacobian", threshold_mode="relative_residual")
    rc = sol.computeSolution(H, H.gradient, None, tolerance=1e-10, maxIterations=5000,
                             zero_initial_guess=True)
    assert rc >= 0
    x = dev.to_host(sol.solution)
    r = dev.to_host(hs.spmv(vector(hs.rows, dev.to_device(x, "float64"))).value) - \
        dev.to_host(hs.gradient.value)
    assert np.linalg.norm(r) <= 1e-9 * np.linalg.norm(dev.to_host(hs.gradient.value))


def test_unsupported_targets_raise():
    from yasps_amd.differentiator import differentiator, secondOrderJacobian
    s0, cs, A, vals = _material_scene(n=2, seed=5)
    d = differentiator()
    with pytest.raises(NotImplementedError):
        d.diff2([A["elastic"]], [A["last"]], [A["pos"]]).compute()     # a DoF as column
    with pytest.raises(ValueError):
        secondOrderJacobian([A["pos"]], [])
    with pytest.raises(ValueError):
        d.diff2([A["elastic"]], [A["pos"]], [A["young"]]) + d.diff2([A["elastic"]], [A["pos"]], [A["poisson"]])
    # a column the energy does not depend on gives a zero product
    from yasps_amd import device as dev
    J = d.diff2([A["inertia"]], [A["pos"]], [A["young"]]).compute()
    out = dev.to_host(J.spmv(dev.to_device(np.ones(J.rows), "float64"), transpose=True).value)
    assert out.shape == (J.cols,) and not out.any()


# ------------------------------------------------------------------ generated kinds
ARITY = {"point_point": 2, "point_edge": 3, "point_triangle": 4, "edge_edge": 4, "bending": 4,
         "baraff_witkin": 3, "stvk_membrane": 3}
PARAMS = {"point_point": ["dhat", "kappa"], "point_edge": ["dhat", "kappa"],
          "point_triangle": ["dhat", "kappa"], "edge_edge": ["dhat", "kappa"],
          "bending": ["k"], "baraff_witkin": ["ks", "kh", "dt"], "stvk_membrane": ["mu", "lam", "dt"]}


@pytest.mark.parametrize("kind", list(ARITY))
def test_generated_kind_scene_jacobians_match_oracle(kind):
    from yasps_amd import scene, energies as en, device as dev
    from yasps_amd.differentiator import differentiator
    print(f"jacobian kind {kind}", flush=True)
    z = np.load(GOLDEN)
    G = {k: z[f"{kind}_{k}"] for k in "XPsR" if f"{kind}_{k}" in z}
    Ar = ARITY[kind]
    X = G["X"].reshape(-1, Ar, 3)
    n = X.shape[0]
    fixed = np.zeros((n, Ar), dtype=bool)
    fixed[::4, 0] = True
    fm = fixed.reshape(-1)
    flat = X.reshape(-1, 3)
    fids, mids = np.flatnonzero(fm), np.flatnonzero(~fm)
    uid = np.empty(flat.shape[0], dtype=np.int64)
    uid[fids] = np.arange(fids.size)
    uid[mids] = fids.size + np.arange(mids.size)
    s0 = scene("k")
    svals = [float(v) for v in G["s"]]
    names = PARAMS[kind]
    sc = {nm: s0.addConstant(nm).updateValue([svals[i]]) for i, nm in enumerate(names)}
    m = s0.addMesh("m")
    fv = m.addPrimitive("fixed", numInstances=fids.size)
    mv = m.addPrimitive("moving", numInstances=mids.size)
    m.addPrimitiveUnion("vertices", [fv, mv])
    fv.addAttribute("position", rows=3, cols=1).updateValue(flat[fids])
    pos = mv.addAttribute("position", rows=3, cols=1).updateValue(flat[mids])
    m.vertices.addAttribute("position")
    el = m.addPrimitive("elems", numInstances=n)
    e2v = el.addConnectivity("e2v", m.vertices, uid.reshape(n, Ar), Ar)
    p = el.addAttribute("positions", through=e2v, source=m.vertices["position"])
    if kind in ("bending", "baraff_witkin", "stvk_membrane"):
        rest = G["R"].reshape(-1, 3)
        fv.addConstant("rest_position", rows=3, cols=1).updateValue(rest[fids])
        mv.addConstant("rest_position", rows=3, cols=1).updateValue(rest[mids])
        m.vertices.addAttribute("rest_position")
        rp = el.addAttribute("rest_positions", through=e2v, source=m.vertices["rest_position"])
        if kind == "bending":
            e = en.bending(p, rp, sc["k"])
        elif kind == "baraff_witkin":
            e = en.baraff_witkin(rp, p, sc["ks"], sc["kh"], 0.003, sc["dt"])
        else:
            e = en.stvk_membrane(rp, p, sc["mu"], sc["lam"], 0.003, sc["dt"])
    else:
        e = getattr(en, kind)(p, sc["dhat"], sc["kappa"])
    att = el.addAttribute(kind, computed_attribute=e)
    s0.addEnergy(att, projection_method=2)
    s0.addMinimizeTarget([pos])
    thetas = [sc[nm] for nm in names]
    J = differentiator().diff2([att], [pos], thetas, dynamic_instances=True).compute()
    # oracle: d g / d s of the golden elements, scattered over the moving rows
    s_full = np.asarray(G["s"], dtype=np.float64)
    Gs = oj.element_scene_gradients(kind, G["X"], G.get("P"), s_full)       # [n, N, S]
    rows = np.where(fixed[:, :, None], -1, 3 * (uid.reshape(n, Ar)[:, :, None] - fids.size)
                    + np.arange(3)).reshape(n, 3 * Ar)
    R = 3 * mids.size
    dense = np.concatenate([oj.dense_jacobian(rows, Gs[:, :, j:j + 1], np.zeros(n, int), R, 1)
                            for j in range(len(names))], axis=1)
    rng = np.random.default_rng(11)
    et, eb = _products(J, dense, rng)
    print(f"  J^T a {et:.2e}, J b {eb:.2e}", flush=True)
    assert et < REL and eb < REL
