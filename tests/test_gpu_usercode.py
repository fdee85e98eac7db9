"""Run-time generated user energies on the GPU. Library energies restated as user
energies (their SymPy definitions compiled at run time into modules of their own) against
the golden vectors of tests/golden/element_energies.npz: per-element energies, the
assembled gradient and the assembled (PSD projected, methods 0 / 1 / 2) Hessian, with
every fourth element's first vertex fixed; a spring energy with no library counterpart
against its closed-form gradient and Hessian; and a Newton solve of a cube with SNH plus
user springs that reaches its residual."""
import numpy as np
import pytest

from oracle import projection

pytestmark = pytest.mark.gpu

REL = 1e-12
SHAPES = {"point_triangle": (4, 0, 2), "edge_edge": (4, 0, 2), "bending": (4, 2, 1),
          "stvk_membrane": (3, 5, 3)}


def _kind(name):
    from yasps_amd import usercode
    kind = usercode.example(f"user_{name}")
    assert (kind.arity, kind.npar, kind.nsc) == SHAPES[name]
    return kind


@pytest.mark.parametrize("method", [0, 1, 2])
@pytest.mark.parametrize("name", list(SHAPES))
def test_user_energy_matches_golden(name, method):
    from yasps_amd import device as dev, scene
    from test_gpu_energies import _golden, _oracle_dense, _rel
    print(f"user energy {name} method {method}", flush=True)
    G = _golden(name)
    arity, npar, nsc = SHAPES[name]
    X = G["X"].reshape(-1, arity, 3)
    n = X.shape[0]
    fixed = np.zeros((n, arity), dtype=bool)
    fixed[::4, 0] = True
    flat, fm = X.reshape(-1, 3), fixed.reshape(-1)
    fixed_ids, moving_ids = np.flatnonzero(fm), np.flatnonzero(~fm)
    uid = np.empty(flat.shape[0], dtype=np.int64)
    uid[fixed_ids] = np.arange(fixed_ids.size)
    uid[moving_ids] = fixed_ids.size + np.arange(moving_ids.size)
    s0 = scene("user")
    m = s0.addMesh("m")
    fv = m.addPrimitive("fixed", numInstances=fixed_ids.size)
    mv = m.addPrimitive("moving", numInstances=moving_ids.size)
    m.addPrimitiveUnion("vertices", [fv, mv])
    fv.addAttribute("position", rows=3, cols=1).updateValue(flat[fixed_ids])
    mv.addAttribute("position", rows=3, cols=1).updateValue(flat[moving_ids])
    m.vertices.addAttribute("position")
    el = m.addPrimitive("elems", numInstances=n)
    e2v = el.addConnectivity("e2v", m.vertices, uid.reshape(n, arity), arity)
    pos = el.addAttribute("positions", through=e2v, source=m.vertices["position"])
    kind = _kind(name)
    e = kind(pos, G["P"] if npar else None, [float(v) for v in G["s"][:nsc]])
    att = el.addAttribute(name, computed_attribute=e)
    s0.addEnergy(att, projection_method=method)
    s0.addMinimizeTarget([mv["position"]])
    hs = s0.minimizer.computeNumericValue()
    grad, D = _oracle_dense(G, method, uid.reshape(n, arity), fixed_ids.size, fixed)
    assert _rel(dev.to_host(att.compute().value), G["E"]) < REL
    assert _rel(dev.to_host(hs.gradient.value), grad) < REL
    assert _rel(hs.dense(), D) < REL


def _spring_scene(n=3, method=0, seed=0, with_snh=False):
    """Cube-mesh edges as springs (rest lengths from a scaled rest shape), bottom fixed;
    optionally SNH and inertia as in the cube scene."""
    from yasps_amd import usercode
    from yasps_amd.scenes import cube_scene, build_yasps_scene
    cs = cube_scene(n, seed=seed)
    if with_snh:
        s0, handles = build_yasps_scene(cs, projection_method=method)
        m = handles["mesh"]
    else:
        from yasps_amd import scene
        s0 = scene("springs")
        m = s0.addMesh("cube")
        fv = m.addPrimitive("fixed_vertices", numInstances=cs.n_fixed)
        mv = m.addPrimitive("moving_vertices", numInstances=cs.n_moving)
        m.addPrimitiveUnion("vertices", [fv, mv])
        fv.addAttribute("position", rows=3, cols=1).updateValue(cs.position[:cs.n_fixed])
        mv.addAttribute("position", rows=3, cols=1).updateValue(cs.position[cs.n_fixed:])
        m.vertices.addAttribute("position")
        s0.addMinimizeTarget([mv["position"]])
    edges = np.unique(np.sort(np.concatenate([cs.tets[:, [a, b]] for a in range(4)
                                              for b in range(a + 1, 4)]), axis=1), axis=0)
    rest_len = 0.9 * np.linalg.norm(cs.rest[edges[:, 1]] - cs.rest[edges[:, 0]], axis=1)
    sp_prim = m.addPrimitive("springs", numInstances=edges.shape[0])
    s2v = sp_prim.addConnectivity("s2v", m.vertices, edges.astype(np.uint32), 2)
    spos = sp_prim.addAttribute("spring_positions", through=s2v, source=m.vertices["position"])
    k = 50.0
    kind = usercode.example("spring")
    att = sp_prim.addAttribute("spring", computed_attribute=kind(spos, rest_len[:, None], [k]))
    s0.addEnergy(att, projection_method=method)
    return cs, s0, att, edges, rest_len, k


@pytest.mark.parametrize("method", [0, 2])
def test_user_spring_matches_closed_form(method):
    """k/2 (l - L)^2 per edge: g = k (l - L) d / l on the second vertex (minus on the
    first) and H = k [(1 - L / l) I + (L / l) d d^T / l^2] on each vertex pair, projected
    with the oracle's method-2 rule when asked."""
    from yasps_amd import device as dev
    print(f"user spring method {method}", flush=True)
    cs, s0, att, edges, L, k = _spring_scene(method=method)
    hs = s0.minimizer.computeNumericValue()
    x = cs.position
    d = x[edges[:, 1]] - x[edges[:, 0]]
    l = np.linalg.norm(d, axis=1)
    E = k / 2 * (l - L) ** 2
    gv = (k * (l - L) / l)[:, None] * d
    Hb = k * ((1 - L / l)[:, None, None] * np.eye(3) + (L / l ** 3)[:, None, None]
              * d[:, :, None] * d[:, None, :])
    He = np.concatenate([np.concatenate([Hb, -Hb], 2), np.concatenate([-Hb, Hb], 2)], 1)
    if method:
        He = projection.project_batch(He, method)
    nf = cs.n_fixed
    nm = x.shape[0] - nf
    grad = np.zeros(3 * nm)
    D = np.zeros((3 * nm, 3 * nm))
    for e, (a, b) in enumerate(edges):
        ge = np.concatenate([-gv[e], gv[e]])
        idx = [(0, a), (1, b)]
        for i, va in idx:
            if va < nf:
                continue
            grad[3 * (va - nf):3 * (va - nf) + 3] += ge[3 * i:3 * i + 3]
            for j, vb in idx:
                if vb < nf:
                    continue
                D[3 * (va - nf):3 * (va - nf) + 3, 3 * (vb - nf):3 * (vb - nf) + 3] += \
                    He[e, 3 * i:3 * i + 3, 3 * j:3 * j + 3]
    rel = lambda a, b: np.abs(a - b).max() / np.abs(b).max()      # noqa: E731
    assert rel(dev.to_host(att.compute().value), E) < REL
    assert rel(dev.to_host(hs.gradient.value), grad) < REL
    assert rel(hs.dense(), D) < REL


def test_newton_solve_with_user_springs():
    """SNH + inertia (library) plus user springs on the n=4 cube: the Newton linear solve
    reaches the requested true residual on the assembled operator."""
    from yasps_amd import device as dev
    from yasps_amd.hessian import vector
    print("newton solve with user springs", flush=True)
    cs, s0, att, *_ = _spring_scene(n=4, method=2, with_snh=True)
    s0.minimizer.setSolver("mas", threshold_mode="relative_residual")
    sol = s0.minimizeEnergy(tolerance=1e-8)
    hs = s0.minimizer.activeHessian
    x = dev.torch().cat([v.reshape(-1) for v in sol])
    r = hs.spmv(vector(hs.rows, x.clone())).value - hs.gradient.value
    rel = float(dev.torch().linalg.norm(r) / dev.torch().linalg.norm(hs.gradient.value))
    assert s0.minimizer.linearSolver.statistics["converged"] and rel < 1.5e-8


PARAMS = {"point_triangle": ["dhat", "kappa"], "edge_edge": ["dhat", "kappa"], "bending": ["k"],
          "stvk_membrane": ["mu", "lam", "dt"]}


@pytest.mark.parametrize("name", list(SHAPES))
def test_user_energy_scene_jacobians_match_oracle(name):
    """d g / d s of the restated library energies through diff2 and the module's mixed
    kernel, with every scene parameter a one-instance constant (a column target), against
    the oracle's scene gradients of the golden elements (as the library kinds' test)."""
    from yasps_amd import device as dev, scene
    from yasps_amd.differentiator import differentiator
    from oracle import jacobian as oj
    from test_gpu_energies import _golden
    from test_gpu_jacobian import _products
    print(f"user energy jacobian {name}", flush=True)
    G = _golden(name)
    arity, npar, nsc = SHAPES[name]
    X = G["X"].reshape(-1, arity, 3)
    n = X.shape[0]
    fixed = np.zeros((n, arity), dtype=bool)
    fixed[::4, 0] = True
    flat, fm = X.reshape(-1, 3), fixed.reshape(-1)
    fids, mids = np.flatnonzero(fm), np.flatnonzero(~fm)
    uid = np.empty(flat.shape[0], dtype=np.int64)
    uid[fids] = np.arange(fids.size)
    uid[mids] = fids.size + np.arange(mids.size)
    s0 = scene("userjac")
    sc = [s0.addConstant(nm).updateValue([float(G["s"][i])]) for i, nm in enumerate(PARAMS[name])]
    m = s0.addMesh("m")
    fv = m.addPrimitive("fixed", numInstances=fids.size)
    mv = m.addPrimitive("moving", numInstances=mids.size)
    m.addPrimitiveUnion("vertices", [fv, mv])
    fv.addAttribute("position", rows=3, cols=1).updateValue(flat[fids])
    pos = mv.addAttribute("position", rows=3, cols=1).updateValue(flat[mids])
    m.vertices.addAttribute("position")
    el = m.addPrimitive("elems", numInstances=n)
    e2v = el.addConnectivity("e2v", m.vertices, uid.reshape(n, arity), arity)
    p = el.addAttribute("positions", through=e2v, source=m.vertices["position"])
    att = el.addAttribute(name, computed_attribute=_kind(name)(p, G["P"] if npar else None, sc))
    s0.addEnergy(att, projection_method=2)
    s0.addMinimizeTarget([pos])
    J = differentiator().diff2([att], [pos], sc).compute()
    Gs = oj.element_scene_gradients(name, G["X"], G.get("P"), np.asarray(G["s"], np.float64))
    rows = np.where(fixed[:, :, None], -1, 3 * (uid.reshape(n, arity)[:, :, None] - fids.size)
                    + np.arange(3)).reshape(n, 3 * arity)
    R = 3 * mids.size
    dense = np.concatenate([oj.dense_jacobian(rows, Gs[:, :, j:j + 1], np.zeros(n, int), R, 1)
                            for j in range(len(sc))], axis=1)
    et, eb = _products(J, dense, np.random.default_rng(11))
    print(f"  J^T a {et:.2e}, J b {eb:.2e}", flush=True)
    assert et < REL and eb < REL


def test_user_spring_jacobian_in_rest_lengths_and_stiffness():
    """A user energy's element parameters as a per-element column target: for the spring
    k/2 (l - L)^2, d g_b / d L = -k d / l on the second vertex (+k d / l on the first), and
    the stiffness as a one-instance target: d g / d k = g / k."""
    from yasps_amd import device as dev, scene, usercode
    from yasps_amd.differentiator import differentiator
    from yasps_amd.scenes import cube_scene
    print("user spring jacobian", flush=True)
    cs = cube_scene(2, seed=3)
    s0 = scene("springjac")
    m = s0.addMesh("cube")
    fv = m.addPrimitive("fixed_vertices", numInstances=cs.n_fixed)
    mv = m.addPrimitive("moving_vertices", numInstances=cs.n_moving)
    m.addPrimitiveUnion("vertices", [fv, mv])
    fv.addAttribute("position", rows=3, cols=1).updateValue(cs.position[:cs.n_fixed])
    pos = mv.addAttribute("position", rows=3, cols=1).updateValue(cs.position[cs.n_fixed:])
    m.vertices.addAttribute("position")
    edges = np.unique(np.sort(np.concatenate([cs.tets[:, [a, b]] for a in range(4)
                                              for b in range(a + 1, 4)]), axis=1), axis=0)
    ne = edges.shape[0]
    L = 0.9 * np.linalg.norm(cs.rest[edges[:, 1]] - cs.rest[edges[:, 0]], axis=1)
    sp_prim = m.addPrimitive("springs", numInstances=ne)
    s2v = sp_prim.addConnectivity("s2v", m.vertices, edges.astype(np.uint32), 2)
    spos = sp_prim.addAttribute("spring_positions", through=s2v, source=m.vertices["position"])
    rest = sp_prim.addAttribute("rest_length", rows=1, cols=1).updateValue(L)
    k = s0.addConstant("k").updateValue([50.0])
    att = sp_prim.addAttribute("spring", computed_attribute=usercode.example("spring")(spos, rest, [k]))
    s0.addEnergy(att, projection_method=2)
    s0.addMinimizeTarget([pos])
    J = differentiator().diff2([att], [pos], [rest, k]).compute()
    x = cs.position
    d = x[edges[:, 1]] - x[edges[:, 0]]
    l = np.linalg.norm(d, axis=1)
    nf = cs.n_fixed
    R = 3 * (x.shape[0] - nf)
    dense = np.zeros((R, ne + 1))
    for e, (a, b) in enumerate(edges):
        g_b = 50.0 * (l[e] - L[e]) / l[e] * d[e]
        for v, sign in ((a, -1.0), (b, 1.0)):
            if v < nf:
                continue
            r = 3 * (v - nf)
            dense[r:r + 3, e] += sign * (-50.0 * d[e] / l[e])
            dense[r:r + 3, ne] += sign * g_b / 50.0
    from test_gpu_jacobian import _products
    et, eb = _products(J, dense, np.random.default_rng(5))
    assert et < REL and eb < REL


def test_user_barriers_on_affine_bodies_match_library():
    """The contact scene (affine and free balls on a cloth) with its point-triangle and
    edge-edge barriers replaced by the user restatements: pairs whose points follow an
    affine body go through the module's general kernel and chain rule, and the assembled
    gradient and projected Hessian equal the library's to 1e-12."""
    from yasps_amd import device as dev, usercode
    from yasps_amd.scenes import contact_scene, build_contact_scene
    print("user barriers on the contact scene", flush=True)
    cs = contact_scene(bodies_x=2, bodies_z=1, cells=4, radius=0.12, cloth_m=9,
                       max_pairs=10 ** 9, seed=0)
    pt, ee = usercode.example("user_point_triangle"), usercode.example("user_edge_edge")
    user = {"pt": lambda pos, dhat, kappa: pt(pos, None, [dhat, kappa]),
            "ee": lambda pos, dhat, kappa: ee(pos, None, [dhat, kappa])}
    out = []
    for barriers in (None, user):
        s0, handles = build_contact_scene(cs, threshold_mode="relative_residual", barriers=barriers)
        hs = s0.minimizer.computeNumericValue()
        out.append((dev.to_host(hs.gradient.value), hs.dense(),
                    dev.to_host(handles["pt"].compute().value)))
    (g0, D0, e0), (g1, D1, e1) = out
    assert cs.n_bodies > 0 and cs.pt.shape[0] > 0
    rel = lambda a, b: np.abs(a - b).max() / np.abs(b).max()      # noqa: E731
    assert rel(e1, e0) < REL and rel(g1, g0) < REL and rel(D1, D0) < REL
