"""Contact scene and its oracle on the host: the barrier pairs the generator emits are
active in the energies' own squared-distance convention, and the oracle's chain-rule
gradient over [moving positions, A, t] (element terms from the committed fixture) equals
central differences of the total energy written with the numpy restatements
(oracle/element_energies.py value_*, oracle/energies.py), so a wrong vertex Jacobian,
index or missing term fails; the assembled projected Hessian is symmetric positive
definite."""
from pathlib import Path

import numpy as np

from oracle import contact, energies, element_energies as ee

GOLDEN = Path(__file__).with_name("golden") / "contact_small.npz"
SCENE = dict(bodies_x=2, bodies_z=1, cells=4, radius=0.12, cloth_m=9, max_pairs=10 ** 9, seed=0)


def _scene():
    from yasps_amd.scenes import contact_scene
    return contact_scene(**SCENE)


def _terms():
    z = np.load(GOLDEN)
    out = {}
    for kind in ("stvk_membrane", "bending", "point_triangle", "edge_edge"):
        g, packed = z[f"{kind}_g"], z[f"{kind}_H"]
        n = g.shape[1]
        iu = np.triu_indices(n)
        H = np.zeros((g.shape[0], n, n))
        H[:, iu[0], iu[1]] = packed
        H[:, iu[1], iu[0]] = packed
        out[kind] = (g, H)
    return out


def _positions(cs, q):
    """Union positions from the target vector q = [moving (3 each), A (9 each), t (3 each)]."""
    off_A, off_t, _ = contact.dof_layout(cs)
    A = q[off_A:off_t].reshape(-1, 3, 3)
    t = q[off_t:].reshape(-1, 3)
    x = cs.position.copy()
    nf, na = cs.n_fixed, cs.n_abd
    s = slice(nf, nf + na)
    x[s] = t[cs.body_of] + np.einsum("bij,bj->bi", A[cs.body_of], cs.rest[s])
    x[nf + na:] = q[:off_A].reshape(-1, 3)
    return x, A


def _energy(cs, q):
    x, A = _positions(cs, q)
    dinv, vol = energies.snh_rest_data(cs.rest[cs.tets])
    e = energies.snh_energy(x[cs.tets], dinv, vol, cs.mu, cs.lam, cs.dt).sum()
    R = cs.rest[cs.triangles]
    P = ee.triangle_rest(R[:, 0], R[:, 1], R[:, 2], cs.thickness)
    X = x[cs.triangles].reshape(-1, 9)
    s = [cs.cloth_mu, cs.cloth_lam, cs.dt]
    e += ee.value_stvk_membrane(X, P, ee.scene_params("stvk_membrane", X, P, s)[0]).sum()
    Q = cs.rest[cs.hinges]
    P = ee.bending_rest(Q[:, 0], Q[:, 1], Q[:, 2], Q[:, 3])
    e += ee.value_bending(x[cs.hinges].reshape(-1, 12), P, [cs.bend]).sum()
    e += ee.value_point_triangle(x[cs.pt].reshape(-1, 12), None, [cs.dhat, cs.kappa]).sum()
    e += ee.value_edge_edge(x[cs.ee].reshape(-1, 12), None, [cs.dhat, cs.kappa]).sum()
    e += ee.value_affine_orthogonality(A.reshape(-1, 9), None, [cs.affine_stiffness]).sum()
    v = np.arange(cs.n_fixed, x.shape[0])
    e += energies.inertia_energy(x[v], cs.last_position[v], cs.velocity[v], cs.mass[v],
                                 cs.dt).sum()
    return e


def test_generated_pairs_are_active():
    cs = _scene()
    assert cs.pt.shape[0] > 20 and cs.ee.shape[0] > 10
    x = cs.positions()
    for kind, pairs in (("point_triangle", cs.pt), ("edge_edge", cs.ee)):
        X = x[pairs].reshape(-1, 12)
        p0, p1, p2, p3 = (X[:, 3 * k:3 * k + 3] for k in range(4))
        if kind == "point_triangle":
            b = np.cross(p2 - p1, p3 - p1)
            a = p0 - p1
        else:
            b = np.cross(p1 - p0, p3 - p2)
            a = p2 - p0
        d = (a * b).sum(1) ** 2 / (b * b).sum(1)
        assert np.all(d > 0) and np.all(d < cs.dhat)
    # point: a body vertex; triangle / edge: cloth vertices (fixed border or moving cloth)
    body = np.zeros(cs.rest.shape[0], dtype=bool)
    body[np.unique(cs.tets)] = True
    assert body[cs.pt[:, 0]].all() and not body[cs.pt[:, 1:]].any()
    assert body[cs.ee[:, :2]].all() and not body[cs.ee[:, 2:]].any()


def test_oracle_gradient_matches_central_differences():
    cs = _scene()
    terms = _terms()
    g, H = contact.assemble(cs, 2, terms)
    off_A, off_t, ndof = contact.dof_layout(cs)
    q = np.concatenate([cs.position[cs.n_fixed + cs.n_abd:].reshape(-1), cs.A.reshape(-1),
                        cs.t.reshape(-1)])
    rng = np.random.default_rng(3)
    # moving body and cloth entries, every affine matrix and translation entry
    idx = np.concatenate([rng.choice(off_A, 24, replace=False), np.arange(off_A, ndof)])
    h = 1e-7
    for i in idx:
        qp, qm = q.copy(), q.copy()
        qp[i] += h
        qm[i] -= h
        fd = (_energy(cs, qp) - _energy(cs, qm)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-5 * np.abs(g).max() + 1e-6 * abs(g[i]), (i, fd, g[i])
    assert np.abs(H - H.T).max() <= 1e-12 * np.abs(H).max()
    assert np.linalg.eigvalsh(H).min() > 0


def test_oracle_hessian_matches_second_differences_off_contact():
    """Out-of-plane motion of cloth vertices that no body touches: the unprojected
    Hessian's quadratic form equals the energy's second difference (the bending term,
    which has no first-order effect on the flat cloth, dominates it)."""
    cs = _scene()
    g, H = contact.assemble(cs, 0, _terms())
    off_A, _, ndof = contact.dof_layout(cs)
    nm0 = cs.n_fixed + cs.n_abd
    touched = np.zeros(cs.rest.shape[0], dtype=bool)
    touched[cs.pt.ravel()] = True
    touched[cs.ee.ravel()] = True
    touched[np.unique(cs.tets)] = True
    verts = np.flatnonzero(~touched)
    verts = verts[verts >= nm0]
    assert verts.size >= 10
    v = np.zeros(ndof)
    v[3 * (verts - nm0) + 1] = np.random.default_rng(0).standard_normal(verts.size)
    q = np.concatenate([cs.position[nm0:].reshape(-1), cs.A.reshape(-1), cs.t.reshape(-1)])
    h = 1e-4
    sd = (_energy(cs, q + h * v) + _energy(cs, q - h * v) - 2.0 * _energy(cs, q)) / h ** 2
    assert abs(sd - v @ H @ v) <= 1e-5 * abs(v @ H @ v)
