"""CPU checks that pin the oracle of the generated energies (oracle/element_energies.py)
and its golden vectors: the numpy restatement against the SymPy one, finite differences
of the energy for gradient and Hessian, translation and rotation invariance, and closed
forms (distances of the barrier geometries, dihedral angles, rest states)."""
from pathlib import Path

import numpy as np
import pytest

from oracle import element_energies as ee

GOLDEN = Path(__file__).with_name("golden") / "element_energies.npz"


@pytest.fixture(scope="module")
def golden():
    if not GOLDEN.exists():
        pytest.skip("golden vectors missing (python tests/golden/make_energy_golden.py)")
    z = np.load(GOLDEN)
    return {k: z[k] for k in z.files}


def _g(golden, kind):
    return {k: golden[f"{kind}_{k}"] for k in "XPsEgHR" if f"{kind}_{k}" in golden}


@pytest.mark.parametrize("kind", ee.KINDS)
def test_numpy_and_sympy_restatements_agree(golden, kind):
    G = _g(golden, kind)
    E = ee.VALUES[kind](G["X"], G["P"], G["s"])
    assert np.abs(E - G["E"]).max() <= 1e-12 * np.abs(G["E"]).max()


def _energy_at(kind, G, X):
    # scene parameters (Baraff-Witkin branch) frozen at the golden point, as the kernel does
    S = ee.scene_params(kind, G["X"], G["P"], G["s"])
    if kind == "baraff_witkin":
        out = []
        for i in range(X.shape[0]):
            F = ee.membrane_f(X[i:i + 1], G["P"][i:i + 1])[0]
            i5u, i5v = np.linalg.norm(F[:, 0]), np.linalg.norm(F[:, 1])
            i6 = F[:, 0] @ F[:, 1]
            st = S[i, 3] * (i5u - 1) ** 2 + S[i, 4] * (i5v - 1) ** 2
            out.append((S[i, 0] * st + S[i, 1] * i6 * i6) * G["P"][i, 4] * S[i, 2] ** 2)
        return np.array(out)
    return ee.VALUES[kind](X, G["P"], G["s"])


def _fd_hessian(kind, G, X, h):
    n = X.shape[1]
    out = np.zeros((X.shape[0], n, n))
    for i in range(n):
        for j in range(i, n):
            ei = np.zeros_like(X)
            ej = np.zeros_like(X)
            ei[:, i] = h
            ej[:, j] = h
            v = (_energy_at(kind, G, X + ei + ej) - _energy_at(kind, G, X + ei - ej)
                 - _energy_at(kind, G, X - ei + ej) + _energy_at(kind, G, X - ei - ej)) / (4 * h * h)
            out[:, i, j] = out[:, j, i] = v
    return out


@pytest.mark.parametrize("kind", ee.KINDS)
def test_gradient_and_hessian_against_finite_differences(golden, kind):
    """Central differences of the numpy energy, step scaled to each element's extent;
    the Hessian uses one Richardson step (error O(h^4)); checked element by element."""
    G = _g(golden, kind)
    X = G["X"]
    n = X.shape[1]
    ext = (X.max(axis=1) - X.min(axis=1))[:, None]
    h = 1e-6 * ext
    fd_g = np.zeros_like(X)
    for i in range(n):
        e = np.zeros_like(X)
        e[:, i] = h[:, 0]
        fd_g[:, i] = (_energy_at(kind, G, X + e) - _energy_at(kind, G, X - e)) / (2 * h[:, 0])
    gs = np.abs(G["g"]).max(axis=1)
    assert (np.abs(fd_g - G["g"]).max(axis=1) <= 1e-5 * gs).all()
    # the barriers curve on the scale of the (small) distances, not of the element
    h2 = (1e-5 if kind in ("point_point", "point_edge", "point_triangle", "edge_edge")
          else 1e-3) * ext[:, 0]
    H1 = _fd_hessian(kind, G, X, h2)
    H2 = _fd_hessian(kind, G, X, h2 / 2)
    fd_H = (4 * H2 - H1) / 3
    hs = np.abs(G["H"]).max(axis=(1, 2))
    err = np.abs(fd_H - G["H"]).max(axis=(1, 2)) / hs
    assert (err <= 1e-5).all(), err.max()
    assert np.abs(G["H"] - np.transpose(G["H"], (0, 2, 1))).max() <= 1e-12 * np.abs(G["H"]).max()


@pytest.mark.parametrize("kind", [k for k in ee.KINDS if k != "affine_orthogonality"])
def test_translation_invariance(golden, kind):
    G = _g(golden, kind)
    arity = ee.SHAPES[kind][0]
    g = G["g"].reshape(-1, arity, 3).sum(axis=1)
    assert np.abs(g).max() <= 1e-9 * np.abs(G["g"]).max()
    Hs = G["H"].reshape(-1, arity, 3, arity, 3).sum(axis=3)
    assert np.abs(Hs).max() <= 1e-9 * np.abs(G["H"]).max()


@pytest.mark.parametrize("kind", ee.KINDS)
def test_rotation_invariance(golden, kind):
    G = _g(golden, kind)
    R = np.linalg.qr(np.random.default_rng(5).standard_normal((3, 3)))[0]
    if kind == "affine_orthogonality":
        X = np.einsum("ij,bjk->bik", R, G["X"].reshape(-1, 3, 3)).reshape(G["X"].shape)
    else:
        X = (G["X"].reshape(-1, 3) @ R.T).reshape(G["X"].shape)
    E = ee.VALUES[kind](X, G["P"], G["s"])
    assert np.abs(E - G["E"]).max() <= 1e-10 * np.abs(G["E"]).max()


def test_barrier_closed_forms():
    dhat, kappa = 0.01, 3.0
    s = np.array([dhat, kappa])

    def b(d):
        return kappa * (d - dhat) ** 2 * np.log(d / dhat) ** 2
    x = np.array([[0.0, 0, 0, np.sqrt(2 * dhat), 0, 0]])
    assert np.isclose(ee.value_point_point(x, None, s)[0], kappa * dhat ** 2 * np.log(2) ** 2)
    h = 0.07
    pe = np.array([[0.3, h, 0.0, -1.0, 0.0, 0.0, 2.0, 0.0, 0.0]])
    assert np.isclose(ee.value_point_edge(pe, None, s)[0], b(h * h))
    pt = np.array([[0.2, 0.1, h, 0.0, 0, 0, 1.0, 0, 0, 0.0, 1.0, 0]])
    assert np.isclose(ee.value_point_triangle(pt, None, s)[0], b(h * h))
    ed = np.array([[0, 0, 0, 1.0, 0, 0, 0.5, -1.0, h, 0.5, 1.0, h]])
    assert np.isclose(ee.value_edge_edge(ed, None, s)[0], b(h * h))


def test_dihedral_angle_closed_form():
    for phi in (0.0, 0.3, 1.2, 2.5):
        q0, q1 = np.array([[0.0, 0, 0]]), np.array([[1.0, 0, 0]])
        q2 = np.array([[0.5, 1.0, 0.0]])
        q3 = np.array([[0.5, -np.cos(phi), np.sin(phi)]])
        assert np.isclose(abs(ee.edge_theta(q0, q1, q2, q3)[0]), phi, atol=1e-12)


def test_membranes_vanish_at_rest_and_under_rotation():
    rng = np.random.default_rng(3)
    v = [rng.standard_normal((6, 3)) for _ in range(3)]
    P = ee.triangle_rest(*v, 0.01)
    X = np.concatenate(v, axis=1)
    R = np.linalg.qr(rng.standard_normal((3, 3)))[0]
    XR = (X.reshape(-1, 3) @ R.T).reshape(X.shape) + 0.7
    for kind, s in (("baraff_witkin", np.array([1e3, 1e2, 0.01, 0, 0])),
                    ("stvk_membrane", np.array([10.0, 20.0, 0.01]))):
        assert np.abs(ee.VALUES[kind](X, P, s)).max() < 1e-20
        assert np.abs(ee.VALUES[kind](XR, P, s)).max() < 1e-20
    # area parameter = thickness * triangle area
    area = 0.5 * np.linalg.norm(np.cross(v[1] - v[0], v[2] - v[0]), axis=1)
    assert np.allclose(P[:, 4], 0.01 * area)


def test_affine_orthogonality_closed_forms():
    s = np.array([2.0])
    R = np.linalg.qr(np.random.default_rng(1).standard_normal((3, 3)))[0]
    assert ee.value_affine_orthogonality(R.reshape(1, 9), None, s)[0] < 1e-28
    A = 2.0 * np.eye(3)
    assert np.isclose(ee.value_affine_orthogonality(A.reshape(1, 9), None, s)[0], 2.0 * 0.5 * 27)
