"""CPU checks of the affine-body oracle (oracle/abd.py): the chain-rule gradient against
central differences of the total energy in the targets (free positions, A, t), the
Hessian against differences of that gradient, symmetry, and the dof layout."""
import numpy as np
import pytest

from oracle import abd, energies, element_energies
from yasps_amd.scenes import abd_cube_scene


def total_energy(sc):
    cs = sc.cube
    x = sc.positions()
    dinv, vol = energies.snh_rest_data(cs.rest[cs.tets])
    e = energies.snh_energy(x[cs.tets], dinv, vol, cs.mu, cs.lam, cs.dt).sum()
    nf = sc.n_fixed
    e += energies.inertia_energy(x[nf:], cs.last_position[nf:], cs.velocity[nf:], cs.mass[nf:],
                                 cs.dt).sum()
    e += element_energies.value_affine_orthogonality(sc.A.reshape(-1, 9), None,
                                                     np.array([sc.affine_stiffness])).sum()
    return e


def perturbed(sc, k, h):
    """Copy of the scene with target dof k moved by h."""
    import copy
    s2 = copy.deepcopy(sc)
    off_A, off_t, _ = abd.dof_layout(sc)
    if k < off_A:
        v = sc.n_fixed + sc.n_abd + k // 3
        s2.cube.position[v, k % 3] += h
    elif k < off_t:
        b, i = divmod(k - off_A, 9)
        s2.A[b].reshape(-1)[i] += h
    else:
        b, i = divmod(k - off_t, 3)
        s2.t[b, i] += h
    return s2


@pytest.fixture(scope="module")
def scene():
    return abd_cube_scene(n=2, abd_layers=1, bodies=2, seed=4)


def test_layout_and_symmetry(scene):
    assert scene.n_abd == 9 and scene.n_free == 9 and scene.n_fixed == 9
    g, H = abd.assemble(scene, method=2)
    assert g.size == 3 * 9 + 12 * 2
    assert np.abs(H - H.T).max() <= 1e-12 * np.abs(H).max()
    assert np.linalg.eigvalsh(H).min() > 0


def test_chain_rule_gradient_and_hessian(scene):
    g, H = abd.assemble(scene, method=0)
    n = g.size
    fd = np.zeros(n)
    Hfd = np.zeros((n, n))
    for k in range(n):
        h = 1e-6
        fd[k] = (total_energy(perturbed(scene, k, h)) - total_energy(perturbed(scene, k, -h))) / (2 * h)
        gp, _ = abd.assemble(perturbed(scene, k, 1e-6), method=0)
        gm, _ = abd.assemble(perturbed(scene, k, -1e-6), method=0)
        Hfd[:, k] = (gp - gm) / 2e-6
    assert np.abs(fd - g).max() <= 1e-6 * np.abs(g).max()
    assert np.abs(Hfd - H).max() <= 1e-6 * np.abs(H).max()
