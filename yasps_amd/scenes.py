"""Synthetic tet-cube scenes used by the solve benchmarks and the parity tests.

The cube scene follows the Newton step of ``examples/one_bunny/one_bunny.py:40-160``:
vertices are split into a fixed and a moving primitive that are stacked in one
union (``one_bunny.py:58-60``), tets join the union vertices, the Stable
Neo-Hookean energy uses the re-parameterised Lame constants of
``one_bunny.py:12-17`` and the inertia term uses a per-vertex lumped mass.

Only numpy is used here: the scene is plain input data for both the HIP path and
the CPU oracle.
"""
from __future__ import annotations

from dataclasses import dataclass, field
import numpy as np


def lame_parameters(young: float, poisson: float):
    """Return (mu, lam) of the Stable Neo-Hookean re-parameterisation.

    ``examples/one_bunny/one_bunny.py:14-17``: mu = 4/3 mu_L and
    lam = lam_L + 5/6 mu_L.
    """
    mu_l = young / (2.0 * (1.0 + poisson))
    lam_l = young * poisson / ((1.0 + poisson) * (1.0 - 2.0 * poisson))
    return 4.0 * mu_l / 3.0, lam_l + 5.0 * mu_l / 6.0


# The six vertex paths from corner (0,0,0) to (1,1,1) of a unit hex, one per
# permutation of the axes (Kuhn / Freudenthal split along the main diagonal).
_AXIS_PERMUTATIONS = ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0))


def cube_tet_mesh(n: int, size: float = 1.0):
    """Split [0,size]^3 into n^3 hexes and each hex into 6 positively oriented tets.

    Returns (positions [V,3] float64, tets [T,4] int64) with lexicographic vertex
    ids ``i + (n+1)*(j + (n+1)*k)`` for grid coordinates (i, j, k).
    """
    if n < 1:
        raise ValueError("cube_tet_mesh: n must be >= 1")
    m = n + 1
    g = np.arange(m, dtype=np.float64) * (size / n)
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    positions = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)

    def vid(i, j, k):
        return i + m * (j + m * k)

    ii, jj, kk = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    ii, jj, kk = ii.ravel(), jj.ravel(), kk.ravel()
    # hex index order: k slowest, then j, then i
    order = np.lexsort((ii, jj, kk))
    ii, jj, kk = ii[order], jj[order], kk[order]
    tets = []
    for perm in _AXIS_PERMUTATIONS:
        c = [ii.copy(), jj.copy(), kk.copy()]
        path = [vid(*c)]
        for axis in perm:
            c[axis] = c[axis] + 1
            path.append(vid(*c))
        tets.append(np.stack(path, axis=1))
    tets = np.stack(tets, axis=1).reshape(-1, 4)
    # make every tet positively oriented: det(Dm) > 0
    p = positions[tets]
    dm = np.stack([p[:, 1] - p[:, 0], p[:, 2] - p[:, 0], p[:, 3] - p[:, 0]], axis=2)
    neg = np.linalg.det(dm) < 0
    tets[neg, 2], tets[neg, 3] = tets[neg, 3].copy(), tets[neg, 2].copy()
    return positions, tets.astype(np.int64)


@dataclass
class CubeScene:
    """A fixed/moving split tet cube ready for one Newton linear solve.

    Vertex ids index the union ``[fixed_vertices, moving_vertices]``: the
    ``n_fixed`` fixed vertices come first, as in ``one_bunny.py:58-60``.
    """
    n: int
    rest: np.ndarray            # [V,3] union order
    position: np.ndarray        # [V,3] union order (fixed rows equal rest)
    last_position: np.ndarray   # [V,3]
    velocity: np.ndarray        # [V,3]
    tets: np.ndarray            # [T,4] union vertex ids
    mass: np.ndarray            # [V] lumped rho*vol/4
    n_fixed: int
    mu: float
    lam: float
    dt: float
    young: float
    poisson: float
    density: float
    extras: dict = field(default_factory=dict)

    @property
    def n_moving(self) -> int:
        return self.rest.shape[0] - self.n_fixed

    @property
    def n_dof(self) -> int:
        return 3 * self.n_moving


def cube_scene(n: int = 8, young: float = 1e5, poisson: float = 0.4, density: float = 1000.0,
               dt: float = 1e-2, noise: float = 0.02, seed: int = 0) -> CubeScene:
    """Tet cube scene: [0,1]^3, n^3 hexes x 6 tets, bottom face (z=0) fixed.

    Moving vertices get positions rest + N(0, (noise*h)^2) from ``default_rng(seed)``;
    fixed vertices stay at rest. last_position = rest, velocity = 0.
    """
    pos, tets = cube_tet_mesh(n)
    h = 1.0 / n
    fixed_mask = pos[:, 2] == 0.0
    order = np.concatenate([np.nonzero(fixed_mask)[0], np.nonzero(~fixed_mask)[0]])
    inv = np.empty_like(order)
    inv[order] = np.arange(order.size)
    rest = pos[order]
    tets_u = inv[tets]
    n_fixed = int(fixed_mask.sum())
    rng = np.random.default_rng(seed)
    position = rest.copy()
    position[n_fixed:] += rng.normal(0.0, noise * h, size=(rest.shape[0] - n_fixed, 3))
    p = rest[tets_u]
    vol = np.linalg.det(np.stack([p[:, 1] - p[:, 0], p[:, 2] - p[:, 0], p[:, 3] - p[:, 0]], axis=2)) / 6.0
    mass = np.zeros(rest.shape[0])
    np.add.at(mass, tets_u.ravel(), np.repeat(density * vol / 4.0, 4))
    mu, lam = lame_parameters(young, poisson)
    return CubeScene(n=n, rest=rest, position=position, last_position=rest.copy(),
                     velocity=np.zeros_like(rest), tets=tets_u, mass=mass, n_fixed=n_fixed,
                     mu=mu, lam=lam, dt=dt, young=young, poisson=poisson, density=density)


def build_yasps_scene(cs: CubeScene, projection_method: int = 2, solver_name: str = "jacobian",
                      **solver_options):
    """The cube scene through the drop-in API, laid out as ``examples/one_bunny/one_bunny.py``
    (fixed/moving vertex primitives, union ``vertices``, tets joining the union, SNH on tets,
    inertia on moving vertices, target = moving positions). Returns (scene, handles)."""
    from . import scene, stable_neo_hookean, inertia
    s0 = scene("scene0")
    dt = s0.addConstant("dt", rows=1, cols=1)
    dt.updateValue([cs.dt])
    m = s0.addMesh("cube")
    fv = m.addPrimitive("fixed_vertices", numInstances=cs.n_fixed)
    mv = m.addPrimitive("moving_vertices", numInstances=cs.n_moving)
    m.addPrimitiveUnion("vertices", [fv, mv])
    tets = m.addPrimitive("tets", numInstances=cs.tets.shape[0])
    nf = cs.n_fixed
    fv.addAttribute("position", rows=3, cols=1).updateValue(cs.position[:nf])
    fv.addConstant("rest_position", rows=3, cols=1).updateValue(cs.rest[:nf])
    mv.addConstant("rest_position", rows=3, cols=1).updateValue(cs.rest[nf:])
    mv.addAttribute("position", rows=3, cols=1).updateValue(cs.position[nf:])
    mv.addConstant("last_position", rows=3, cols=1).updateValue(cs.last_position[nf:])
    mv.addConstant("velocity", rows=3, cols=1).updateValue(cs.velocity[nf:])
    mv.addConstant("mass", rows=1, cols=1).updateValue(cs.mass[nf:])
    m.vertices.addAttribute("position")
    m.vertices.addAttribute("rest_position")
    mu = m.addConstant("mu", rows=1, cols=1)
    mu.updateValue([cs.mu])
    lam = m.addConstant("lam", rows=1, cols=1)
    lam.updateValue([cs.lam])
    tet2v = tets.addConnectivity("tet2v", m.vertices, cs.tets, 4)
    tp = tets.addAttribute("positions", through=tet2v, source=m.vertices["position"])
    trp = tets.addAttribute("rest_positions", through=tet2v, source=m.vertices["rest_position"])
    snh = tets.addAttribute("stable_neo_hookean",
                            computed_attribute=stable_neo_hookean(trp, tp, mu, lam, dt))
    ine = mv.addAttribute("inertia", computed_attribute=inertia(
        mv["last_position"], mv["velocity"], dt, mv["position"], mv["mass"]))
    s0.addEnergy(snh, projection_method=projection_method)
    s0.addEnergy(ine, projection_method=projection_method)
    s0.addMinimizeTarget([mv["position"]])
    s0.minimizer.setSolver(solver_name, **solver_options)
    return s0, dict(mesh=m, tets=tets, moving=mv, fixed=fv, snh=snh, inertia=ine, tet2v=tet2v,
                    mu=mu, lam=lam, dt=dt)


@dataclass
class AbdScene:
    """Tet cube with a fixed bottom face, affine bodies over the top layers and free
    vertices in between (``examples/one_bunny_partial_abd/one_bunny_partial_abd.py``).
    Vertex order: fixed, affine-body vertices, free vertices."""
    cube: CubeScene
    n_fixed: int
    n_abd: int
    n_free: int
    body_of: np.ndarray          # [n_abd] body index of every affine vertex
    A: np.ndarray                # [bodies, 3, 3]
    t: np.ndarray                # [bodies, 3]
    affine_stiffness: float

    @property
    def n_bodies(self) -> int:
        return self.A.shape[0]

    def positions(self) -> np.ndarray:
        """Current positions of all vertices (affine ones from t + A r)."""
        cs = self.cube
        x = cs.position.copy()
        s = slice(self.n_fixed, self.n_fixed + self.n_abd)
        r = cs.rest[s]
        x[s] = self.t[self.body_of] + np.einsum("bij,bj->bi", self.A[self.body_of], r)
        return x


def abd_cube_scene(n: int = 4, abd_layers: int = 2, bodies: int = 2, noise: float = 0.02,
                   affine_noise: float = 0.03, affine_stiffness: float = 1e3, seed: int = 0,
                   **kw) -> AbdScene:
    """Cube scene whose top ``abd_layers`` vertex layers are split (by x) into ``bodies``
    affine bodies; A = I + N(0, affine_noise^2), t = N(0, (affine_noise h)^2)."""
    cs = cube_scene(n, noise=noise, seed=seed, **kw)
    nf = cs.n_fixed
    rest = cs.rest
    top = rest[:, 2] >= 1.0 - (abd_layers - 1) / n - 1e-12
    top[:nf] = False
    body = np.minimum((rest[:, 0] * bodies).astype(np.int64), bodies - 1)
    free = ~top
    free[:nf] = False
    order = np.concatenate([np.arange(nf), np.flatnonzero(top), np.flatnonzero(free)])
    inv = np.empty_like(order)
    inv[order] = np.arange(order.size)
    cs.rest = rest[order]
    cs.position = cs.position[order]
    cs.last_position = cs.last_position[order]
    cs.velocity = cs.velocity[order]
    cs.mass = cs.mass[order]
    cs.tets = inv[cs.tets]
    n_abd = int(top.sum())
    rng = np.random.default_rng(seed + 1)
    A = np.eye(3) + affine_noise * rng.standard_normal((bodies, 3, 3))
    t = affine_noise / n * rng.standard_normal((bodies, 3))
    return AbdScene(cube=cs, n_fixed=nf, n_abd=n_abd, n_free=cs.rest.shape[0] - nf - n_abd,
                    body_of=body[order][nf:nf + n_abd], A=A, t=t,
                    affine_stiffness=affine_stiffness)


def build_abd_scene(sc: AbdScene, projection_method: int = 2, solver_name: str = "jacobian",
                    **solver_options):
    """The affine-body scene through the drop-in API, as
    ``examples/one_bunny_partial_abd/one_bunny_partial_abd.py:102-258``: union
    [fixed, abd, moving] vertices, abd positions = translation + affine_matrix * rest through
    a vertex -> body connectivity, SNH on tets, inertia on abd and moving vertices, affine
    orthogonality on the bodies; targets [moving position, affine_matrix, translation]."""
    from . import scene, stable_neo_hookean, inertia
    from .energies import affine_energy
    cs = sc.cube
    s0 = scene("scene0")
    dt = s0.addConstant("dt", rows=1, cols=1)
    dt.updateValue([cs.dt])
    m = s0.addMesh("bunny")
    fv = m.addPrimitive("fixed_vertices", numInstances=sc.n_fixed)
    av = m.addPrimitive("abd_vertices", numInstances=sc.n_abd)
    ab = m.addPrimitive("affine_body", numInstances=sc.n_bodies)
    mv = m.addPrimitive("moving_vertices", numInstances=sc.n_free)
    m.addPrimitiveUnion("vertices", [fv, av, mv])
    tets = m.addPrimitive("tets", numInstances=cs.tets.shape[0])
    ab.addAttribute("affine_matrix", rows=3, cols=3).updateValue(sc.A.reshape(-1))
    ab.addAttribute("translation", rows=3, cols=1).updateValue(sc.t.reshape(-1))
    bav2ab = av.addConnectivity("bav2ab", ab, sc.body_of, 1)
    bavam = av.addAttribute("affine_matrix", through=bav2ab, source=ab["affine_matrix"]).resize(3, 3)
    bavt = av.addAttribute("translation", through=bav2ab, source=ab["translation"]).resize(3, 1)
    nf, na = sc.n_fixed, sc.n_abd
    sa = slice(nf, nf + na)
    sm = slice(nf + na, None)
    av.addConstant("rest_position", rows=3, cols=1).updateValue(cs.rest[sa])
    av.addAttribute("position", computed_attribute=bavt + bavam * av["rest_position"])
    av.addConstant("last_position", rows=3, cols=1).updateValue(cs.last_position[sa])
    av.addConstant("velocity", rows=3, cols=1).updateValue(cs.velocity[sa])
    av.addConstant("mass", rows=1, cols=1).updateValue(cs.mass[sa])
    fv.addAttribute("position", rows=3, cols=1).updateValue(cs.position[:nf])
    fv.addConstant("rest_position", rows=3, cols=1).updateValue(cs.rest[:nf])
    mv.addAttribute("position", rows=3, cols=1).updateValue(cs.position[sm])
    mv.addConstant("rest_position", rows=3, cols=1).updateValue(cs.rest[sm])
    mv.addConstant("last_position", rows=3, cols=1).updateValue(cs.last_position[sm])
    mv.addConstant("velocity", rows=3, cols=1).updateValue(cs.velocity[sm])
    mv.addConstant("mass", rows=1, cols=1).updateValue(cs.mass[sm])
    m.vertices.addAttribute("position")
    m.vertices.addAttribute("rest_position")
    mu = m.addConstant("mu", rows=1, cols=1)
    mu.updateValue([cs.mu])
    lam = m.addConstant("lam", rows=1, cols=1)
    lam.updateValue([cs.lam])
    tet2v = tets.addConnectivity("tet2v", m.vertices, cs.tets, 4)
    tp = tets.addAttribute("positions", through=tet2v, source=m.vertices["position"])
    trp = tets.addAttribute("rest_positions", through=tet2v, source=m.vertices["rest_position"])
    snh = tets.addAttribute("stable_neo_hookean",
                            computed_attribute=stable_neo_hookean(trp, tp, mu, lam, dt))
    aff = m.addAttribute("affine_energy",
                         computed_attribute=affine_energy(ab["affine_matrix"], sc.affine_stiffness))
    ia = av.addAttribute("inertia", computed_attribute=inertia(
        av["last_position"], av["velocity"], dt, av["position"], av["mass"]))
    im = mv.addAttribute("inertia", computed_attribute=inertia(
        mv["last_position"], mv["velocity"], dt, mv["position"], mv["mass"]))
    s0.addEnergy(snh, projection_method=projection_method)
    s0.addEnergy(aff, projection_method=projection_method)
    s0.addEnergy(ia, projection_method=-1)
    s0.addEnergy(im, projection_method=-1)
    s0.addMinimizeTarget([mv["position"], ab["affine_matrix"], ab["translation"]])
    s0.minimizer.setSolver(solver_name, **solver_options)
    return s0, dict(mesh=m, snh=snh, affine=aff, inertia_abd=ia, inertia_moving=im)


@dataclass
class ClothScene:
    """An m x m vertex grid in the plane z = 0 (spacing 1/(m-1)) cut into triangles, with
    its hinges (interior edges: the shared edge x0-x1 and the two opposite vertices, the
    layout of ``examples/many_bunnies_one_cloth/helpers.py:193-212``). Positions are the
    rest grid plus N(0, (noise h)^2) in every coordinate."""
    m: int
    rest: np.ndarray          # [m*m, 3]
    position: np.ndarray      # [m*m, 3]
    triangles: np.ndarray     # [2 (m-1)^2, 3], counter-clockwise seen from +z
    hinges: np.ndarray        # [interior edges, 4]


def cloth_scene(m: int, noise: float = 0.02, seed: int = 0) -> ClothScene:
    h = 1.0 / (m - 1)
    j, i = np.meshgrid(np.arange(m), np.arange(m), indexing="ij")
    rest = np.stack([i.ravel() * h, j.ravel() * h, np.zeros(m * m)], axis=1)
    v = lambda a, b: b * m + a                                      # noqa: E731
    tri = []
    for b in range(m - 1):
        for a in range(m - 1):
            tri.append((v(a, b), v(a + 1, b), v(a + 1, b + 1)))
            tri.append((v(a, b), v(a + 1, b + 1), v(a, b + 1)))
    tri = np.array(tri, dtype=np.int64)
    # a directed edge u -> v of one triangle meets v -> u of its neighbour
    opposite = {}
    for t in tri:
        for k in range(3):
            opposite[(int(t[k]), int(t[(k + 1) % 3]))] = int(t[(k + 2) % 3])
    hinges = [(a, b, c, opposite[(b, a)]) for (a, b), c in opposite.items() if (b, a) in opposite
              and a < b]
    rng = np.random.default_rng(seed)
    position = rest + rng.normal(0.0, noise * h, size=rest.shape)
    return ClothScene(m=m, rest=rest, position=position, triangles=tri,
                      hinges=np.array(hinges, dtype=np.int64))
