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


# ------------------------------------------------------------------ contact scene
def sphere_tet_mesh(cells: int, radius: float):
    """A ball voxelised on a cells^3 grid of [-radius, radius]^3: the hexes whose centre
    lies inside the ball, each split into 6 positively oriented tets. Returns
    (positions [V,3] centred on the origin, tets [T,4])."""
    pos, tets = cube_tet_mesh(cells, size=2.0 * radius)
    pos = pos - radius
    h = 2.0 * radius / cells
    centre = pos[tets[::6, 0]] + 0.5 * h          # every hex's 6 tets start at its min corner
    keep = np.repeat((centre ** 2).sum(axis=1) <= radius * radius, 6)
    tets = tets[keep]
    used = np.unique(tets)
    remap = np.full(pos.shape[0], -1, dtype=np.int64)
    remap[used] = np.arange(used.size)
    return pos[used], remap[tets]


def boundary_faces(tets: np.ndarray) -> np.ndarray:
    """Faces [F,3] that belong to exactly one tet."""
    f = np.concatenate([tets[:, [0, 1, 2]], tets[:, [0, 1, 3]], tets[:, [0, 2, 3]],
                        tets[:, [1, 2, 3]]])
    key = np.sort(f, axis=1)
    _, idx, cnt = np.unique(key, axis=0, return_index=True, return_counts=True)
    return f[idx[cnt == 1]]


def grid_cloth(m: int, noise: float, rng):
    """An m x m cloth in the plane y = 0 over [0,1]^2 (x, z), triangles oriented with
    normal +y, hinges (a, b, c, d) for the interior edges (shared edge a-b with a < b,
    c opposite in the triangle holding a -> b, d opposite in the one holding b -> a, as
    ``examples/many_bunnies_one_cloth/helpers.py:193-212``). In-plane noise only, so
    every triangle stays in y = 0."""
    h = 1.0 / (m - 1)
    j, i = np.meshgrid(np.arange(m), np.arange(m), indexing="ij")
    rest = np.stack([i.ravel() * h, np.zeros(m * m), j.ravel() * h], axis=1)
    a, b = np.meshgrid(np.arange(m - 1), np.arange(m - 1), indexing="xy")
    a, b = a.ravel(), b.ravel()
    v00, v10, v11, v01 = b * m + a, b * m + a + 1, (b + 1) * m + a + 1, (b + 1) * m + a
    tri = np.stack([np.stack([v00, v11, v10], 1), np.stack([v00, v01, v11], 1)], 1).reshape(-1, 3)
    src = np.concatenate([tri[:, 0], tri[:, 1], tri[:, 2]])
    dst = np.concatenate([tri[:, 1], tri[:, 2], tri[:, 0]])
    opp = np.concatenate([tri[:, 2], tri[:, 0], tri[:, 1]])
    key = np.minimum(src, dst) * (m * m) + np.maximum(src, dst)
    order = np.argsort(key, kind="stable")
    k = key[order]
    first = np.flatnonzero(k[:-1] == k[1:])
    e0, e1 = order[first], order[first + 1]
    fwd = np.where(src[e0] < dst[e0], e0, e1)      # the directed edge a -> b with a < b
    bwd = np.where(src[e0] < dst[e0], e1, e0)
    hinges = np.stack([src[fwd], dst[fwd], opp[fwd], opp[bwd]], axis=1)
    position = rest.copy()
    position[:, [0, 2]] += rng.normal(0.0, noise * h, size=(m * m, 2))
    return rest, position, tri.astype(np.int64), hinges.astype(np.int64)


def _point_triangle_dist2_flat(p, t0, t1, t2):
    """Squared distance from points p [n,3] to triangles lying in y = 0 (footprints in x, z)."""
    q = p[:, [0, 2]]
    a, b, c = t0[:, [0, 2]], t1[:, [0, 2]], t2[:, [0, 2]]

    def cross(u, v):
        return u[:, 0] * v[:, 1] - u[:, 1] * v[:, 0]

    s0, s1, s2 = cross(b - a, q - a), cross(c - b, q - b), cross(a - c, q - c)
    inside = ((s0 >= 0) & (s1 >= 0) & (s2 >= 0)) | ((s0 <= 0) & (s1 <= 0) & (s2 <= 0))

    def seg2(u, v):
        e = v - u
        t = np.clip(((q - u) * e).sum(1) / (e * e).sum(1), 0.0, 1.0)
        d = q - (u + t[:, None] * e)
        return (d * d).sum(1)

    planar = np.where(inside, 0.0, np.minimum(np.minimum(seg2(a, b), seg2(b, c)), seg2(c, a)))
    return planar + p[:, 1] ** 2


@dataclass
class ContactScene:
    """Bodies resting on a cloth, laid out for one Newton linear solve with contact.

    Union vertex order: [fixed (the cloth's border), affine-body vertices, moving (free-body
    vertices, then the cloth's interior)]. Affine-body vertices follow x = t + A r through a
    vertex -> body connectivity, as
    ``examples/one_bunny_partial_abd/one_bunny_partial_abd.py:112-144``, with r the rest
    position in the body's frame (centred on the body; t starts at the centre), so that A
    can depart from the identity without moving the body onto the cloth."""
    rest: np.ndarray
    position: np.ndarray
    last_position: np.ndarray
    velocity: np.ndarray
    mass: np.ndarray
    tets: np.ndarray              # body tets (union ids)
    triangles: np.ndarray         # cloth triangles (union ids)
    hinges: np.ndarray            # cloth hinges (union ids)
    pt: np.ndarray                # [n, 4] point (body surface vertex), cloth triangle
    ee: np.ndarray                # [n, 4] body surface edge, cloth edge
    n_fixed: int
    n_abd: int
    n_free: int
    body_of: np.ndarray           # [n_abd] affine body of every affine vertex
    A: np.ndarray                 # [bodies, 3, 3]
    t: np.ndarray                 # [bodies, 3]
    n_free_bodies: int
    mu: float
    lam: float
    dt: float
    dhat: float
    kappa: float
    cloth_mu: float
    cloth_lam: float
    thickness: float
    bend: float
    affine_stiffness: float

    @property
    def n_bodies(self) -> int:
        return self.A.shape[0]

    def positions(self) -> np.ndarray:
        x = self.position.copy()
        s = slice(self.n_fixed, self.n_fixed + self.n_abd)
        x[s] = self.t[self.body_of] + np.einsum("bij,bj->bi", self.A[self.body_of], self.rest[s])
        return x


def contact_scene(bodies_x: int = 5, bodies_z: int = 8, cells: int = 16, radius: float = 0.045,
                  cloth_m: int = 256, dhat: float = 1e-3, kappa: float = 1e-2,
                  max_pairs: int = 1_500_000, young: float = 1e5, poisson: float = 0.4,
                  density: float = 1000.0, dt: float = 1e-2, seed: int = 0) -> ContactScene:
    """Contact scene with mixed DoFs: bodies_x x bodies_z voxelised balls on a jittered grid
    above a cloth_m x cloth_m cloth (StVK membrane + dihedral bending, border pinned), every
    other ball a 12-DoF affine body, the others free vertices. Every ball's lowest point
    sits U(0.2, 2) * dhat above the cloth. Barrier pairs (squared-distance convention of
    ``examples/one_bunny/helpers.py:74-120``: active when d < dhat): point-triangle pairs
    of body surface vertices and cloth triangles whose squared distance is below dhat, and
    edge-edge pairs of body surface edges and cloth edges whose closest points are interior
    to both edges, not parallel, with squared distance below dhat; at most max_pairs in
    total (a uniform sample when more are active)."""
    rng = np.random.default_rng(seed)
    sp, st = sphere_tet_mesh(cells, radius)
    bottom = -sp[:, 1].min()
    sfaces = boundary_faces(st)
    ssurf = np.unique(sfaces)
    sedges = np.unique(np.sort(np.concatenate([sfaces[:, [0, 1]], sfaces[:, [1, 2]],
                                               sfaces[:, [0, 2]]]), axis=1), axis=0)
    nb = bodies_x * bodies_z
    affine = np.arange(nb) % 2 == 0
    jitter = 0.1 * min(1.0 / bodies_x, 1.0 / bodies_z)
    centres = []
    for j in range(bodies_z):
        for i in range(bodies_x):
            gap = rng.uniform(0.2, 2.0) * dhat
            centres.append([(i + 0.5) / bodies_x + rng.uniform(-jitter, jitter), bottom + gap,
                            (j + 0.5) / bodies_z + rng.uniform(-jitter, jitter)])
    centres = np.array(centres)
    c_rest, c_pos, c_tri, c_hinge = grid_cloth(cloth_m, 0.005, rng)
    border = ((c_rest[:, 0] == 0.0) | (c_rest[:, 2] == 0.0) | (c_rest[:, 0] >= 1.0 - 1e-12)
              | (c_rest[:, 2] >= 1.0 - 1e-12))
    nvs = sp.shape[0]
    # union order: fixed cloth, affine bodies, free bodies, moving cloth
    abd_bodies = np.flatnonzero(affine)
    free_bodies = np.flatnonzero(~affine)
    n_fixed = int(border.sum())
    n_abd = abd_bodies.size * nvs
    base = {}
    off = n_fixed
    for b in abd_bodies:
        base[b] = off
        off += nvs
    for b in free_bodies:
        base[b] = off
        off += nvs
    cloth_id = np.empty(cloth_m * cloth_m, dtype=np.int64)
    cloth_id[border] = np.arange(n_fixed)
    cloth_id[~border] = off + np.arange(int((~border).sum()))
    nv = off + int((~border).sum())
    rest = np.zeros((nv, 3))
    position = np.zeros((nv, 3))
    rest[cloth_id] = c_rest
    position[cloth_id] = c_pos
    tets = []
    for b in range(nb):
        s = slice(base[b], base[b] + nvs)
        # affine bodies keep body-frame rest coordinates r (x = t + A r, t = the centre)
        rest[s] = sp if affine[b] else sp + centres[b]
        position[s] = sp + centres[b] + (0.0 if affine[b] else rng.normal(
            0.0, 1e-3 * 2.0 * radius / cells, size=(nvs, 3)))
        tets.append(st + base[b])
    tets = np.concatenate(tets)
    A = np.eye(3) + 1e-3 * rng.standard_normal((abd_bodies.size, 3, 3))
    t = centres[abd_bodies] + 1e-5 * rng.standard_normal((abd_bodies.size, 3))
    body_of = np.repeat(np.arange(abd_bodies.size), nvs)
    sa = slice(n_fixed, n_fixed + n_abd)
    position[sa] = t[body_of] + np.einsum("bij,bj->bi", A[body_of], rest[sa])
    # masses: rho vol / 4 per tet vertex, cloth areal density 0.2 kg/m^2 / 3 per triangle
    p = rest[tets]
    vol = np.linalg.det(np.stack([p[:, 1] - p[:, 0], p[:, 2] - p[:, 0], p[:, 3] - p[:, 0]], 2)) / 6.0
    mass = np.zeros(nv)
    np.add.at(mass, tets.ravel(), np.repeat(density * vol / 4.0, 4))
    tri_u = cloth_id[c_tri]
    tr = rest[tri_u]
    area = 0.5 * np.linalg.norm(np.cross(tr[:, 1] - tr[:, 0], tr[:, 2] - tr[:, 0]), axis=1)
    np.add.at(mass, tri_u.ravel(), np.repeat(0.2 * area / 3.0, 3))
    # ---- barrier pairs on the current positions
    hc = 1.0 / (cloth_m - 1)
    reach = np.sqrt(dhat)
    rc = int(np.ceil(reach / hc)) + 1
    pts, tris = [], []
    edges_b, edges_c = [], []
    cedge = np.concatenate([c_tri[:, [0, 1]], c_tri[:, [1, 2]], c_tri[:, [0, 2]]])
    cedge = np.unique(np.sort(cedge, axis=1), axis=0)
    ce_mid = 0.5 * (c_pos[cedge[:, 0]] + c_pos[cedge[:, 1]])
    ce_cell = (np.clip(np.floor(ce_mid[:, 0] / hc), 0, cloth_m - 2).astype(np.int64) +
               (cloth_m - 1) * np.clip(np.floor(ce_mid[:, 2] / hc), 0, cloth_m - 2).astype(np.int64))
    ce_order = np.argsort(ce_cell, kind="stable")
    ce_start = np.searchsorted(ce_cell[ce_order], np.arange((cloth_m - 1) ** 2 + 1))
    for b in range(nb):
        x = position[base[b]:base[b] + nvs]
        near = ssurf[x[ssurf, 1] < reach]
        if near.size:
            q = x[near]
            ca = np.floor(q[:, 0] / hc).astype(np.int64)
            cb = np.floor(q[:, 2] / hc).astype(np.int64)
            da, db = np.meshgrid(np.arange(-rc, rc + 1), np.arange(-rc, rc + 1), indexing="ij")
            a = ca[:, None] + da.ravel()[None]
            bb = cb[:, None] + db.ravel()[None]
            ok = (a >= 0) & (a < cloth_m - 1) & (bb >= 0) & (bb < cloth_m - 1)
            vi = np.broadcast_to(near[:, None], a.shape)[ok]
            cell = (bb * (cloth_m - 1) + a)[ok]
            vi = np.concatenate([vi, vi])
            tri_i = np.concatenate([2 * cell, 2 * cell + 1])
            t3 = c_tri[tri_i]
            d2 = _point_triangle_dist2_flat(x[vi], c_pos[t3[:, 0]], c_pos[t3[:, 1]], c_pos[t3[:, 2]])
            keep = d2 < dhat
            pts.append(vi[keep] + base[b])
            tris.append(tri_i[keep])
        # edge-edge
        e = sedges[(x[sedges[:, 0], 1] < reach) & (x[sedges[:, 1], 1] < reach)]
        if e.size:
            mid = 0.5 * (x[e[:, 0]] + x[e[:, 1]])
            ca = np.floor(mid[:, 0] / hc).astype(np.int64)
            cb = np.floor(mid[:, 2] / hc).astype(np.int64)
            cand_e, cand_c = [], []
            for da in range(-2, 3):
                for db in range(-2, 3):
                    a = ca + da
                    bb = cb + db
                    ok = (a >= 0) & (a < cloth_m - 1) & (bb >= 0) & (bb < cloth_m - 1)
                    cell = bb[ok] * (cloth_m - 1) + a[ok]
                    cnt = ce_start[cell + 1] - ce_start[cell]
                    ei = np.repeat(np.flatnonzero(ok), cnt)
                    first = np.repeat(ce_start[cell], cnt)
                    within = np.arange(cnt.sum()) - np.repeat(np.cumsum(cnt) - cnt, cnt)
                    cand_e.append(ei)
                    cand_c.append(ce_order[first + within])
            ei = np.concatenate(cand_e)
            ci = np.concatenate(cand_c)
            p0, p1 = x[e[ei, 0]], x[e[ei, 1]]
            p2, p3 = c_pos[cedge[ci, 0]], c_pos[cedge[ci, 1]]
            u, v, w = p1 - p0, p3 - p2, p0 - p2
            aa, bq, cc = (u * u).sum(1), (u * v).sum(1), (v * v).sum(1)
            dd, ee_ = (u * w).sum(1), (v * w).sum(1)
            den = aa * cc - bq * bq
            sin2 = den / (aa * cc)
            with np.errstate(divide="ignore", invalid="ignore"):
                s_ = (bq * ee_ - cc * dd) / den
                t_ = (aa * ee_ - bq * dd) / den
            cp = w + s_[:, None] * u - t_[:, None] * v
            d2 = (cp * cp).sum(1)
            keep = ((sin2 > 0.01) & (s_ > 0.05) & (s_ < 0.95) & (t_ > 0.05) & (t_ < 0.95)
                    & (d2 < dhat))
            edges_b.append(e[ei[keep]] + base[b])
            edges_c.append(cedge[ci[keep]])
    cat = lambda parts, shape: (np.concatenate(parts) if parts            # noqa: E731
                                else np.zeros(shape, dtype=np.int64))
    pt = np.concatenate([cat(pts, (0,))[:, None], cloth_id[c_tri[cat(tris, (0,))]]], 1)
    ee = np.concatenate([cat(edges_b, (0, 2)), cloth_id[cat(edges_c, (0, 2))]], 1)
    total = pt.shape[0] + ee.shape[0]
    if total > max_pairs:
        keep = np.sort(rng.choice(total, size=max_pairs, replace=False))
        npt = pt.shape[0]
        pt = pt[keep[keep < npt]]
        ee = ee[keep[keep >= npt] - npt]
    mu, lam = lame_parameters(young, poisson)
    return ContactScene(rest=rest, position=position, last_position=position.copy(),
                        velocity=np.zeros_like(rest), mass=mass, tets=tets,
                        triangles=tri_u, hinges=cloth_id[c_hinge], pt=pt.astype(np.int64),
                        ee=ee.astype(np.int64), n_fixed=n_fixed, n_abd=n_abd,
                        n_free=nv - n_fixed - n_abd, body_of=body_of, A=A, t=t,
                        n_free_bodies=free_bodies.size, mu=mu, lam=lam, dt=dt, dhat=dhat,
                        kappa=kappa, cloth_mu=1e3, cloth_lam=1e3, thickness=1e-3, bend=1e-4,
                        affine_stiffness=1e3)


def build_contact_scene(cs: ContactScene, projection_method: int = 2, solver_name: str = "jacobian",
                        barriers=None, **solver_options):
    """The contact scene through the drop-in API: union [fixed, abd, moving] vertices
    (``one_bunny_partial_abd.py:102-258`` for the affine part), SNH on the body tets, StVK
    membrane and dihedral bending on the cloth, IPC point-triangle and edge-edge barriers on
    dynamic pair primitives (``dynamic_instances=True``, re-indexed on every assembly as
    ``hessian.pyx:846-880``), inertia on abd and moving vertices, affine orthogonality on
    the bodies; targets [moving position, affine_matrix, translation]. ``barriers`` may
    replace the point-triangle / edge-edge energy constructors ({"pt": fn, "ee": fn},
    fn(positions, dhat, kappa)), e.g. by user energies."""
    from . import scene, stable_neo_hookean, inertia
    from . import energies as en
    barriers = dict(barriers or {})
    s0 = scene("contact")
    dt = s0.addConstant("dt", rows=1, cols=1)
    dt.updateValue([cs.dt])
    m = s0.addMesh("world")
    fv = m.addPrimitive("fixed_vertices", numInstances=cs.n_fixed)
    av = m.addPrimitive("abd_vertices", numInstances=cs.n_abd)
    ab = m.addPrimitive("affine_body", numInstances=cs.n_bodies)
    mv = m.addPrimitive("moving_vertices", numInstances=cs.n_free)
    m.addPrimitiveUnion("vertices", [fv, av, mv])
    ab.addAttribute("affine_matrix", rows=3, cols=3).updateValue(cs.A.reshape(-1))
    ab.addAttribute("translation", rows=3, cols=1).updateValue(cs.t.reshape(-1))
    bav2ab = av.addConnectivity("bav2ab", ab, cs.body_of, 1)
    bavam = av.addAttribute("affine_matrix", through=bav2ab, source=ab["affine_matrix"]).resize(3, 3)
    bavt = av.addAttribute("translation", through=bav2ab, source=ab["translation"]).resize(3, 1)
    nf, na = cs.n_fixed, cs.n_abd
    sa, sm = slice(nf, nf + na), slice(nf + na, None)
    fv.addAttribute("position", rows=3, cols=1).updateValue(cs.position[:nf])
    fv.addConstant("rest_position", rows=3, cols=1).updateValue(cs.rest[:nf])
    av.addConstant("rest_position", rows=3, cols=1).updateValue(cs.rest[sa])
    av.addAttribute("position", computed_attribute=bavt + bavam * av["rest_position"])
    mv.addAttribute("position", rows=3, cols=1).updateValue(cs.position[sm])
    mv.addConstant("rest_position", rows=3, cols=1).updateValue(cs.rest[sm])
    for prim, s in ((av, sa), (mv, sm)):
        prim.addConstant("last_position", rows=3, cols=1).updateValue(cs.last_position[s])
        prim.addConstant("velocity", rows=3, cols=1).updateValue(cs.velocity[s])
        prim.addConstant("mass", rows=1, cols=1).updateValue(cs.mass[s])
    m.vertices.addAttribute("position")
    m.vertices.addAttribute("rest_position")
    mu = m.addConstant("mu", rows=1, cols=1)
    mu.updateValue([cs.mu])
    lam = m.addConstant("lam", rows=1, cols=1)
    lam.updateValue([cs.lam])
    handles = dict(mesh=m)
    tets = m.addPrimitive("tets", numInstances=cs.tets.shape[0])
    tet2v = tets.addConnectivity("tet2v", m.vertices, cs.tets, 4)
    tp = tets.addAttribute("positions", through=tet2v, source=m.vertices["position"])
    trp = tets.addAttribute("rest_positions", through=tet2v, source=m.vertices["rest_position"])
    handles["snh"] = tets.addAttribute("stable_neo_hookean",
                                       computed_attribute=stable_neo_hookean(trp, tp, mu, lam, dt))
    for name, conn, arity in (("cloth_triangles", cs.triangles, 3), ("cloth_hinges", cs.hinges, 4)):
        el = m.addPrimitive(name, numInstances=conn.shape[0])
        e2v = el.addConnectivity(name + "2v", m.vertices, conn, arity)
        pos = el.addAttribute("positions", through=e2v, source=m.vertices["position"])
        rp = el.addAttribute("rest_positions", through=e2v, source=m.vertices["rest_position"])
        if arity == 3:
            e = en.stvk_membrane(rp, pos, cs.cloth_mu, cs.cloth_lam, cs.thickness, cs.dt)
            handles["stvk"] = el.addAttribute("stvk_membrane", computed_attribute=e)
        else:
            handles["bending"] = el.addAttribute("bending",
                                                 computed_attribute=en.bending(pos, rp, cs.bend))
    for name, pairs, fn in (("pt", cs.pt, barriers.get("pt", en.point_triangle)),
                            ("ee", cs.ee, barriers.get("ee", en.edge_edge))):
        prim = m.addPrimitive(name, numInstances=pairs.shape[0], isDynamic=True)
        conn = prim.addConnectivity(name + "2v", m.vertices, pairs, 4)
        pos = prim.addAttribute("positions", through=conn, source=m.vertices["position"])
        handles[name] = prim.addAttribute(name + "_barrier",
                                          computed_attribute=fn(pos, cs.dhat, cs.kappa))
        handles[name + "_primitive"] = prim
        handles[name + "_connectivity"] = conn
    handles["affine"] = m.addAttribute("affine_energy", computed_attribute=en.affine_energy(
        ab["affine_matrix"], cs.affine_stiffness))
    handles["inertia_abd"] = av.addAttribute("inertia", computed_attribute=inertia(
        av["last_position"], av["velocity"], dt, av["position"], av["mass"]))
    handles["inertia_moving"] = mv.addAttribute("inertia", computed_attribute=inertia(
        mv["last_position"], mv["velocity"], dt, mv["position"], mv["mass"]))
    for key in ("snh", "stvk", "bending", "affine"):
        s0.addEnergy(handles[key], projection_method=projection_method)
    for key in ("pt", "ee"):
        s0.addEnergy(handles[key], projection_method=projection_method, dynamic_instances=True)
    s0.addEnergy(handles["inertia_abd"], projection_method=-1)
    s0.addEnergy(handles["inertia_moving"], projection_method=-1)
    s0.addMinimizeTarget([mv["position"], ab["affine_matrix"], ab["translation"]])
    s0.minimizer.setSolver(solver_name, **solver_options)
    return s0, handles


# --------------------------------------------------------------- drop onto a floor
@dataclass
class DropScene:
    """A soft tet cube, turned and tilted, its lowest corner `height` above a fixed
    triangulated floor (the plane y = 0; gravity acts along -y in the inertia term). Vertices: the floor's first (fixed), then the cube's. Contact
    barriers come from the collision module each trial step: pp / pe / pt / ee pairs below
    dhat (squared distance) with stiffness kappa."""
    rest: np.ndarray
    position: np.ndarray
    tets: np.ndarray              # union ids of the cube's tets
    mass: np.ndarray              # moving (cube) vertices
    n_fixed: int
    faces: np.ndarray             # floor triangles then the cube's boundary (union ids)
    edges: np.ndarray             # unique edges of those faces
    surface: np.ndarray           # every floor vertex and the cube's boundary vertices
    mesh_ids: np.ndarray          # 1 floor, 2 cube: no self-contact
    mu: float
    lam: float
    dt: float
    dhat: float
    kappa: float

    @property
    def n_moving(self) -> int:
        return self.rest.shape[0] - self.n_fixed


def drop_scene(n: int = 3, side: float = 0.2, height: float = 0.02, floor_n: int = 8,
               floor_size: float = 0.8, young: float = 1e5, poisson: float = 0.4,
               density: float = 1000.0, dt: float = 1e-2, dhat: float = 1e-4,
               kappa: float = 10.0) -> DropScene:
    g = np.linspace(-(floor_size - side) / 2, (floor_size + side) / 2, floor_n)
    fx, fz = np.meshgrid(g, g, indexing="ij")
    floor = np.stack([fx.ravel(), np.zeros(fx.size), fz.ravel()], axis=1)
    ff = []
    for i in range(floor_n - 1):
        for j in range(floor_n - 1):
            a = i * floor_n + j
            ff += [[a, a + 1, a + floor_n + 1], [a, a + floor_n + 1, a + floor_n]]
    ff = np.array(ff, dtype=np.int64)
    pos, tets = cube_tet_mesh(n, side)
    pos = pos[:, [0, 2, 1]]                       # the cube's third axis becomes y
    # the axis swap mirrors the mesh: swap two vertices of every tet back to det > 0
    tets = tets[:, [0, 1, 3, 2]]
    # turned about y and tilted about x, so that no cube edge runs parallel to a floor edge
    # (edge-edge distances of near-parallel lines are ill-conditioned without a mollifier)
    a, b = np.radians(30.0), np.radians(12.0)
    ry = np.array([[np.cos(a), 0, np.sin(a)], [0, 1, 0], [-np.sin(a), 0, np.cos(a)]])
    rx = np.array([[1, 0, 0], [0, np.cos(b), -np.sin(b)], [0, np.sin(b), np.cos(b)]])
    c = pos.mean(axis=0)
    pos = (pos - c) @ (rx @ ry).T + [side / 2, 0.0, side / 2]
    pos[:, 1] += height - pos[:, 1].min()
    nf = floor.shape[0]
    rest = np.concatenate([floor, pos])
    tets = tets + nf
    bf = boundary_faces(tets)
    faces = np.concatenate([ff, bf])
    edges = np.array(sorted({tuple(sorted((int(t[a]), int(t[b])))) for t in faces
                             for a, b in ((0, 1), (1, 2), (2, 0))}), dtype=np.int64)
    surface = np.concatenate([np.arange(nf), np.unique(bf)])
    p = rest[tets]
    vol = np.linalg.det(np.stack([p[:, 1] - p[:, 0], p[:, 2] - p[:, 0],
                                  p[:, 3] - p[:, 0]], axis=2)) / 6.0
    assert (vol > 0).all()
    mass = np.zeros(rest.shape[0])
    np.add.at(mass, tets.ravel(), np.repeat(density * vol / 4.0, 4))
    mu, lam = lame_parameters(young, poisson)
    mesh_ids = np.r_[np.ones(nf), 2 * np.ones(pos.shape[0])].astype(np.uint32)
    return DropScene(rest=rest, position=rest.copy(), tets=tets, mass=mass[nf:], n_fixed=nf,
                     faces=faces, edges=edges, surface=surface, mesh_ids=mesh_ids, mu=mu, lam=lam,
                     dt=dt, dhat=dhat, kappa=kappa)


def build_drop_scene(ds: DropScene, projection_method: int = 2, solver_name: str = "jacobian",
                     **solver_options):
    """The drop scene through the drop-in API as ``examples/one_bunny/one_bunny.py`` lays
    out its contact scene: fixed/moving vertex union, SNH on the tets, inertia (with
    gravity) on the moving vertices and four dynamic barrier primitives pp / pe / pt / ee
    over the union (empty until the collision module fills them). Returns (scene, handles)."""
    from . import scene, stable_neo_hookean, inertia
    from . import energies as en
    s0 = scene("drop")
    dt = s0.addConstant("dt", rows=1, cols=1)
    dt.updateValue([ds.dt])
    m = s0.addMesh("drop")
    nf = ds.n_fixed
    fv = m.addPrimitive("fixed_vertices", numInstances=nf)
    mv = m.addPrimitive("moving_vertices", numInstances=ds.n_moving)
    m.addPrimitiveUnion("vertices", [fv, mv])
    fv.addAttribute("position", rows=3, cols=1).updateValue(ds.position[:nf])
    fv.addConstant("rest_position", rows=3, cols=1).updateValue(ds.rest[:nf])
    mv.addAttribute("position", rows=3, cols=1).updateValue(ds.position[nf:])
    mv.addConstant("rest_position", rows=3, cols=1).updateValue(ds.rest[nf:])
    mv.addConstant("last_position", rows=3, cols=1).updateValue(ds.position[nf:])
    mv.addConstant("velocity", rows=3, cols=1).updateValue(np.zeros((ds.n_moving, 3)))
    mv.addConstant("mass", rows=1, cols=1).updateValue(ds.mass)
    m.vertices.addAttribute("position")
    m.vertices.addAttribute("rest_position")
    mu = m.addConstant("mu", rows=1, cols=1)
    mu.updateValue([ds.mu])
    lam = m.addConstant("lam", rows=1, cols=1)
    lam.updateValue([ds.lam])
    tets = m.addPrimitive("tets", numInstances=ds.tets.shape[0])
    t2v = tets.addConnectivity("tet2v", m.vertices, ds.tets, 4)
    tp = tets.addAttribute("positions", through=t2v, source=m.vertices["position"])
    trp = tets.addAttribute("rest_positions", through=t2v, source=m.vertices["rest_position"])
    handles = {"mesh": m, "moving": mv, "fixed": fv}
    handles["snh"] = tets.addAttribute("snh", computed_attribute=stable_neo_hookean(trp, tp, mu, lam, dt))
    handles["inertia"] = mv.addAttribute("inertia", computed_attribute=inertia(
        mv["last_position"], mv["velocity"], dt, mv["position"], mv["mass"]))
    s0.addEnergy(handles["snh"], projection_method=projection_method)
    s0.addEnergy(handles["inertia"], projection_method=-1)
    for name, arity, fn in (("pp", 2, en.point_point), ("pe", 3, en.point_edge),
                            ("pt", 4, en.point_triangle), ("ee", 4, en.edge_edge)):
        prim = m.addPrimitive(name, numInstances=0, isDynamic=True)
        conn = prim.addConnectivity(name + "2v", m.vertices, np.zeros((0, arity), dtype=np.int64), arity)
        pos = prim.addAttribute("positions", through=conn, source=m.vertices["position"])
        handles[name] = prim.addAttribute(name + "_barrier", computed_attribute=fn(pos, ds.dhat, ds.kappa))
        handles[name + "_primitive"] = prim
        handles[name + "_connectivity"] = conn
        s0.addEnergy(handles[name], projection_method=projection_method, dynamic_instances=True)
    s0.addMinimizeTarget([mv["position"]])
    s0.minimizer.setSolver(solver_name, **solver_options)
    return s0, handles
