"""CPU restatement of the contact-set module: proximity pairs by distance type, swept
broad-phase candidates and the additive CCD step bound (examples/ccd/mlbvh.cu,
examples/ccd/ACCD.cu, driven by examples/ccd/ccd.py). Test infrastructure only: the
product path is yasps_amd/csrc/ccd.hip; nothing here is imported by it.

Semantics restated (the reference's file:line for each):

* squared distances `_d_PP`, `_d_PE` (to the edge's line), `_d_PT` (to the triangle's
  plane), `_d_EE` (between the edges' lines) (mlbvh.cu:404-433);
* the point-triangle distance type (mlbvh.cu:453-522): for each edge (t0 t1), (t1 t2),
  (t2 t0) in turn, s = the point's coordinate along the edge and the sign of its offset
  along e x n (n the triangle normal); 0 < s < 1 and offset >= 0 selects that edge
  (types 3, 4, 5); otherwise a vertex whose two incident edges place the point before /
  after it (types 0, 1, 2); otherwise the face (6). The reference solves the same
  coordinates with Cramer's rule on the orthogonal frame (e, e x n, n);
* the edge-edge distance type (mlbvh.cu:524-591): the closest-point parameters of the two
  segments (numerators sN, tN over D = ac - b^2, the parallel fallback), types 0-8;
* proximity: every (surface vertex, face) pair without a shared vertex and every edge
  pair i < j without a shared vertex, unless all four vertices carry the same nonzero mesh
  id, emits its distance-type pair when the squared distance is below dhat
  (mlbvh.cu:594-725, :1015-1137); the face BVH only skips pairs that cannot pass that test,
  so a brute-force loop gives the same multiset;
* swept candidates: boxes of x and x - alpha d per vertex, overlap with the gap
  sqrt(dhat) (`overlap`, mlbvh.cu:225-232; leaf boxes :95-119; point query :1015-1080; edge
  query i < j :1082-1137);
* additive CCD (ACCD.cu:258-420, kernel :676-730, host :776-805): per candidate, the
  displacement is -d (the step moves x to x - t d), centred by the mean of the four
  displacements; conservative advancement by (1 - eta) (d^2 - xi^2) / ((d + xi) l_max)
  with eta = 1 - slackness, xi = 0, until the distance falls below eta d0 or the time
  passes 1; the step bound is the minimum over the candidates (1 without candidates).
"""
from __future__ import annotations

import numpy as np


# ------------------------------------------------------------------- distances
def _cross(a, b):
    return np.cross(a, b)


def _dot(a, b):
    return np.sum(a * b, axis=-1)


def d_pp(v0, v1):
    return _dot(v0 - v1, v0 - v1)


def d_pe(v0, v1, v2):
    c = _cross(v1 - v0, v2 - v0)
    e = v2 - v1
    return _dot(c, c) / _dot(e, e)


def d_pt(v0, v1, v2, v3):
    b = _cross(v2 - v1, v3 - v1)
    a = _dot(v0 - v1, b)
    return a * a / _dot(b, b)


def d_ee(v0, v1, v2, v3):
    b = _cross(v1 - v0, v3 - v2)
    a = _dot(v2 - v0, b)
    return a * a / _dot(b, b)


# ---------------------------------------------------------------- distance types
def _edge_coords(p, a, b, n):
    """Coordinate of p - a along the edge a -> b and the sign side of its offset along
    (b - a) x n."""
    e = b - a
    w = p - a
    s = _dot(w, e) / _dot(e, e)
    side = _dot(w, _cross(e, n))
    return s, side


def pt_type(p, t0, t1, t2):
    """Distance type of point p against triangle (t0, t1, t2), scalar inputs."""
    n = _cross(t1 - t0, t2 - t0)
    s0, o0 = _edge_coords(p, t0, t1, n)
    if 0.0 < s0 < 1.0 and o0 >= 0.0:
        return 3
    s1, o1 = _edge_coords(p, t1, t2, n)
    if 0.0 < s1 < 1.0 and o1 >= 0.0:
        return 4
    s2, o2 = _edge_coords(p, t2, t0, n)
    if 0.0 < s2 < 1.0 and o2 >= 0.0:
        return 5
    if s0 <= 0.0 and s2 >= 1.0:
        return 0
    if s1 <= 0.0 and s0 >= 1.0:
        return 1
    if s2 <= 0.0 and s1 >= 1.0:
        return 2
    return 6


def ee_type(v0, v1, v2, v3):
    """Distance type of edges (v0, v1) and (v2, v3), scalar inputs."""
    u = v1 - v0
    v = v3 - v2
    w = v0 - v2
    a = _dot(u, u)
    b = _dot(u, v)
    c = _dot(v, v)
    d = _dot(u, w)
    e = _dot(v, w)
    D = a * c - b * b
    tD = D
    case = 8
    sN = b * e - c * d
    if sN <= 0.0:
        tN, tD, case = e, c, 2
    elif sN >= D:
        tN, tD, case = e + b, c, 5
    else:
        tN = a * e - b * d
        uv = _cross(u, v)
        if 0.0 < tN < tD and (_dot(w, uv) == 0.0 or _dot(uv, uv) < 1.0e-20 * a * c):
            if sN < D / 2:
                tN, tD, case = e, c, 2
            else:
                tN, tD, case = e + b, c, 5
    if tN <= 0.0:
        if -d <= 0.0:
            return 0
        if -d >= a:
            return 3
        return 6
    if tN >= tD:
        if -d + b <= 0.0:
            return 1
        if -d + b >= a:
            return 4
        return 7
    return case


def pt_pair(x, p, f):
    """(case, tuple, squared distance) of a point-face pair; case 0 PP, 1 PE, 2 PT."""
    a, b, c = int(f[0]), int(f[1]), int(f[2])
    t = pt_type(x[p], x[a], x[b], x[c])
    if t <= 2:
        q = (a, b, c)[t]
        return 0, (p, q), float(d_pp(x[p], x[q]))
    if t <= 5:
        e0, e1 = ((a, b), (b, c), (c, a))[t - 3]
        return 1, (p, e0, e1), float(d_pe(x[p], x[e0], x[e1]))
    return 2, (p, a, b, c), float(d_pt(x[p], x[a], x[b], x[c]))


def ee_pair(x, e, g):
    """(case, tuple, squared distance) of an edge pair; case 0 PP, 1 PE, 3 EE."""
    i0, i1, j0, j1 = int(e[0]), int(e[1]), int(g[0]), int(g[1])
    t = ee_type(x[i0], x[i1], x[j0], x[j1])
    table = {0: (0, (i0, j0)), 1: (0, (i0, j1)), 2: (1, (i0, j0, j1)), 3: (0, (i1, j0)),
             4: (0, (i1, j1)), 5: (1, (i1, j0, j1)), 6: (1, (j0, i0, i1)), 7: (1, (j1, i0, i1)),
             8: (3, (i0, i1, j0, j1))}
    case, tup = table[t]
    if case == 0:
        d = d_pp(x[tup[0]], x[tup[1]])
    elif case == 1:
        d = d_pe(x[tup[0]], x[tup[1]], x[tup[2]])
    else:
        d = d_ee(x[i0], x[i1], x[j0], x[j1])
    return case, tup, float(d)


def _same_mesh(mesh, ids):
    if mesh is None:
        return False
    m = mesh[ids[0]]
    return m != 0 and all(mesh[i] == m for i in ids[1:])


def proximity(x, faces, surface, edges, dhat, mesh=None):
    """Active pairs by case: {'pp': [(p, q)], 'pe': [(p, e0, e1)], 'pt': [(p, a, b, c)],
    'ee': [(a0, a1, b0, b1)]} (multisets, as lists in loop order)."""
    out = {"pp": [], "pe": [], "pt": [], "ee": []}
    names = ("pp", "pe", "pt", "ee")
    for p in surface:
        p = int(p)
        for f in faces:
            if p in (int(f[0]), int(f[1]), int(f[2])) or _same_mesh(mesh, [p, *map(int, f)]):
                continue
            case, tup, d = pt_pair(x, p, f)
            if d < dhat:
                out[names[case]].append(tup)
    ne = len(edges)
    for i in range(ne):
        for j in range(i + 1, ne):
            e, g = edges[i], edges[j]
            if len({int(e[0]), int(e[1])} & {int(g[0]), int(g[1])}) or \
                    _same_mesh(mesh, [int(e[0]), int(e[1]), int(g[0]), int(g[1])]):
                continue
            case, tup, d = ee_pair(x, e, g)
            if d < dhat:
                out[names[case]].append(tup)
    return out


# ------------------------------------------------------------------ swept boxes
def _boxes(x, d, alpha, prims):
    """Boxes of the primitives' vertices at x and at x - alpha d."""
    y = x - d * alpha
    pts = np.concatenate([x[prims], y[prims]], axis=1)
    return pts.min(axis=1), pts.max(axis=1)


def _overlap(lo_a, hi_a, lo_b, hi_b, gap):
    return not (np.any(lo_b - hi_a >= gap) or np.any(lo_a - hi_b >= gap))


def swept_candidates(x, d, alpha, faces, surface, edges, dhat, mesh=None):
    """Compact candidates as the reference stores them: (-(p + 1), face) for point-face
    pairs and (i, j), i < j, for edge pairs, in loop order."""
    gap = np.sqrt(dhat)
    flo, fhi = _boxes(x, d, alpha, np.asarray(faces, dtype=np.int64))
    plo, phi = _boxes(x, d, alpha, np.asarray(surface, dtype=np.int64)[:, None])
    out = []
    for k, p in enumerate(surface):
        p = int(p)
        for fi, f in enumerate(faces):
            if p in (int(f[0]), int(f[1]), int(f[2])) or _same_mesh(mesh, [p, *map(int, f)]):
                continue
            if _overlap(plo[k], phi[k], flo[fi], fhi[fi], gap):
                out.append((-p - 1, fi))
    elo, ehi = _boxes(x, d, alpha, np.asarray(edges, dtype=np.int64))
    for i in range(len(edges)):
        for j in range(i + 1, len(edges)):
            e, g = edges[i], edges[j]
            if len({int(e[0]), int(e[1])} & {int(g[0]), int(g[1])}) or \
                    _same_mesh(mesh, [int(e[0]), int(e[1]), int(g[0]), int(g[1])]):
                continue
            if _overlap(elo[i], ehi[i], elo[j], ehi[j], gap):
                out.append((i, j))
    return out


# ---------------------------------------------------------------- additive CCD
def pt_distance(p, t0, t1, t2):
    t = pt_type(p, t0, t1, t2)
    if t <= 2:
        return d_pp(p, (t0, t1, t2)[t])
    if t <= 5:
        a, b = ((t0, t1), (t1, t2), (t2, t0))[t - 3]
        return d_pe(p, a, b)
    return d_pt(p, t0, t1, t2)


def ee_distance(a0, a1, b0, b1):
    t = ee_type(a0, a1, b0, b1)
    return [lambda: d_pp(a0, b0), lambda: d_pp(a0, b1), lambda: d_pe(a0, b0, b1),
            lambda: d_pp(a1, b0), lambda: d_pp(a1, b1), lambda: d_pe(a1, b0, b1),
            lambda: d_pe(b0, a0, a1), lambda: d_pe(b1, a0, a1),
            lambda: d_ee(a0, a1, b0, b1)][t]()


def _endpoint_min(a0, a1, b0, b1):
    return min(d_pp(a0, b0), d_pp(a0, b1), d_pp(a1, b0), d_pp(a1, b1))


MAX_ADVANCES = 1 << 20


def accd(xs, ds, eta, edge_edge):
    """Time of impact bound of one candidate: positions xs[4], displacements ds[4] (the
    motion is x + t ds), eta = 1 - slackness, thickness 0. A pair already in contact
    (zero distance) gets 0, and the advancement stops at the time reached if the distance
    reaches zero or after MAX_ADVANCES steps (where the reference would not return)."""
    xs = [np.array(v, dtype=np.float64) for v in xs]
    ds = [np.array(v, dtype=np.float64) for v in ds]
    mov = (ds[0] + ds[1] + ds[2] + ds[3]) * -0.25
    ds = [v + mov for v in ds]
    if edge_edge:
        lmax = np.sqrt(max(_dot(ds[0], ds[0]), _dot(ds[1], ds[1]))) + \
            np.sqrt(max(_dot(ds[2], ds[2]), _dot(ds[3], ds[3])))
    else:
        lmax = np.sqrt(_dot(ds[0], ds[0])) + \
            np.sqrt(max(_dot(ds[1], ds[1]), _dot(ds[2], ds[2]), _dot(ds[3], ds[3])))
    if lmax == 0.0:
        return 1.0

    def dist2(v):
        if edge_edge:
            r = ee_distance(*v)
            return r if r > 0.0 else _endpoint_min(*v)
        return pt_distance(*v)
    d2 = dist2(xs)
    if not d2 > 0.0:
        return 0.0
    d = np.sqrt(d2)
    gap = eta * d2 / d
    toc = 0.0
    for _ in range(MAX_ADVANCES):
        lb = (1.0 - eta) * d2 / (d * lmax)
        xs = [v + w * lb for v, w in zip(xs, ds)]
        d2 = dist2(xs)
        if not d2 > 0.0:                # contact reached: stop at the time accumulated
            break
        d = np.sqrt(d2)
        if toc != 0.0 and d2 / d < gap:
            break
        toc += lb
        if toc > 1.0:
            return 1.0
    return toc


def largest_step(x, d, faces, edges, candidates, slackness):
    """Minimum ACCD time over the candidates (the displacement is -d), 1 without any."""
    eta = 1.0 - slackness
    best = 1.0
    for a, b in candidates:
        if a < 0:
            p = -a - 1
            ids = [p, *map(int, faces[b])]
            t = accd([x[i] for i in ids], [-d[i] for i in ids], eta, False)
        else:
            ids = [*map(int, edges[a]), *map(int, edges[b])]
            t = accd([x[i] for i in ids], [-d[i] for i in ids], eta, True)
        best = min(best, t)
    return best


def swept_candidates_np(x, d, alpha, faces, surface, edges, dhat, mesh=None):
    """swept_candidates with the box tests vectorised (same rule), as a sorted int array
    [candidates, 2]; d may be None (no sweep)."""
    gap = np.sqrt(dhat)
    faces = np.asarray(faces, dtype=np.int64).reshape(-1, 3)
    surface = np.asarray(surface, dtype=np.int64).reshape(-1)
    edges = np.asarray(edges, dtype=np.int64).reshape(-1, 2)
    dd = np.zeros_like(x) if d is None else d
    out = []
    if len(faces) and len(surface):
        flo, fhi = _boxes(x, dd, alpha, faces)
        plo, phi = _boxes(x, dd, alpha, surface[:, None])
        for s in range(0, len(surface), 256):
            lo, hi = plo[s:s + 256, None, :], phi[s:s + 256, None, :]
            ok = ~(np.any(flo[None] - hi >= gap, axis=2) | np.any(lo - fhi[None] >= gap, axis=2))
            p = surface[s:s + 256, None]
            ok &= (faces[None, :, 0] != p) & (faces[None, :, 1] != p) & (faces[None, :, 2] != p)
            if mesh is not None:
                m = mesh[p]
                ok &= ~((m != 0) & (mesh[faces[None, :, 0]] == m) & (mesh[faces[None, :, 1]] == m)
                        & (mesh[faces[None, :, 2]] == m))
            ii, ff = np.nonzero(ok)
            out.append(np.stack([-surface[s + ii] - 1, ff], axis=1))
    if len(edges) > 1:
        elo, ehi = _boxes(x, dd, alpha, edges)
        for s in range(0, len(edges), 256):
            lo, hi = elo[s:s + 256, None, :], ehi[s:s + 256, None, :]
            ok = ~(np.any(elo[None] - hi >= gap, axis=2) | np.any(lo - ehi[None] >= gap, axis=2))
            i = np.arange(s, min(s + 256, len(edges)))[:, None]
            ok &= np.arange(len(edges))[None, :] > i
            a, b = edges[i, 0], edges[i, 1]
            c, e = edges[None, :, 0], edges[None, :, 1]
            ok &= (a != c) & (a != e) & (b != c) & (b != e)
            if mesh is not None:
                m = mesh[a]
                ok &= ~((m != 0) & (mesh[b] == m) & (mesh[c] == m) & (mesh[e] == m))
            ii, jj = np.nonzero(ok)
            out.append(np.stack([s + ii, jj], axis=1))
    if not out:
        return np.zeros((0, 2), dtype=np.int64)
    r = np.concatenate(out)
    return r[np.lexsort((r[:, 1], r[:, 0]))]


def active_from_candidates(x, faces, edges, candidates, dhat):
    """The narrow phase over given candidates: active pairs by case, as proximity()."""
    out = {"pp": [], "pe": [], "pt": [], "ee": []}
    names = ("pp", "pe", "pt", "ee")
    for a, b in candidates:
        a, b = int(a), int(b)
        if a < 0:
            case, tup, dist = pt_pair(x, -a - 1, faces[b])
        else:
            case, tup, dist = ee_pair(x, edges[a], edges[b])
        if dist < dhat:
            out[names[case]].append(tup)
    return out


def filter_candidates(x, d, alpha, faces, edges, candidates, dhat):
    """The candidates whose boxes (swept to x - alpha d when d is given) pass the gap test:
    from a superset built with larger boxes, exactly the set swept_candidates_np gives."""
    gap = np.sqrt(dhat)
    faces = np.asarray(faces, dtype=np.int64).reshape(-1, 3)
    edges = np.asarray(edges, dtype=np.int64).reshape(-1, 2)
    c = np.asarray(candidates, dtype=np.int64).reshape(-1, 2)
    dd = np.zeros_like(x) if d is None else d
    keep = np.zeros(len(c), dtype=bool)
    pf = c[:, 0] < 0
    if pf.any():
        alo, ahi = _boxes(x, dd, alpha, (-c[pf, 0] - 1)[:, None])
        blo, bhi = _boxes(x, dd, alpha, faces[c[pf, 1]])
        keep[pf] = ~(np.any(blo - ahi >= gap, axis=1) | np.any(alo - bhi >= gap, axis=1))
    if (~pf).any():
        alo, ahi = _boxes(x, dd, alpha, edges[c[~pf, 0]])
        blo, bhi = _boxes(x, dd, alpha, edges[c[~pf, 1]])
        keep[~pf] = ~(np.any(blo - ahi >= gap, axis=1) | np.any(alo - bhi >= gap, axis=1))
    return c[keep]
