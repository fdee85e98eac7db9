"""The contact-set oracle (oracle/ccd.py) against independent mathematics: closest points
on triangles and segments computed another way (barycentric regions of the triangle, the
convex quadratic of two segment parameters minimised on its box), closed forms of the
additive CCD on head-on approaches (t = slackness * D / v), its two guarantees on random
motions (no contact before the returned time, distance there at least (1 - slackness) of
the initial one), and conservativeness of the swept candidates (every pair that comes
within dhat anywhere along the sweep is a candidate)."""
import numpy as np
import pytest

from oracle import ccd

RNG = np.random.default_rng(331)


def _closest_on_triangle(p, a, b, c):
    """Closest point of triangle abc to p by barycentric regions (vertex, edge and face
    regions tested in turn) and the feature it lies on: ('v', i), ('e', (i, j)), ('f',)."""
    ab, ac, ap = b - a, c - a, p - a
    d1, d2 = ab @ ap, ac @ ap
    if d1 <= 0 and d2 <= 0:
        return a, ("v", 0)
    bp = p - b
    d3, d4 = ab @ bp, ac @ bp
    if d3 >= 0 and d4 <= d3:
        return b, ("v", 1)
    vc = d1 * d4 - d3 * d2
    if vc <= 0 and d1 >= 0 and d3 <= 0:
        return a + d1 / (d1 - d3) * ab, ("e", (0, 1))
    cp = p - c
    d5, d6 = ab @ cp, ac @ cp
    if d6 >= 0 and d5 <= d6:
        return c, ("v", 2)
    vb = d5 * d2 - d1 * d6
    if vb <= 0 and d2 >= 0 and d6 <= 0:
        return a + d2 / (d2 - d6) * ac, ("e", (0, 2))
    va = d3 * d6 - d5 * d4
    if va <= 0 and (d4 - d3) >= 0 and (d5 - d6) >= 0:
        return b + (d4 - d3) / ((d4 - d3) + (d5 - d6)) * (c - b), ("e", (1, 2))
    den = 1.0 / (va + vb + vc)
    return a + ab * (vb * den) + ac * (vc * den), ("f",)


def _segment_distance2(p0, p1, q0, q1):
    """min over s, t in [0, 1] of |p(s) - q(t)|^2: the interior stationary point if it lies
    in the box, else the minimum over the four sides (each a clamped 1-D quadratic)."""
    u, v, w = p1 - p0, q1 - q0, p0 - q0
    a, b, c, d, e = u @ u, u @ v, v @ v, u @ w, v @ w

    def f(s, t):
        r = w + s * u - t * v
        return r @ r
    cands = []
    den = a * c - b * b
    if den > 1e-14 * a * c:
        s = (b * e - c * d) / den
        t = (a * e - b * d) / den
        if 0 <= s <= 1 and 0 <= t <= 1:
            cands.append(f(s, t))
    for s in (0.0, 1.0):
        t = np.clip((e + s * b) / c, 0, 1)
        cands.append(f(s, t))
    for t in (0.0, 1.0):
        s = np.clip((t * b - d) / a, 0, 1)
        cands.append(f(s, t))
    return min(cands)


def test_distances_closed_forms():
    t0, t1, t2 = np.array([0., 0, 0]), np.array([2., 0, 0]), np.array([0., 2, 0])
    assert ccd.d_pt(np.array([0.3, 0.4, 1.5]), t0, t1, t2) == pytest.approx(2.25, rel=1e-15)
    assert ccd.d_pe(np.array([1.0, -3.0, 4.0]), t0, t1) == pytest.approx(25.0, rel=1e-15)
    assert ccd.d_pp(np.array([1., 2, 2]), t0) == 9.0
    # perpendicular lines x-axis and a line parallel to y at height 3
    assert ccd.d_ee(t0, t1, np.array([0.5, -1, 3]), np.array([0.5, 1, 3])) == pytest.approx(9.0)


def test_point_triangle_type_is_the_closest_feature():
    names = {0: ("v", 0), 1: ("v", 1), 2: ("v", 2), 3: ("e", (0, 1)), 4: ("e", (1, 2)),
             5: ("e", (0, 2)), 6: ("f",)}
    seen = set()
    for _ in range(3000):
        tri = RNG.normal(size=(3, 3))
        p = RNG.normal(size=3) * 1.5
        q, feat = _closest_on_triangle(p, *tri)
        t = ccd.pt_type(p, *tri)
        seen.add(t)
        assert names[t] == feat
        assert ccd.pt_distance(p, *tri) == pytest.approx(float((p - q) @ (p - q)), rel=1e-9, abs=1e-14)
    assert seen == set(range(7))


def test_edge_edge_distance_is_the_segment_distance():
    seen = set()
    for _ in range(3000):
        v = RNG.normal(size=(4, 3))
        t = ccd.ee_type(*v)
        seen.add(t)
        assert ccd.ee_distance(*v) == pytest.approx(_segment_distance2(*v), rel=1e-8, abs=1e-14)
    assert seen == set(range(9))


@pytest.mark.parametrize("slack", [0.8, 0.5, 0.9])
def test_accd_head_on_closed_form(slack):
    """A point falling straight onto a triangle's interior and an edge falling onto a
    perpendicular edge: the advancement is exact, so t = slackness * D / v."""
    tri = [np.array([0., 0, 0]), np.array([1., 0, 0]), np.array([0., 1, 0])]
    D, v = 0.3, 0.7
    p = np.array([0.2, 0.3, D])
    t = ccd.accd([p, *tri], [np.array([0, 0, -v]), *(np.zeros(3),) * 3], 1 - slack, False)
    assert t == pytest.approx(slack * D / v, rel=1e-12)
    e0, e1 = np.array([-1., 0.1, 0]), np.array([1., 0.1, 0])
    f0, f1 = np.array([0.2, -1, D]), np.array([0.2, 1, D])
    t = ccd.accd([e0, e1, f0, f1], [np.zeros(3), np.zeros(3), np.array([0, 0, -v]),
                                    np.array([0, 0, -v])], 1 - slack, True)
    assert t == pytest.approx(slack * D / v, rel=1e-12)
    # far apart: the bound is 1
    assert ccd.accd([p + [0, 0, 10.0], *tri], [np.array([0, 0, -v]), *(np.zeros(3),) * 3],
                    1 - slack, False) == 1.0


def _dist(xs, edge_edge):
    return ccd.ee_distance(*xs) if edge_edge else ccd.pt_distance(*xs)


def test_accd_guarantees_on_random_motions():
    slack = 0.8
    hits = 0
    for k in range(400):
        ee = bool(k % 2)
        xs = [RNG.normal(size=3) for _ in range(4)]
        ds = [RNG.normal(size=3) * 2.0 for _ in range(4)]
        d0 = _dist(xs, ee)
        t = ccd.accd(xs, ds, 1 - slack, ee)
        assert 0.0 < t <= 1.0
        ts = np.linspace(0.0, t, 200)
        dmin = min(_dist([a + s * b for a, b in zip(xs, ds)], ee) for s in ts)
        assert dmin > 0.0
        if t < 1.0:
            hits += 1
            dt = _dist([a + t * b for a, b in zip(xs, ds)], ee)
            assert np.sqrt(dt) >= (1 - slack) * np.sqrt(d0) * (1 - 1e-9)
    assert hits > 50


def _mesh():
    """Two small triangulated patches facing each other (a square of 2 triangles over a
    square of 8 triangles), all vertices on the surface."""
    g = np.array([[i, j] for j in range(3) for i in range(3)], dtype=np.float64) * 0.5
    lower = np.concatenate([g, np.zeros((9, 1))], axis=1)
    faces = []
    for j in range(2):
        for i in range(2):
            a = 3 * j + i
            faces += [[a, a + 1, a + 4], [a, a + 4, a + 3]]
    upper = np.array([[0.1, 0.1, 0.3], [0.9, 0.1, 0.3], [0.9, 0.9, 0.32], [0.1, 0.9, 0.31]])
    x = np.concatenate([lower, upper])
    faces += [[9, 10, 11], [9, 11, 12]]
    faces = np.array(faces)
    edges = sorted({tuple(sorted((int(f[a]), int(f[b])))) for f in faces
                    for a, b in ((0, 1), (1, 2), (2, 0))})
    return x, faces, np.arange(13), np.array(edges)


def test_proximity_matches_true_closest_distances():
    x, faces, surf, edges = _mesh()
    x = x + RNG.normal(size=x.shape) * 0.02
    dhat = 0.35 ** 2
    out = ccd.proximity(x, faces, surf, edges, dhat)
    # every point-face pair: emitted iff its true closest distance is below sqrt(dhat)
    n_pf = 0
    for p in surf:
        for f in faces:
            if p in f:
                continue
            q, _ = _closest_on_triangle(x[p], *x[f])
            n_pf += float((x[p] - q) @ (x[p] - q)) < dhat
    assert n_pf > 0
    n_ee = sum(1 for i in range(len(edges)) for j in range(i + 1, len(edges))
               if not set(edges[i]) & set(edges[j]) and _segment_distance2(*x[edges[i]], *x[edges[j]]) < dhat)
    assert n_ee > 0 and n_pf + n_ee == sum(len(v) for v in out.values())
    # a nonzero mesh id shared by all four vertices removes the pair
    mesh = np.ones(len(x), dtype=np.uint32)
    assert sum(len(v) for v in ccd.proximity(x, faces, surf, edges, dhat, mesh).values()) == 0


def test_swept_candidates_cover_the_sweep():
    x, faces, surf, edges = _mesh()
    d = RNG.normal(size=x.shape) * 0.1
    d[:9] = 0.0
    d[9:] = [0.0, 0.0, 0.4]                  # the upper patch moves down (x - t d)
    dhat = 0.05 ** 2
    alpha = 1.0
    cands = set(ccd.swept_candidates(x, d, alpha, faces, surf, edges, dhat))
    for t in np.linspace(0.0, alpha, 11):
        y = x - t * d
        for p in surf:
            for fi, f in enumerate(faces):
                if p in f:
                    continue
                q, _ = _closest_on_triangle(y[p], *y[f])
                if float((y[p] - q) @ (y[p] - q)) < dhat:
                    assert (-int(p) - 1, fi) in cands
        for i in range(len(edges)):
            for j in range(i + 1, len(edges)):
                if set(edges[i]) & set(edges[j]):
                    continue
                if _segment_distance2(*y[edges[i]], *y[edges[j]]) < dhat:
                    assert (i, j) in cands
    assert cands


def test_vectorised_candidates_equal_the_loop_restatement():
    x, faces, surf, edges = _mesh()
    d = np.zeros_like(x)
    d[9:] = [0.05, -0.02, 0.35]
    d[9:] += RNG.normal(size=(4, 3)) * 0.02
    mesh = np.r_[np.ones(9), 2 * np.ones(4)].astype(np.uint32)
    for m in (None, mesh):
        loop = np.array(sorted(ccd.swept_candidates(x, d, 0.8, faces, surf, edges, 0.01, m)))
        vec = ccd.swept_candidates_np(x, d, 0.8, faces, surf, edges, 0.01, m)
        np.testing.assert_array_equal(loop.reshape(-1, 2), vec)
        assert len(vec) > 0
    cross = ccd.swept_candidates_np(x, d, 0.8, faces, surf, edges, 0.01, mesh)
    assert len(cross) < len(ccd.swept_candidates_np(x, d, 0.8, faces, surf, edges, 0.01))
