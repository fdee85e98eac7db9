"""The contact module on the GPU (yasps_amd.ccd.CCD over csrc/ccd.hip) against the oracle
(oracle/ccd.py): proximity pairs by distance type as bit-exact multisets of vertex-id
tuples, swept candidates as exact sets (a small scene and one of about 6k faces and 9k
edges against the vectorised box rule), per-candidate additive-CCD times and the step
bound (1e-12 relative), mesh ids, buffer growth and the capacity error."""
import numpy as np
import pytest

from oracle import ccd as occd

pytestmark = pytest.mark.gpu


def _grid(n, z, jitter, rng, shift=(0.0, 0.0), size=1.0):
    g = np.array([[i, j] for j in range(n) for i in range(n)], dtype=np.float64) * (size / (n - 1))
    x = np.concatenate([g + shift, np.full((n * n, 1), z)], axis=1)
    x += rng.normal(size=x.shape) * jitter
    f = []
    for j in range(n - 1):
        for i in range(n - 1):
            a = n * j + i
            f += [[a, a + 1, a + n + 1], [a, a + n + 1, a + n]]
    return x, np.array(f, dtype=np.int64)


def _scene(n_low, n_up, h, jitter, seed):
    """A cloth-like lower grid and a smaller grid hovering h above it."""
    rng = np.random.default_rng(seed)
    xl, fl = _grid(n_low, 0.0, jitter, rng)
    xu, fu = _grid(n_up, h, jitter, rng, shift=(0.21, 0.17), size=0.6)
    x = np.concatenate([xl, xu])
    faces = np.concatenate([fl, fu + len(xl)])
    edges = np.array(sorted({tuple(sorted((int(t[a]), int(t[b])))) for t in faces
                             for a, b in ((0, 1), (1, 2), (2, 0))}), dtype=np.int64)
    mesh = np.r_[np.ones(len(xl)), 2 * np.ones(len(xu))].astype(np.uint32)
    return x, faces, edges, mesh


def _ccd(x, faces, edges, mesh=None, **kw):
    from yasps_amd import device as dev
    from yasps_amd.ccd import CCD
    nv = len(x)
    c = CCD(nv, nv, mesh_indices=mesh, print_timings=False, **kw)
    xd = dev.to_device(x.reshape(-1), "float64")
    c.init_faces(xd, faces.reshape(-1), np.arange(nv), len(faces))
    c.init_edges(xd, xd, edges.reshape(-1), len(edges))
    return c, xd


def _tuples(flat, w):
    from yasps_amd import device as dev
    a = dev.to_host(flat).astype(np.int64).reshape(-1, w)
    return sorted(map(tuple, a))


@pytest.mark.parametrize("use_mesh", [False, True])
def test_proximity_pairs_match_oracle(use_mesh):
    print(f"ccd proximity mesh={use_mesh}", flush=True)
    x, faces, edges, mesh = _scene(9, 5, 0.03, 0.008, seed=11)
    mesh[:40] = 2                  # half the lower grid shares the upper grid's id
    m = mesh if use_mesh else None
    dhat = 0.045 ** 2
    c, xd = _ccd(x, faces, edges, m)
    counts = c.cd(xd, dhat)
    ref = occd.proximity(x, faces, np.arange(len(x)), edges, dhat, m)
    assert counts == [len(ref[k]) for k in ("pp", "pe", "pt", "ee")] == c.separated_counts
    assert min(counts) > 0
    for k, w in (("pp", 2), ("pe", 3), ("pt", 4), ("ee", 4)):
        assert _tuples(getattr(c, k), w) == sorted(ref[k])
    assert c.case_storage.numel() == 2 * counts[0] + 3 * counts[1] + 4 * counts[2] + 4 * counts[3]


def test_swept_candidates_and_step_bound_match_oracle():
    from yasps_amd import device as dev
    print("ccd sweep and step bound", flush=True)
    x, faces, edges, mesh = _scene(9, 5, 0.03, 0.008, seed=12)
    rng = np.random.default_rng(5)
    d = np.zeros_like(x)
    nl = 81
    d[nl:] = [0.01, -0.004, 0.06]                # the upper grid falls through: x - t d
    d += rng.normal(size=d.shape) * 0.002
    dhat = 0.01 ** 2
    c, xd = _ccd(x, faces, edges)
    dd = dev.to_device(d.reshape(-1), "float64")
    n = c.ccd(xd, dhat, dd, 1.0)
    got = dev.to_host(c.ccd_candidates).astype(np.int64).reshape(-1, 2)
    got = got[np.lexsort((got[:, 1], got[:, 0]))]
    ref = occd.swept_candidates_np(x, d, 1.0, faces, np.arange(len(x)), edges, dhat)
    assert n == len(ref) > 0
    np.testing.assert_array_equal(got, ref)
    # per-candidate times and the bound
    times = dev.to_host(c.candidate_times(0.8))
    cand = dev.to_host(c.ccd_candidates).astype(np.int64).reshape(-1, 2)
    for (a, b), t in zip(cand, times):
        if a < 0:
            ids = [-a - 1, *faces[b]]
            r = occd.accd([x[i] for i in ids], [-d[i] for i in ids], 0.2, False)
        else:
            ids = [*edges[a], *edges[b]]
            r = occd.accd([x[i] for i in ids], [-d[i] for i in ids], 0.2, True)
        assert t == pytest.approx(r, rel=1e-12, abs=0), (a, b)
    step = c.compute_largest_step_size(0.8, xd, dd)
    assert step == pytest.approx(min(1.0, float(times.min())), rel=0, abs=0)
    assert 0.0 < step < 1.0
    # no contact at the bounded step: every pair keeps a positive distance
    y = x - step * d
    near = occd.proximity(y, faces, np.arange(len(x)), edges, 1e-9)
    assert sum(len(v) for v in near.values()) == 0
    # the cached candidates hold every pair below dhat along the sweep
    from yasps_amd.ccd import CCD  # noqa: F401
    yd = dev.to_device((x - 0.5 * step * d).reshape(-1), "float64")
    counts = c.cd_from_cached_ccd(yd, dhat, 0.5 * step)
    ref = occd.proximity(x - 0.5 * step * d, faces, np.arange(len(x)), edges, dhat)
    assert counts == [len(ref[k]) for k in ("pp", "pe", "pt", "ee")]


def test_large_scene_candidates_match_bruteforce():
    """About 6k faces and 9k edges: the BVH queries give exactly the brute-force set."""
    from yasps_amd import device as dev
    print("ccd large scene", flush=True)
    x, faces, edges, mesh = _scene(48, 30, 0.004, 0.0015, seed=13)
    d = np.zeros_like(x)
    d[48 * 48:] = [0.0, 0.0, 0.01]
    dhat = 0.003 ** 2
    c, xd = _ccd(x, faces, edges, mesh)
    dd = dev.to_device(d.reshape(-1), "float64")
    n = c.ccd(xd, dhat, dd, 1.0)
    got = dev.to_host(c.ccd_candidates).astype(np.int64).reshape(-1, 2)
    got = got[np.lexsort((got[:, 1], got[:, 0]))]
    ref = occd.swept_candidates_np(x, d, 1.0, faces, np.arange(len(x)), edges, dhat, mesh)
    assert n == len(ref) > 10000
    np.testing.assert_array_equal(got, ref)
    # proximity at the start against the narrow phase of the brute-force proximity set (the
    # swept candidates whose static boxes pass the gap test)
    counts = c.cd(xd, dhat)
    prox = occd.filter_candidates(x, None, 0.0, faces, edges, ref, dhat)
    ref = occd.active_from_candidates(x, faces, edges, prox, dhat)
    assert counts == [len(ref[k]) for k in ("pp", "pe", "pt", "ee")]
    for k, w in (("pp", 2), ("pe", 3), ("pt", 4), ("ee", 4)):
        assert _tuples(getattr(c, k), w) == sorted(ref[k])


def test_capacity_growth_and_overflow():
    from yasps_amd import device as dev
    x, faces, edges, _ = _scene(20, 12, 0.004, 0.0015, seed=14)
    d = np.zeros_like(x)
    d[400:] = [0.0, 0.0, 0.01]
    dd = dev.to_device(d.reshape(-1), "float64")
    c, xd = _ccd(x, faces, edges)
    n = c.ccd(xd, 0.003 ** 2, dd, 1.0)
    assert n > 1024                       # grew past the first buffer
    small, xs = _ccd(x, faces, edges, max_ccd_pairs=100)
    with pytest.raises(OverflowError):
        small.ccd(xs, 0.003 ** 2, dd, 1.0)
    with pytest.raises(RuntimeError):
        small.compute_largest_step_size(0.8, xs, dd)
