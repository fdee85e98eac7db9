"""CPU checks of the MAS hierarchy, the device tables built from it, and the MAS oracle:
partition invariants, the Galerkin identity R A R^T, and an emulation of the device
apply (table semantics of the restrict / coarsen / scatter / local-solve / prolong
kernels, FP64) against the dense oracle apply."""
import numpy as np
import pytest

from oracle import mas as omas, newton, pcg
from yasps_amd import mas_hierarchy as mh
from yasps_amd.mas import DOMAIN_SCALARS, NO_DEST, build_layout, level_weights, static_pairs
from yasps_amd.scenes import cube_scene


@pytest.fixture(scope="module")
def cube():
    sc = cube_scene(n=4, noise=0.05, seed=3)
    st = newton.fast_structure(sc)
    asm = newton.evaluate(sc, st, 2)
    A = newton.bsr_matrix(st, asm.values)
    return sc, A, asm.gradient


def _hierarchy(A, **kw):
    nb = A.shape[0] // 3
    src = np.zeros(A.indices.size, dtype=np.uint32)
    pairs = static_pairs(A.indptr, A.indices, src)
    kw = {"max_domain_dofs": 32, "max_levels": 4, "target_nodes_per_partition": 8} | kw
    return mh.build_hierarchy(np.full(nb, 3), pairs, **kw)


def test_graph_from_pairs_is_symmetric_without_loops():
    xadj, adj = mh.graph_from_pairs(4, np.array([[0, 1], [1, 0], [2, 2], [3, 1], [0, 1]]))
    nbrs = [sorted(adj[xadj[i]:xadj[i + 1]].tolist()) for i in range(4)]
    assert nbrs == [[1], [0, 3], [], [1]]


def test_bfs_groups_are_connected_and_bounded():
    # 6 x 6 grid graph
    n = 36
    pairs = [(i, i + 1) for i in range(n) if (i + 1) % 6] + [(i, i + 6) for i in range(n - 6)]
    xadj, adj = mh.graph_from_pairs(n, np.array(pairs))
    groups = mh.bfs_groups(xadj, adj, 5)
    assert sorted(v for g in groups for v in g) == list(range(n))
    for g in groups:
        assert 1 <= len(g) <= 5
        # connected: BFS inside the group reaches every member
        seen, stack, members = {g[0]}, [g[0]], set(g)
        while stack:
            v = stack.pop()
            for u in adj[xadj[v]:xadj[v + 1]]:
                if u in members and u not in seen:
                    seen.add(int(u))
                    stack.append(int(u))
        assert seen == members


def test_pack_domain_groups_never_crosses_groups_or_capacity():
    dims = np.array([3, 3, 3, 12, 3, 40, 3])
    doms = mh.pack_domain_groups([[0, 1, 2, 3], [4, 5, 6]], dims, 12)
    assert doms == [[0, 1, 2], [3], [4], [5], [6]]


def test_hierarchy_partition_invariants(cube):
    _, A, _ = cube
    h = _hierarchy(A)
    assert h.num_levels >= 2
    for l, lv in enumerate(h.levels):
        nodes = np.concatenate(lv.domains)
        assert np.array_equal(np.sort(nodes), np.arange(lv.num_nodes))      # a partition
        assert all(lv.dims[d].sum() <= 32 for d in lv.domains)
        if lv.fine_to_parent is not None:
            m = lv.fine_to_parent
            # merged nodes share a domain and a dimension
            dom = np.empty(lv.num_nodes, dtype=np.int64)
            for k, d in enumerate(lv.domains):
                dom[d] = k
            for J in range(m.max() + 1):
                kids = np.flatnonzero(m == J)
                assert np.unique(dom[kids]).size == 1
                assert np.unique(lv.dims[kids]).size == 1
            assert np.array_equal(h.composed[l + 1], m[h.composed[l]])


def test_galerkin_levels_are_restrictions(cube):
    _, A, _ = cube
    h = _hierarchy(A)
    Ad = A.toarray()
    mats = omas.level_matrices(h, Ad)
    n = Ad.shape[0]
    for l in range(h.num_levels):
        m = omas.scalar_map(h, l)
        R = np.zeros((h.levels[l].num_dofs, n))
        R[m, np.arange(n)] = 1.0
        # brute force: entry (a, b) of level l sums every fine entry mapping onto it
        B = np.zeros_like(mats[l])
        np.add.at(B, (m[:, None], m[None, :]), Ad)
        assert np.allclose(B, mats[l], rtol=0, atol=1e-9 * np.abs(Ad).max())
        assert np.allclose(mats[l], mats[l].T)


def test_mas_operator_is_spd_and_inverts_domains(cube):
    _, A, _ = cube
    h = _hierarchy(A)
    mats = omas.level_matrices(h, A.toarray())
    invs = omas.domain_inverses(h, mats)
    for l, lv in enumerate(h.levels):
        for d, inv in zip(lv.domains, invs[l]):
            s = omas.domain_scalars(lv, d)
            assert np.allclose(inv @ mats[l][np.ix_(s, s)], np.eye(s.size), atol=1e-9)
    M = omas.dense_operator(h, invs, omas.level_weights(h))
    assert np.allclose(M, M.T, atol=1e-12 * np.abs(M).max())
    assert np.linalg.eigvalsh(0.5 * (M + M.T)).min() > 0


def emulate_device_apply(layout, block_values, r, weights, duplicate_weight=1.0):
    """FP64 restatement of the device kernels over the layout tables."""
    D = layout.domain_scalars
    b0 = layout.blocks[0]
    if b0.coarsen_offsets is not None:          # repeated (row, col) entries summed first
        block_values = np.stack([block_values[b0.coarsen_items[b0.coarsen_offsets[e]:
                                                               b0.coarsen_offsets[e + 1]]].sum(0)
                                 for e in range(b0.nnz)])
    vals = [block_values]
    for b in layout.blocks[1:]:
        prev = vals[-1]
        out = np.zeros((b.nnz, 3, 3))
        for e in range(b.nnz):
            out[e] = prev[b.coarsen_items[b.coarsen_offsets[e]:b.coarsen_offsets[e + 1]]].sum(0)
        vals.append(out)
    dom = np.zeros((layout.num_domains * D * D,))
    for b, v in zip(layout.blocks, vals):
        for e in np.flatnonzero(b.dest != NO_DEST):
            base = int(b.dest[e])
            for i in range(3):
                dom[base + i * D: base + i * D + 3] = v[e, i]
    dom = dom.reshape(-1, D, D)
    for k, P in enumerate(layout.domain_sizes):
        dom[k, np.arange(P, D), np.arange(P, D)] = 1.0
    inv = np.linalg.inv(dom)
    nc = int(layout.level_offsets[-1])
    rc = np.zeros((nc, 3))
    r3 = r.reshape(-1, 3)
    for l in range(layout.num_coarse_levels):
        for J in range(layout.level_offsets[l], layout.level_offsets[l + 1]):
            kids = layout.restrict_items[layout.restrict_offsets[J]:layout.restrict_offsets[J + 1]]
            rc[J] = r3[kids].sum(0) if l < 3 else rc[kids].sum(0)
    c0 = np.zeros_like(r3)
    cc = np.zeros((nc, 3))
    for d in range(layout.num_domains):
        nodes = layout.domain_nodes[layout.domain_node_offsets[d]:layout.domain_node_offsets[d + 1]]
        src, dst = (r3, c0) if d < layout.level0_domains else (rc, cc)
        loc = np.zeros(D)
        loc[:3 * nodes.size] = src[nodes].reshape(-1)
        scale = duplicate_weight if layout.duplicate[d] else 1.0
        dst[nodes] = scale * (inv[d] @ loc)[:3 * nodes.size].reshape(-1, 3)
    z = weights[0] * c0
    for l in range(layout.num_coarse_levels):
        z = z + weights[l + 1] * cc[layout.prolong_maps[l]]
    return z.reshape(-1)


@pytest.mark.parametrize("levels,coarsest,target",
                         [(6, 2.5, 3), (4, 2.5, 8), (2, 1.0, 8), (1, 2.5, 8)])
def test_device_tables_reproduce_the_oracle_apply(cube, levels, coarsest, target):
    _, A, _ = cube
    h = _hierarchy(A, max_levels=levels, target_nodes_per_partition=target)
    assert h.num_levels == levels
    layout = build_layout(h, A.indptr, A.indices)
    w = level_weights(h.num_levels, coarsest).astype(np.float64)
    assert np.array_equal(w, np.asarray(omas.level_weights(h, coarsest)))
    mats = omas.level_matrices(h, A.toarray())
    invs = omas.domain_inverses(h, mats)
    rng = np.random.default_rng(7)
    for _ in range(3):
        r = rng.standard_normal(A.shape[0])
        ref = omas.apply(h, invs, r, w)
        got = emulate_device_apply(layout, A.data, r, w)
        assert np.abs(got - ref).max() <= 1e-10 * np.abs(ref).max()


def test_deep_levels_restrict_from_level_three(cube):
    """Nodes of level 4 and deeper list level-3 coarse ids, whose level-0 descendants are
    exactly their own (the hierarchy is nested), so the two-stage sum equals the flat one."""
    _, A, _ = cube
    h = _hierarchy(A, max_levels=6, target_nodes_per_partition=3)
    assert h.num_levels >= 5
    layout = build_layout(h, A.indptr, A.indices)
    n2, n3 = layout.level_offsets[2], layout.level_offsets[3]
    off, items = layout.restrict_offsets, layout.restrict_items
    assert np.all(items[:off[n3]] < A.shape[0] // 3)
    deep = items[off[n3]:]
    assert deep.size and np.all((n2 <= deep) & (deep < n3))
    for l in range(4, h.num_levels):
        for J in range(layout.level_offsets[l - 1], layout.level_offsets[l]):
            kids = items[off[J]:off[J + 1]] - n2
            flat = np.flatnonzero(h.composed[l] == J - layout.level_offsets[l - 1])
            assert np.array_equal(np.sort(np.flatnonzero(np.isin(h.composed[3], kids))), flat)


def test_repeated_entries_are_summed(cube):
    """An operator whose block-CSR repeats (row, col) pairs (a block held by both the
    static and the dynamic part) gives the MAS of the summed operator: every block of the
    cube is entered twice, as 0.3 B and 0.7 B."""
    _, A, _ = cube
    h = _hierarchy(A)
    nnz = A.indices.size
    rows = np.repeat(np.arange(A.shape[0] // 3), np.diff(A.indptr))
    order = np.argsort(np.concatenate([rows, rows]), kind="stable")
    col2 = np.concatenate([A.indices, A.indices])[order]
    data2 = np.concatenate([0.3 * A.data, 0.7 * A.data])[order]
    indptr2 = 2 * A.indptr
    layout = build_layout(h, indptr2, col2)
    assert layout.blocks[0].coarsen_offsets is not None and layout.blocks[0].nnz == nnz
    w = level_weights(h.num_levels, 2.5).astype(np.float64)
    invs = omas.domain_inverses(h, omas.level_matrices(h, A.toarray()))
    r = np.random.default_rng(11).standard_normal(A.shape[0])
    ref = omas.apply(h, invs, r, w)
    got = emulate_device_apply(layout, data2, r, w)
    assert np.abs(got - ref).max() <= 1e-10 * np.abs(ref).max()


def test_layout_destinations_stay_inside_their_domain(cube):
    _, A, _ = cube
    layout = build_layout(_hierarchy(A), A.indptr, A.indices)
    D = layout.domain_scalars
    for b in layout.blocks:
        d = b.dest[b.dest != NO_DEST].astype(np.int64)
        dom = d // (D * D)
        row, col = (d % (D * D)) // D, d % D
        size = layout.domain_sizes[dom]
        assert np.all(row + 2 < size) and np.all(col + 2 < size)
        assert np.all(row % 3 == 0) and np.all(col % 3 == 0)
    assert layout.domain_sizes.max() <= D


def test_level_weight_rules():
    assert level_weights(3).tolist() == [1.0, 1.0, 2.5]
    assert level_weights(1).tolist() == [2.5]
    assert level_weights(2, weights=(1.0, 2.0, 3.0)).tolist() == [1.0, 3.0]
    with pytest.raises(ValueError):
        level_weights(3, weights=(1.0,))


def test_mas_pcg_oracle_converges(cube):
    _, A, b = cube
    h = _hierarchy(A)
    mats = omas.level_matrices(h, A.toarray())
    invs = omas.domain_inverses(h, mats)
    w = omas.level_weights(h, 1.0)
    st_m, x_m, it_m, _ = pcg.pcg(lambda v: A @ v, lambda v: omas.apply(h, invs, v, w), b, 1e-12, 500)
    assert st_m == it_m > 0                         # reference code: iterations on success
    assert np.linalg.norm(A @ x_m - b) <= 1e-5 * np.linalg.norm(b)


@pytest.mark.parametrize("levels", [1, 2])
def test_mas_single_domain_closed_form(levels):
    """n=1 cube (4 moving vertices, 12 scalars): one level-0 domain holding everything
    gives z = w A^-1 r with w the coarsest weight (one level), and with a second level the
    four nodes collapse into one coarse node: z = A^-1 r + w R^T (R A R^T)^-1 R r with
    R = [I I I I]. A mis-weighted level or a wrong restriction fails this."""
    sc = cube_scene(n=1, noise=0.05, seed=2)
    st = newton.fast_structure(sc)
    asm = newton.evaluate(sc, st, 2)
    A = newton.bsr_matrix(st, asm.values)
    h = _hierarchy(A, max_levels=levels)
    assert h.num_levels == levels and len(h.levels[0].domains) == 1
    Ad = A.toarray()
    w = level_weights(h.num_levels, 2.5).astype(np.float64)
    r = np.random.default_rng(3).standard_normal(Ad.shape[0])
    if levels == 1:
        expect = 2.5 * np.linalg.solve(Ad, r)
    else:
        R = np.tile(np.eye(3), (1, Ad.shape[0] // 3))
        expect = np.linalg.solve(Ad, r) + 2.5 * R.T @ np.linalg.solve(R @ Ad @ R.T, R @ r)
    invs = omas.domain_inverses(h, omas.level_matrices(h, Ad))
    got = omas.apply(h, invs, r, w)
    assert np.abs(got - expect).max() <= 1e-10 * np.abs(expect).max()
    layout = build_layout(h, A.indptr, A.indices)
    emu = emulate_device_apply(layout, A.data, r, w)
    assert np.abs(emu - expect).max() <= 1e-10 * np.abs(expect).max()


def test_duplicate_domain_rule_on_a_built_hierarchy():
    """Level 0: domains {0, 1} and {2}; level 1: node 0 = {0, 1}, node 1 = {2}, each its
    own domain. Level-1 domain {1} repeats level-0 domain {2} (one child, the whole
    domain); {0} has two children. Level 2 merges both: not a duplicate."""
    L0 = mh.Level(np.full(3, 3), mh.offsets_of(np.full(3, 3)), np.zeros(4, np.int64),
                  np.zeros(0, np.int64), [np.array([0, 1]), np.array([2])],
                  fine_to_parent=np.array([0, 0, 1]))
    L1 = mh.Level(np.full(2, 3), mh.offsets_of(np.full(2, 3)), np.zeros(3, np.int64),
                  np.zeros(0, np.int64), [np.array([0]), np.array([1])],
                  fine_to_parent=np.array([0, 0]))
    L2 = mh.Level(np.full(1, 3), mh.offsets_of(np.full(1, 3)), np.zeros(2, np.int64),
                  np.zeros(0, np.int64), [np.array([0])])
    h = mh.Hierarchy([L0, L1, L2], [np.arange(3), np.array([0, 0, 1]), np.zeros(3, np.int64)])
    from yasps_amd.mas import duplicate_domains
    expect = [[False, False], [False, True], [False]]
    assert [list(map(bool, d)) for d in duplicate_domains(h)] == expect
    assert omas.duplicate_domains(h) == expect


@pytest.mark.parametrize("levels,target", [(6, 3), (4, 8)])
def test_duplicate_rule_matches_oracle_and_weights_apply(cube, levels, target):
    """The product's duplicate flags equal the oracle's restatement on cube hierarchies,
    and the device tables with a 0.5 topology weight reproduce the oracle apply."""
    from yasps_amd.mas import duplicate_domains
    _, A, _ = cube
    h = _hierarchy(A, max_levels=levels, target_nodes_per_partition=target)
    prod = [list(map(bool, d)) for d in duplicate_domains(h)]
    assert prod == omas.duplicate_domains(h)
    layout = build_layout(h, A.indptr, A.indices)
    w = level_weights(h.num_levels, 2.5).astype(np.float64)
    mats = omas.level_matrices(h, A.toarray())
    invs = omas.domain_inverses(h, mats)
    scales = [[0.5 if f else 1.0 for f in lvl] for lvl in prod]
    r = np.random.default_rng(9).standard_normal(A.shape[0])
    ref = omas.apply(h, invs, r, w, scales)
    got = emulate_device_apply(layout, A.data, r, w, duplicate_weight=0.5)
    assert np.abs(got - ref).max() <= 1e-10 * np.abs(ref).max()


def _typed_operator(nv=60, nb=3, seed=0):
    """An SPD operator over typed nodes as an affine-body scene has them: nv 3-DoF vertex
    nodes (type 0), nb 9-DoF affine matrices (type 1) and nb 3-DoF translations (type 2),
    vertices coupled along a chain and a stride-6 lattice, every body's matrix and
    translation to each other and to six vertices. Returns dims, types, the block-row
    pairs and the 3x3 BSR operator."""
    import scipy.sparse as sp
    rng = np.random.default_rng(seed)
    dims = np.array([3] * nv + [9] * nb + [3] * nb)
    types = np.array([0] * nv + [1] * nb + [2] * nb)
    off = mh.offsets_of(dims)
    edges = [(i, i + 1) for i in range(nv - 1)] + [(i, i + 6) for i in range(nv - 6)]
    for b in range(nb):
        a, t = nv + b, nv + nb + b
        edges.append((a, t))
        edges += [(a, int(v)) for v in rng.choice(nv, 6, replace=False)]
        edges += [(t, int(v)) for v in rng.choice(nv, 6, replace=False)]
    n = int(dims.sum())
    A = np.zeros((n, n))
    for i, j in edges:
        M = 0.2 * rng.standard_normal((dims[i], dims[j]))
        A[off[i]:off[i] + dims[i], off[j]:off[j] + dims[j]] += M
        A[off[j]:off[j] + dims[j], off[i]:off[i] + dims[i]] += M.T
    for i in range(dims.size):
        B = rng.standard_normal((dims[i], dims[i]))
        A[off[i]:off[i] + dims[i], off[i]:off[i] + dims[i]] += 0.1 * (B + B.T)
    A += np.diag(np.abs(A).sum(1) + 1.0)
    bsr = sp.bsr_matrix(A, blocksize=(3, 3))
    bsr.sort_indices()
    return dims, types, bsr


def _typed_hierarchy(dims, types, bsr, **kw):
    node_of_row = np.repeat(np.arange(dims.size), dims // 3)
    pairs = node_of_row[static_pairs(bsr.indptr, bsr.indices,
                                     np.zeros(bsr.indices.size, dtype=np.uint32))]
    kw = {"max_domain_dofs": 32, "max_levels": 4, "target_nodes_per_partition": 4} | kw
    return mh.build_hierarchy(dims, pairs, type_ids=types, **kw)


def test_typed_nodes_never_mix_types_or_dimensions():
    """Coarse nodes are unions of nodes of one type and one dimension: every node's
    parent has its type and its dimension, and every level keeps the scalar count of
    each type (the reference's rule, solver/mas/hierarchy.py:105-147)."""
    dims, types, bsr = _typed_operator()
    h = _typed_hierarchy(dims, types, bsr)
    assert h.num_levels >= 3
    for l in range(1, h.num_levels):
        fine, coarse = h.levels[l - 1], h.levels[l]
        m = fine.fine_to_parent
        assert np.array_equal(coarse.types[m], fine.types)
        assert np.array_equal(coarse.dims[m], fine.dims)
        for t in range(3):
            assert coarse.dims[coarse.types == t].sum() <= fine.dims[fine.types == t].sum()
        assert coarse.num_nodes < fine.num_nodes
    assert set(h.levels[0].dims[h.levels[0].types == 1].tolist()) == {9}


def test_slot_expansion_gives_the_same_preconditioner():
    """The hierarchy restated over 3-scalar slots applies the same operator as the typed
    one (oracle apply on both, to rounding), a 9-DoF node's three slots sit in one domain
    and aggregate slot by slot, and the device tables built from the slots (emulated)
    match the typed oracle apply."""
    dims, types, bsr = _typed_operator()
    h = _typed_hierarchy(dims, types, bsr)
    hs = mh.expand_to_slots(h)
    assert hs.num_levels == h.num_levels
    assert all(np.all(lv.dims == 3) for lv in hs.levels)
    Ad = bsr.toarray()
    w = level_weights(h.num_levels, 2.5).astype(np.float64)
    invs = omas.domain_inverses(h, omas.level_matrices(h, Ad))
    invs_s = omas.domain_inverses(hs, omas.level_matrices(hs, Ad))
    r = np.random.default_rng(5).standard_normal(Ad.shape[0])
    ref = omas.apply(h, invs, r, w)
    got = omas.apply(hs, invs_s, r, w)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
    # the slots of every 9-DoF level-0 node: one domain, consecutive parent slots
    lv0 = h.levels[0]
    slot_dom = np.full(hs.levels[0].num_nodes, -1)
    for k, d in enumerate(hs.levels[0].domains):
        slot_dom[d] = k
    for a in np.flatnonzero(lv0.dims == 9):
        s0 = lv0.offsets[a] // 3
        assert len(set(slot_dom[s0:s0 + 3].tolist())) == 1
        if hs.num_levels > 1:
            p = hs.levels[0].fine_to_parent[s0:s0 + 3]
            assert np.array_equal(p, p[0] + np.arange(3))
    layout = build_layout(hs, bsr.indptr, bsr.indices)
    emu = emulate_device_apply(layout, bsr.data, r, w)
    assert np.abs(emu - ref).max() <= 1e-10 * np.abs(ref).max()


def test_48_scalar_domains(cube):
    """Domains of up to 48 scalars (the reference's default size) padded to 48: the
    emulated device tables apply the oracle's operator, and a 32-padded layout refuses
    such a hierarchy."""
    from yasps_amd.mas import padded_domain_size
    _, A, _ = cube
    h = _hierarchy(A, max_domain_dofs=48, target_nodes_per_partition=16)
    widest = max(3 * d.size for lv in h.levels for d in lv.domains)
    assert 32 < widest <= 48 and padded_domain_size(widest) == 48
    assert padded_domain_size(30) == 32
    with pytest.raises(ValueError):
        padded_domain_size(51)
    with pytest.raises(ValueError):
        build_layout(h, A.indptr, A.indices, 32)
    layout = build_layout(h, A.indptr, A.indices, 48)
    assert layout.domain_scalars == 48
    w = level_weights(h.num_levels, 2.5).astype(np.float64)
    invs = omas.domain_inverses(h, omas.level_matrices(h, A.toarray()))
    r = np.random.default_rng(8).standard_normal(A.shape[0])
    ref = omas.apply(h, invs, r, w)
    got = emulate_device_apply(layout, A.data, r, w)
    assert np.abs(got - ref).max() <= 1e-10 * np.abs(ref).max()
