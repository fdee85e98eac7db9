"""Integer restatement of the sparsity pipeline (bit-exact target).

Oracle module (test infrastructure only).

Four stages, each following the named reference code:

1. ``get_indices``  - ``kernel/Coordinate/gradientIndicesKernel.pyx:546-600`` (sizes)
   and ``:621-950`` (emission). Every element gets K slots of
   (index, size): index = wrt start + instance * size + 2 for a DATA leaf that is a
   minimisation target (``:699-706``), 1 for a DATA leaf that is not (``:707-713``),
   0 for the padding slot a UNION adds after its children (``:571,809-811``), whose
   size is the union's max child gradient size minus the active child's.
2. ``compress_indices`` - ``gradientIndicesKernel.pyx:62-214``: first occurrence of
   an index >= 2 (and every index-1 slot) gets perm = running offset + 1, a repeat
   gets -(perm of the first); padding keeps perm 0. Gradient size counts every non
   padding first occurrence; m counts unique indices >= 2. Elements are grouped by
   compressed size with a stable sort; coordinate counts are m(m+1)/2, prefix summed.
3. ``compute_coordinates`` - ``gradientIndicesKernel.pyx:234-296``: for slots i <= j
   (slot order) with perm > 0 and index >= 2, the pair (min, max) of raw indices - 2
   and the matching (h, w) sizes.
4. ``unique_coordinates`` - ``kernel/Coordinate/coordinateCompressionKernel.pyx:34-144``
   (sort by (h, w, row, col), unique, per-occurrence ordinal) and ``:264-419``
   (value offsets = exclusive scan of h*w, per-(h,w) category starts/counts/dims,
   lookup = value offset of the occurrence's block).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional, Sequence
import numpy as np


# ---------------------------------------------------------------------------
# Index program nodes
# ---------------------------------------------------------------------------
@dataclass
class Leaf:
    """A DATA attribute reached on a path. ``wrt_pos`` is its position in the
    minimisation target list or None; ``per_instance`` is False for scene/mesh level
    attributes (multiplier 0, ``gradientIndicesKernel.pyx:706``)."""
    size: int
    wrt_pos: Optional[int] = None
    per_instance: bool = True

    def sizes(self):
        return self.size, 1


@dataclass
class Join:
    """JOIN through a fixed-arity connectivity ``conn`` [N, d] (row-major)."""
    conn: np.ndarray
    children: List[object]

    @property
    def dim(self):
        return int(np.asarray(self.conn).reshape(len(self.conn), -1).shape[1])

    def sizes(self):
        g = sum(c.sizes()[0] for c in self.children) * self.dim
        k = sum(c.sizes()[1] for c in self.children) * self.dim
        return g, k


@dataclass
class Union:
    """UNION over stacked child primitives. ``members[c]`` lists the path nodes of
    child c (empty when the child contributes no path)."""
    counts: Sequence[int]
    members: List[List[object]]

    def child_sizes(self, c):
        g = sum(n.sizes()[0] for n in self.members[c])
        k = sum(n.sizes()[1] for n in self.members[c])
        return g, k

    def sizes(self):
        g = 0
        k = 0
        for c in range(len(self.members)):
            cg, ck = self.child_sizes(c)
            g = max(g, cg)
            k = max(k, ck + 1)
        return g, k


@dataclass
class Concat:
    """A computed attribute: its path children evaluated at the same instance."""
    children: List[object]

    def sizes(self):
        return (sum(c.sizes()[0] for c in self.children),
                sum(c.sizes()[1] for c in self.children))


def _emit(node, idx, wrt_start, out_i, out_s, off):
    if isinstance(node, Leaf):
        if node.wrt_pos is not None:
            mult = node.size if node.per_instance else 0
            out_i[off] = wrt_start[node.wrt_pos] + idx * mult + 2
        else:
            out_i[off] = 1
        out_s[off] = node.size
    elif isinstance(node, Join):
        d = node.dim
        conn = np.asarray(node.conn).reshape(-1)
        per = node.sizes()[1] // d
        for i in range(d):
            acc = 0
            target = int(conn[idx * d + i])
            for child in node.children:
                _emit(child, target, wrt_start, out_i, out_s, off + per * i + acc)
                acc += child.sizes()[1]
    elif isinstance(node, Union):
        total = 0
        which = 0
        for c, cnt in enumerate(node.counts):
            if total + cnt > idx:
                which = c
                break
            total += cnt
        shifted = idx - total
        g_union, k_union = node.sizes()
        shift = 0
        actual = 0
        for member in node.members[which]:
            _emit(member, shifted, wrt_start, out_i, out_s, off + shift)
            shift += member.sizes()[1]
            actual += member.sizes()[0]
        out_s[off + k_union - 1] = g_union - actual
    elif isinstance(node, Concat):
        acc = 0
        for child in node.children:
            _emit(child, idx, wrt_start, out_i, out_s, off + acc)
            acc += child.sizes()[1]
    else:
        raise TypeError(f"unknown index node {node!r}")


def get_indices(program, num_instances: int, wrt_start: Sequence[int]):
    """Return (indices uint32 [N, K], sizes uint16 [N, K]); buffers start zeroed as in
    ``gradientIndicesKernel.pyx:969-972``."""
    _, K = program.sizes()
    out_i = np.zeros((num_instances, K), dtype=np.int64)
    out_s = np.zeros((num_instances, K), dtype=np.int64)
    for e in range(num_instances):
        _emit(program, e, wrt_start, out_i[e], out_s[e], 0)
    return out_i.astype(np.uint32), out_s.astype(np.uint16)


# ---------------------------------------------------------------------------
# Local compression
# ---------------------------------------------------------------------------
@dataclass
class Compression:
    permutations: np.ndarray        # int16 [N, K]
    gradient_sizes: np.ndarray      # uint16 [N]
    unique_counts: np.ndarray       # uint32 [N]  (m)
    coordinates_outer: np.ndarray   # uint32 [N+1]
    unique_sizes: np.ndarray        # uint16 [U]
    grouped_inner: np.ndarray       # uint32 [N]
    grouped_outer: np.ndarray       # uint32 [U+1]


def compress_indices(indices, sizes) -> Compression:
    indices = np.asarray(indices, dtype=np.int64)
    sizes = np.asarray(sizes, dtype=np.int64)
    N, K = indices.shape
    perm = np.zeros((N, K), dtype=np.int64)
    gsize = np.zeros(N, dtype=np.int64)
    m = np.zeros(N, dtype=np.int64)
    for e in range(N):
        offset = 0
        for i in range(K):
            idx = indices[e, i]
            if idx == 0:
                continue
            found = False
            if idx >= 2:
                for j in range(i):
                    if indices[e, j] == idx:
                        perm[e, i] = -perm[e, j]
                        found = True
                        break
            if not found:
                perm[e, i] = offset + 1
                offset += sizes[e, i]
                gsize[e] += sizes[e, i]
                if idx >= 2:
                    m[e] += 1
    order = np.argsort(gsize, kind="stable")
    uniq, counts = np.unique(gsize[order], return_counts=True)
    outer = np.zeros(len(uniq) + 1, dtype=np.int64)
    outer[1:] = np.cumsum(counts)
    coord_outer = np.zeros(N + 1, dtype=np.int64)
    coord_outer[1:] = np.cumsum(m * (m + 1) // 2)
    return Compression(perm.astype(np.int16), gsize.astype(np.uint16), m.astype(np.uint32),
                       coord_outer.astype(np.uint32), uniq.astype(np.uint16),
                       order.astype(np.uint32), outer.astype(np.uint32))


def compute_coordinates(indices, permutations, sizes, coordinates_outer):
    """Return (coords uint32 [2C], dims uint16 [2C])."""
    indices = np.asarray(indices, dtype=np.int64)
    perm = np.asarray(permutations, dtype=np.int64)
    sizes = np.asarray(sizes, dtype=np.int64)
    N, K = indices.shape
    C = int(coordinates_outer[-1])
    coords = np.zeros(2 * C, dtype=np.int64)
    dims = np.zeros(2 * C, dtype=np.int64)
    for e in range(N):
        t = 2 * int(coordinates_outer[e])
        for i in range(K):
            ii = indices[e, i]
            if not (perm[e, i] > 0 and ii >= 2):
                continue
            for j in range(i, K):
                jj = indices[e, j]
                if not (perm[e, j] > 0 and jj >= 2):
                    continue
                if ii < jj:
                    coords[t], coords[t + 1] = ii - 2, jj - 2
                    dims[t], dims[t + 1] = sizes[e, i], sizes[e, j]
                else:
                    coords[t], coords[t + 1] = jj - 2, ii - 2
                    dims[t], dims[t + 1] = sizes[e, j], sizes[e, i]
                t += 2
    return coords.astype(np.uint32), dims.astype(np.uint16)


# ---------------------------------------------------------------------------
# Global dedupe
# ---------------------------------------------------------------------------
@dataclass
class BlockLayout:
    """Block storage layout of one (static or dynamic) Hessian part
    (attribute names of ``matrixAndVector/matrix/hessian.pyx:783-808``)."""
    block_positions: np.ndarray      # uint32 [2*nb] (row, col) scalar starts
    blocks_start_indices: List[int]  # value offset (doubles) per category, + total
    block_counts: List[int]
    block_dimensions: List[int]      # flattened (h, w) per category
    lookups: List[np.ndarray]        # per input term: uint32 value offset per coordinate
    num_unique: int = 0
    block_offsets: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint32))

    @property
    def total_size(self) -> int:
        return int(self.blocks_start_indices[-1]) if self.blocks_start_indices else 0


def unique_coordinates(coords_list, dims_list) -> BlockLayout:
    """Dedupe the concatenated coordinates of all terms (term order)."""
    nums = [len(c) // 2 for c in coords_list]
    if sum(nums) == 0:
        return BlockLayout(np.zeros(0, np.uint32), [0], [], [], [np.zeros(0, np.uint32) for _ in nums])
    coords = np.concatenate([np.asarray(c, dtype=np.int64) for c in coords_list]).reshape(-1, 2)
    dims = np.concatenate([np.asarray(d, dtype=np.int64) for d in dims_list]).reshape(-1, 2)
    # sort key (h, w, row, col); lexsort uses the last key as primary
    order = np.lexsort((coords[:, 1], coords[:, 0], dims[:, 1], dims[:, 0]))
    sc, sd = coords[order], dims[order]
    new_run = np.ones(len(order), dtype=bool)
    new_run[1:] = np.any(sc[1:] != sc[:-1], axis=1) | np.any(sd[1:] != sd[:-1], axis=1)
    ordinal = np.cumsum(new_run) - 1
    lookup_ord = np.empty(len(order), dtype=np.int64)
    lookup_ord[order] = ordinal
    uc, ud = sc[new_run], sd[new_run]
    sizes = ud[:, 0] * ud[:, 1]
    offsets = np.zeros(len(uc), dtype=np.int64)
    offsets[1:] = np.cumsum(sizes)[:-1]
    starts_cat = np.ones(len(uc), dtype=bool)
    starts_cat[1:] = np.any(ud[1:] != ud[:-1], axis=1)
    cat_first = np.nonzero(starts_cat)[0]
    counts = np.diff(np.append(cat_first, len(uc)))
    starts = list(offsets[cat_first].tolist()) + [int(offsets[-1] + sizes[-1])]
    block_dims = ud[cat_first].reshape(-1).tolist()
    lookup = offsets[lookup_ord]
    lookups = []
    pos = 0
    for n in nums:
        lookups.append(lookup[pos:pos + n].astype(np.uint32))
        pos += n
    return BlockLayout(uc.reshape(-1).astype(np.uint32), [int(s) for s in starts],
                       [int(c) for c in counts], [int(d) for d in block_dims], lookups,
                       num_unique=len(uc), block_offsets=offsets.astype(np.uint32))


@dataclass
class TermIndices:
    """Everything the element scatter of one term needs."""
    indices: np.ndarray
    sizes: np.ndarray
    compression: Compression
    coords: np.ndarray
    dims: np.ndarray

    @property
    def K(self) -> int:
        return self.indices.shape[1]


def term_indices(program, num_instances, wrt_start) -> TermIndices:
    idx, sz = get_indices(program, num_instances, wrt_start)
    comp = compress_indices(idx, sz)
    coords, dims = compute_coordinates(idx, comp.permutations, sz, comp.coordinates_outer)
    return TermIndices(idx, sz, comp, coords, dims)


def build_layout(terms: List[TermIndices]) -> BlockLayout:
    """``hessian.getSparseIndices`` (``hessian.pyx:783-808``) for one part."""
    return unique_coordinates([t.coords for t in terms], [t.dims for t in terms])
