"""Host-side multilevel additive Schwarz (MAS) hierarchy.

Built once per static sparsity structure, as the reference does
(``solver/mas/hierarchy.py:150-271``):

* level 0 nodes are the target variable blocks, edges the off-diagonal static blocks
  (``solver/mas/block_graph.py:13-48``);
* nodes are grouped into connected, locality-preserving groups of about
  ``target_nodes_per_partition`` nodes. The reference uses METIS recursive
  partitioning (``solver/mas/metis.py:42-90``), which is not available here; groups are
  grown greedily by breadth-first search from the lowest unassigned node instead, so
  MAS domains (and iteration counts) differ from the reference's;
* groups are packed into domains of at most ``max_domain_dofs`` scalar DoF without
  crossing group boundaries (``solver/mas/partition.py:41-73``, same rule);
* the next level's nodes are unions of adjacent nodes of one domain with equal
  dimension and equal type (``hierarchy.py:105-147``, same union-find rule), the
  transfer is identity aggregation (``solver/mas/transfer.py``);
* a last level that fits in one domain becomes a single global domain
  (``hierarchy.py:257-269``).

Level-0 nodes are the targets' instances (``solver/yaspsMatrixView.py:69-78``): a vertex
position is a 3-DoF node, an affine body's matrix a 9-DoF node and its translation a
3-DoF node, each typed by its target, so a body's matrix DoFs always share a domain and
coarse nodes never mix targets. ``expand_to_slots`` restates such a hierarchy over
3-scalar slots (a d-DoF node is d/3 consecutive slots, slot k of a node aggregating into
slot k of its parent), which is the form the device tables take.
"""
from __future__ import annotations

from collections import deque
from dataclasses import dataclass, field
from typing import List

import numpy as np


@dataclass
class Level:
    dims: np.ndarray                 # [nodes]
    offsets: np.ndarray              # scalar offset of every node
    xadj: np.ndarray                 # CSR adjacency (no self loops)
    adjncy: np.ndarray
    domains: List[np.ndarray]        # node lists
    fine_to_parent: np.ndarray = None  # map to the next level's nodes
    types: np.ndarray = None         # type id of every node (None: one type)

    @property
    def num_nodes(self) -> int:
        return int(self.dims.size)

    @property
    def num_dofs(self) -> int:
        return int(self.dims.sum())


@dataclass
class Hierarchy:
    levels: List[Level]
    composed: List[np.ndarray] = field(default_factory=list)   # level-0 node -> level node

    @property
    def num_levels(self) -> int:
        return len(self.levels)


def graph_from_pairs(n: int, pairs: np.ndarray):
    """CSR adjacency of the undirected graph with the given (i, j) pairs (i != j)."""
    pairs = np.asarray(pairs, dtype=np.int64).reshape(-1, 2)
    pairs = pairs[pairs[:, 0] != pairs[:, 1]]
    a = np.concatenate([pairs[:, 0], pairs[:, 1]])
    b = np.concatenate([pairs[:, 1], pairs[:, 0]])
    key = np.unique(a * n + b)
    a, b = key // n, key % n
    xadj = np.zeros(n + 1, dtype=np.int64)
    np.add.at(xadj, a + 1, 1)
    return np.cumsum(xadj), b.astype(np.int64)


def bfs_groups(xadj, adjncy, target: int) -> List[List[int]]:
    """Connected groups of up to ``target`` nodes, grown breadth first from the lowest
    unassigned node (replacement of the METIS recursive partition)."""
    n = xadj.size - 1
    assigned = np.zeros(n, dtype=bool)
    groups = []
    for s in range(n):
        if assigned[s]:
            continue
        group = [s]
        assigned[s] = True
        q = deque([s])
        while q and len(group) < target:
            v = q.popleft()
            for u in adjncy[xadj[v]:xadj[v + 1]]:
                if not assigned[u]:
                    assigned[u] = True
                    group.append(int(u))
                    q.append(u)
                    if len(group) >= target:
                        break
        groups.append(group)
    return groups


def pack_domain_groups(groups, dims, max_domain_dofs: int) -> List[List[int]]:
    """Greedy packing of each group into domains of at most ``max_domain_dofs``
    scalars, never crossing a group boundary; an oversized node is its own domain."""
    domains = []
    for group in groups:
        cur, cur_dofs = [], 0
        for node in group:
            d = int(dims[node])
            if cur and cur_dofs + d > max_domain_dofs:
                domains.append(cur)
                cur, cur_dofs = [], 0
            cur.append(int(node))
            cur_dofs += d
            if d > max_domain_dofs:
                domains.append(cur)
                cur, cur_dofs = [], 0
        if cur:
            domains.append(cur)
    return domains


def collapse_compatible(xadj, adjncy, domains, dims, types=None):
    """Union of adjacent nodes of one domain with equal dimension and type -> next level
    nodes: (fine-to-coarse map, coarse dims, coarse types)."""
    n = dims.size
    parent = np.arange(n)

    def find(v):
        while parent[v] != v:
            parent[v] = parent[parent[v]]
            v = parent[v]
        return v

    node_domain = np.full(n, -1, dtype=np.int64)
    for d, dom in enumerate(domains):
        node_domain[np.asarray(dom)] = d
    src = np.repeat(np.arange(n), np.diff(xadj))
    for i, j in zip(src, adjncy):
        if (i < j and node_domain[i] == node_domain[j] and dims[i] == dims[j]
                and (types is None or types[i] == types[j])):
            a, b = find(i), find(j)
            if a != b:
                parent[max(a, b)] = min(a, b)
    roots = np.array([find(v) for v in range(n)])
    uniq, mapping = np.unique(roots, return_inverse=True)
    # compact ids in order of first appearance of each root (node order)
    first = np.full(uniq.size, n, dtype=np.int64)
    np.minimum.at(first, mapping, np.arange(n))
    order = np.argsort(first, kind="stable")
    rank = np.empty_like(order)
    rank[order] = np.arange(order.size)
    mapping = rank[mapping]
    coarse_dims = np.zeros(uniq.size, dtype=np.int64)
    coarse_dims[mapping] = dims
    coarse_types = None
    if types is not None:
        coarse_types = np.zeros(uniq.size, dtype=np.int64)
        coarse_types[mapping] = types
    return mapping.astype(np.int64), coarse_dims, coarse_types


def offsets_of(dims):
    return np.concatenate([[0], np.cumsum(dims)[:-1]]).astype(np.int64)


def build_hierarchy(dims, pairs, max_domain_dofs: int = 32, max_levels: int = 4,
                    target_nodes_per_partition: int = 10,
                    global_coarse_domain: bool = True,
                    minimum_domains_for_next_level: int = 0,
                    domain_dof_schedule=None, target_node_schedule=None,
                    type_ids=None) -> Hierarchy:
    """``dims``: dimension of every level-0 node; ``pairs``: off-diagonal node pairs of
    the static blocks; ``type_ids``: type of every level-0 node (None: one type). The level
    loop and its stopping rules follow
    ``solver/mas/hierarchy.py:193-270``: per-level domain sizes and partition targets from
    the schedules (the last entry repeats), and no further level once the level before
    the last one (from level 3 on) has fewer than ``minimum_domains_for_next_level``
    domains."""
    if minimum_domains_for_next_level < 0:
        raise ValueError("minimum_domains_for_next_level must be non-negative")
    for sched, what in ((domain_dof_schedule, "domain DOF"), (target_node_schedule, "target-node")):
        if sched is not None and (not sched or any(int(v) <= 0 for v in sched)):
            raise ValueError(f"{what} schedule must contain positive values")

    def at(sched, default, level):
        return default if sched is None else int(sched[min(level, len(sched) - 1)])

    dims = np.asarray(dims, dtype=np.int64)
    n = dims.size
    types = None if type_ids is None else np.asarray(type_ids, dtype=np.int64).reshape(-1)
    if types is not None and types.size != n:
        raise ValueError("type_ids needs one entry per node")
    xadj, adjncy = graph_from_pairs(n, pairs)
    levels: List[Level] = []
    composed = [np.arange(n, dtype=np.int64)]
    while len(levels) < max_levels:
        li = len(levels)
        groups = bfs_groups(xadj, adjncy, at(target_node_schedule, target_nodes_per_partition, li))
        domains = pack_domain_groups(groups, dims, at(domain_dof_schedule, max_domain_dofs, li))
        lvl = Level(dims.copy(), offsets_of(dims), xadj, adjncy,
                    [np.asarray(d, dtype=np.int64) for d in domains],
                    types=None if types is None else types.copy())
        levels.append(lvl)
        if len(levels) >= max_levels or dims.size <= 1 or adjncy.size == 0:
            break
        if (len(levels) >= 3 and minimum_domains_for_next_level
                and len(levels[-2].domains) < minimum_domains_for_next_level):
            break
        mapping, cdims, ctypes = collapse_compatible(xadj, adjncy, domains, dims, types)
        if cdims.size == dims.size:
            break
        lvl.fine_to_parent = mapping
        composed.append(mapping[composed[-1]])
        src = np.repeat(np.arange(dims.size), np.diff(xadj))
        cp = np.stack([mapping[src], mapping[adjncy]], 1)
        xadj, adjncy = graph_from_pairs(cdims.size, cp)
        dims, types = cdims, ctypes
    last = levels[-1]
    cap = at(domain_dof_schedule, max_domain_dofs, len(levels) - 1)
    if global_coarse_domain and len(last.domains) > 1 and last.num_dofs <= cap:
        last.domains = [np.arange(last.num_nodes, dtype=np.int64)]
    return Hierarchy(levels, composed)


def expand_to_slots(h: Hierarchy, width: int = 3) -> Hierarchy:
    """The same hierarchy over ``width``-scalar slots: node a of dimension d becomes the
    d / width slots offsets[a] / width + k, slot k of a maps to slot k of a's parent (a
    parent has its children's dimension), and a domain lists its nodes' slots in node
    order. Restriction, prolongation, domain matrices and the apply are unchanged (every
    scalar keeps its index, its domain and its parent scalar); a hierarchy whose nodes
    are all ``width`` wide is returned as is."""
    if all(np.all(lv.dims == width) for lv in h.levels):
        return h
    levels = []
    for l, lv in enumerate(h.levels):
        if np.any(lv.dims % width) or np.any(lv.offsets % width):
            raise ValueError(f"node dimensions must be multiples of {width}")
        first = lv.offsets // width
        count = lv.dims // width
        total = int(count.sum())
        node = np.repeat(np.arange(lv.num_nodes, dtype=np.int64), count)
        k = np.arange(total, dtype=np.int64) - first[node]
        domains = [np.concatenate([first[a] + np.arange(count[a], dtype=np.int64) for a in d])
                   for d in lv.domains]
        parent = None
        if lv.fine_to_parent is not None:
            nxt = h.levels[l + 1]
            parent = nxt.offsets[lv.fine_to_parent[node]] // width + k
        empty = np.zeros(0, dtype=np.int64)
        levels.append(Level(np.full(total, width, dtype=np.int64),
                            width * np.arange(total, dtype=np.int64),
                            np.zeros(total + 1, dtype=np.int64), empty, domains, parent,
                            None if lv.types is None else lv.types[node]))
    composed = [np.arange(levels[0].num_nodes, dtype=np.int64)]
    for lv in levels[:-1]:
        if lv.fine_to_parent is None:
            break
        composed.append(lv.fine_to_parent[composed[-1]])
    return Hierarchy(levels, composed)
