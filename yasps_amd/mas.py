"""Multilevel additive Schwarz PCG solver (``solver/masSolver.py``, ``solver/mas/``).

``masSolver`` keeps the reference's split between a hierarchy built once per static
sparsity structure (``solver/mas/solver.py:266-295``) and numeric work on every solve
(``solver/mas/numeric_assembly.py``, ``local_inverse.py``): on the device, each solve
coarsens the current Hessian level by level (identity-aggregation Galerkin sums),
scatters the in-domain blocks into dense 32 x 32 (or 48 x 48) domain matrices, inverts them (FP64
Gauss-Jordan, FP32 storage) and runs PCG with z = sum_l w_l P_l B_l^-1 R_l r
(``solver/mas/preconditioner.py:128-154``, ``solver/mas/pcg.py:23-102``).

Level-0 nodes are the targets' instances, typed by target, as in the reference: a
vertex or an affine body's translation is a 3-DoF node, an affine body's matrix a 9-DoF
node whose three block rows stay in one domain at every level. Target sizes must be
multiples of 3 and domains hold at most 48 scalars (padded to 32 when every domain fits
in 32). The node grouping is a breadth-first
replacement of METIS (see ``mas_hierarchy``), so domains and iteration counts differ from
the reference's; the preconditioner is still symmetric positive definite and PCG reaches
the same stated residual.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass, field
from typing import List

import numpy as np

from . import _lib, device as dev
from .hessian import vector
from .mas_hierarchy import Hierarchy, build_hierarchy, expand_to_slots
from .refine import RefinementBuffers, refine_solve


class LocalInverseError(np.linalg.LinAlgError):
    """A MAS domain matrix could not be inverted (``solver/mas/local_inverse.py:15``): a
    Gauss-Jordan pivot at or below 1e-12 of the domain's largest diagonal entry, as for a
    singular (PSD but not definite) or indefinite operator. The reference raises it from its
    numeric update (``solver/mas/cuda_runtime.py:2904-2920``)."""

_MODES = {"reference": 0, "relative_residual": 1}
DOMAIN_SCALARS = 32          # padded domain size of the device kernels (the default)
MAX_DOMAIN_SCALARS = 48      # the larger padded size (the reference's default domain)


def padded_domain_size(scalars: int) -> int:
    """Padded domain size of the device kernels for domains of up to ``scalars``."""
    if scalars > MAX_DOMAIN_SCALARS:
        raise ValueError(f"MAS domains hold at most {MAX_DOMAIN_SCALARS} scalars on this device path")
    return DOMAIN_SCALARS if scalars <= DOMAIN_SCALARS else MAX_DOMAIN_SCALARS
NO_DEST = np.uint32(0xFFFFFFFF)


@dataclass
class LevelBlocks:
    """Blocks of one level: (row, col) node pairs and, for coarse levels, the CSR of the
    finer level's blocks summed into each block (at level 0, when the operator repeats a
    (row, col) pair, the CSR of its entries summed into each unique block)."""
    rows: np.ndarray
    cols: np.ndarray
    coarsen_offsets: np.ndarray = None
    coarsen_items: np.ndarray = None
    dest: np.ndarray = None          # offset into the domain matrices or NO_DEST

    @property
    def nnz(self) -> int:
        return int(self.rows.size)


@dataclass
class MasLayout:
    hierarchy: Hierarchy
    blocks: List[LevelBlocks]
    level_offsets: np.ndarray            # coarse id ranges [num_coarse_levels + 1]
    restrict_offsets: np.ndarray
    restrict_items: np.ndarray
    prolong_maps: np.ndarray             # [num_coarse_levels, nv]
    domain_nodes: np.ndarray
    domain_node_offsets: np.ndarray
    domain_sizes: np.ndarray             # scalars per domain
    level0_domains: int
    weights: np.ndarray = field(default=None)
    duplicate: np.ndarray = field(default=None)   # per domain: repeats a previous-level domain
    weights_base: np.ndarray = field(default=None)   # configured level weights
    domain_scalars: int = DOMAIN_SCALARS          # padded domain size D (32 or 48)

    @property
    def num_domains(self) -> int:
        return int(self.domain_sizes.size)

    @property
    def num_coarse_levels(self) -> int:
        return int(self.level_offsets.size - 1)


def _csr_by_key(keys: np.ndarray, num_keys: int):
    """Stable grouping of positions by key: offsets [num_keys + 1], items (positions)."""
    order = np.argsort(keys, kind="stable")
    counts = np.bincount(keys, minlength=num_keys)
    offsets = np.zeros(num_keys + 1, dtype=np.int64)
    np.cumsum(counts, out=offsets[1:])
    return offsets, order


def level_weights(num_levels: int, coarsest_level_weight: float = 2.5,
                  weights=None) -> np.ndarray:
    """(1, ..., 1, coarsest) or configured weights cut to the constructed levels
    (``solver/mas/solver.py:297-317``, ``preconditioner.py:27-36``)."""
    if weights is None:
        w = [1.0] * (num_levels - 1) + [float(coarsest_level_weight)]
        if num_levels == 1:
            w = [float(coarsest_level_weight)]
        return np.asarray(w, dtype=np.float32)
    weights = tuple(float(x) for x in weights)
    if len(weights) == num_levels:
        return np.asarray(weights, dtype=np.float32)
    if len(weights) > num_levels:
        w = weights[:1] if num_levels == 1 else weights[:num_levels - 1] + weights[-1:]
        return np.asarray(w, dtype=np.float32)
    raise ValueError("level_weights must contain at least one weight per constructed "
                     f"hierarchy level ({num_levels})")


def static_pairs(row_ptr: np.ndarray, col: np.ndarray, src: np.ndarray,
                 dynamic_bit: int = 30) -> np.ndarray:
    """Off-diagonal node pairs of the static blocks of a full block-CSR (src bit 30 marks
    dynamic blocks; bit 43 in the split block-CSR of affine-body layouts)."""
    nbr = row_ptr.size - 1
    rows = np.repeat(np.arange(nbr, dtype=np.int64), np.diff(row_ptr.astype(np.int64)))
    keep = ((src.astype(np.int64) >> dynamic_bit) & 1) == 0
    keep &= rows != col
    return np.stack([rows[keep], col.astype(np.int64)[keep]], 1)


def target_nodes(h, n: int):
    """Level-0 node dimensions and type ids of a Hessian's targets
    (``solver/yaspsMatrixView.py:69-78``: one node per instance of each target, typed by
    the target's position in the list), and the node of every 3-scalar block row."""
    dims = np.concatenate([np.full(int(w.numInstances), int(w.size), dtype=np.int64)
                           for w in h.wrt])
    types = np.concatenate([np.full(int(w.numInstances), t, dtype=np.int64)
                            for t, w in enumerate(h.wrt)])
    if np.any(dims % 3) or int(dims.sum()) != n:
        raise NotImplementedError("masSolver: the device MAS path needs target sizes that are "
                                  "multiples of 3")
    return dims, types, np.repeat(np.arange(dims.size, dtype=np.int64), dims // 3)


def duplicate_domains(h: Hierarchy) -> List[np.ndarray]:
    """Per level, per domain: True when the domain only repeats a domain of the previous
    level, i.e. each of its nodes has exactly one child and the children are exactly the
    nodes of one previous-level domain (the topology rule of
    ``solver/mas/cuda_runtime.py:1285-1340``; such a domain's correction is weighted by
    ``duplicate_level_weight``). Level 0 has none."""
    out = [np.zeros(len(h.levels[0].domains), dtype=bool)]
    for l in range(1, h.num_levels):
        prev, cur = h.levels[l - 1], h.levels[l]
        mapping = np.asarray(prev.fine_to_parent, dtype=np.int64)
        nchild = np.bincount(mapping, minlength=cur.num_nodes)
        only_child = np.full(cur.num_nodes, -1, dtype=np.int64)
        only_child[mapping] = np.arange(mapping.size)
        prev_dom = np.full(prev.num_nodes, -1, dtype=np.int64)
        for k, d in enumerate(prev.domains):
            prev_dom[d] = k
        dup = np.zeros(len(cur.domains), dtype=bool)
        for k, d in enumerate(cur.domains):
            if np.any(nchild[d] != 1):
                continue
            kids = only_child[d]
            src = prev_dom[kids[0]]
            if np.all(prev_dom[kids] == src) and set(kids.tolist()) == set(prev.domains[src].tolist()):
                dup[k] = True
        out.append(dup)
    return out


def build_layout(h: Hierarchy, row_ptr: np.ndarray, col: np.ndarray,
                 domain_scalars: int = DOMAIN_SCALARS) -> MasLayout:
    """Device tables of the MAS apply and numeric update for the full block-CSR
    (row_ptr, col) of the current operator and a hierarchy over its block rows, with
    domains padded to ``domain_scalars`` (32 or 48)."""
    if domain_scalars not in (DOMAIN_SCALARS, MAX_DOMAIN_SCALARS):
        raise ValueError("MAS layout: domains are padded to 32 or 48 scalars")
    D = domain_scalars
    lv0 = h.levels[0]
    nv = lv0.num_nodes
    if np.any(lv0.dims != 3) or row_ptr.size != nv + 1:
        raise ValueError("MAS layout: every node must be a 3-vector block row")
    if h.num_levels - 1 > _lib.MAS_MAX_COARSE_LEVELS:
        raise ValueError(f"MAS layout: at most {_lib.MAS_MAX_COARSE_LEVELS + 1} levels")
    sizes = [lv.num_nodes for lv in h.levels]
    bases = np.zeros(h.num_levels, dtype=np.int64)          # global coarse id of node 0
    bases[1:] = np.concatenate([[0], np.cumsum(sizes[1:])[:-1]]) if h.num_levels > 1 else []
    level_offsets = np.concatenate([[0], np.cumsum(sizes[1:])]).astype(np.int64)

    # restriction: level-0 descendants of the nodes of levels 1 to 3; for deeper levels
    # the level-3 descendants as coarse ids (the hierarchy is nested, so their residuals
    # sum to the same level-0 sum, read from a few entries instead of every vertex)
    r_off, r_items = [np.zeros(1, dtype=np.int64)], []
    for l in range(1, h.num_levels):
        if l <= 3:
            off, items = _csr_by_key(h.composed[l], sizes[l])
        else:
            anc = np.full(sizes[3], -1, dtype=np.int64)
            anc[h.composed[3]] = h.composed[l]
            if np.any(anc < 0) or not np.array_equal(anc[h.composed[3]], h.composed[l]):
                raise ValueError(f"MAS layout: level {l} is not nested over level 3")
            off, items = _csr_by_key(anc, sizes[l])
            items = items + bases[3]
        r_off.append(off[1:] + r_off[-1][-1])
        r_items.append(items)
    restrict_offsets = np.concatenate(r_off)
    restrict_items = (np.concatenate(r_items) if r_items else np.zeros(0, dtype=np.int64))
    prolong = (np.stack([bases[l] + h.composed[l] for l in range(1, h.num_levels)])
               if h.num_levels > 1 else np.zeros((0, nv), dtype=np.int64))

    # domains, numbered level by level
    dom_nodes, dom_off, dom_sizes = [], [0], []
    node_dom, node_pos = [], []
    for l, lv in enumerate(h.levels):
        nd = np.full(lv.num_nodes, -1, dtype=np.int64)
        npos = np.zeros(lv.num_nodes, dtype=np.int64)
        for d in lv.domains:
            g = len(dom_sizes)
            if 3 * d.size > D:
                raise ValueError(f"MAS layout: a domain holds {3 * d.size} scalars (max {D})")
            nd[d] = g
            npos[d] = np.arange(d.size)
            dom_nodes.append(d + (bases[l] if l else 0))
            dom_off.append(dom_off[-1] + d.size)
            dom_sizes.append(3 * d.size)
        if np.any(nd < 0):
            raise ValueError(f"MAS layout: level {l} has nodes outside every domain")
        node_dom.append(nd)
        node_pos.append(npos)

    # blocks of every level and their destinations in the domain matrices
    rows0 = np.repeat(np.arange(nv, dtype=np.int64), np.diff(row_ptr.astype(np.int64)))
    cols0 = col.astype(np.int64)
    key0 = rows0 * nv + cols0
    uniq0, inv0 = np.unique(key0, return_inverse=True)
    if uniq0.size < key0.size:
        # a block present in the static and the dynamic part appears twice in the full
        # block-CSR: the domain matrices take the sum, so level 0 is first summed into its
        # unique blocks (the domain scatter stores each block once)
        off0, items0 = _csr_by_key(inv0.reshape(-1), uniq0.size)
        blocks = [LevelBlocks(uniq0 // nv, uniq0 % nv, off0, items0)]
    else:
        blocks = [LevelBlocks(rows0, cols0)]
    for l in range(1, h.num_levels):
        prev = blocks[-1]
        m = h.levels[l - 1].fine_to_parent
        n_l = sizes[l]
        key = m[prev.rows] * n_l + m[prev.cols]
        uniq, inv = np.unique(key, return_inverse=True)
        off, items = _csr_by_key(inv.reshape(-1), uniq.size)
        blocks.append(LevelBlocks(uniq // n_l, uniq % n_l, off, items))
    for l, b in enumerate(blocks):
        same = node_dom[l][b.rows] == node_dom[l][b.cols]
        dest = (node_dom[l][b.rows] * (D * D) + 3 * node_pos[l][b.rows] * D
                + 3 * node_pos[l][b.cols])
        b.dest = np.where(same, dest, np.int64(NO_DEST)).astype(np.uint32)
    return MasLayout(h, blocks, level_offsets, restrict_offsets, restrict_items, prolong,
                     np.concatenate(dom_nodes), np.asarray(dom_off, dtype=np.int64),
                     np.asarray(dom_sizes, dtype=np.int64), len(h.levels[0].domains),
                     duplicate=np.concatenate(duplicate_domains(h)), domain_scalars=D)


class _DeviceMas:
    """Device copies of a layout plus the per-solve numeric buffers."""

    def __init__(self, layout: MasLayout, weights: np.ndarray, n: int, coarse_bank: str = "fp32",
                 duplicate_level_weight: float = 1.0):
        self.layout = layout
        u32 = lambda a: dev.to_device(np.asarray(a, dtype=np.uint32), "uint32")   # noqa: E731
        self.dest = [u32(b.dest) for b in layout.blocks]
        self.coarsen = [(u32(b.coarsen_offsets), u32(b.coarsen_items))
                        if b.coarsen_offsets is not None else (None, None) for b in layout.blocks]
        self.merged0 = layout.blocks[0].coarsen_offsets is not None
        self.values = ([dev.empty(9 * layout.blocks[0].nnz, "float64") if self.merged0 else None]
                       + [dev.empty(9 * b.nnz, "float64") for b in layout.blocks[1:]])
        nc = int(layout.level_offsets[-1])
        self.restrict_offsets = u32(layout.restrict_offsets)
        self.restrict_items = u32(layout.restrict_items)
        self.prolong = u32(layout.prolong_maps.reshape(-1))
        self.domain_nodes = u32(layout.domain_nodes)
        self.domain_node_offsets = u32(layout.domain_node_offsets)
        self.domain_sizes = u32(layout.domain_sizes)
        D = self.D = layout.domain_scalars
        self.domains = dev.empty(layout.num_domains * D * D, "float64")
        # domains >= bf16_first keep BF16 inverses when the coarse bank is BF16
        self.bf16_first = layout.level0_domains if coarse_bank == "bf16" else layout.num_domains
        self.inverses = dev.empty(max(1, self.bf16_first) * D * D, "float32")
        nb16 = layout.num_domains - self.bf16_first
        self.inverses_bf16 = dev.empty(nb16 * D * D, "int16") if nb16 else None
        self.status = dev.zeros(1, "int32")
        self.weights = dev.to_device(weights.astype(np.float32), "float32")
        self.weights_host = weights.astype(np.float32).copy()
        self.domain_scale = None
        if duplicate_level_weight != 1.0 and layout.duplicate is not None and layout.duplicate.any():
            scale = np.where(layout.duplicate, duplicate_level_weight, 1.0).astype(np.float32)
            self.domain_scale = dev.to_device(scale, "float32")
        self.coarse_residual = dev.empty(max(1, 3 * nc), "float64")
        self.coarse_correction = dev.empty(max(1, 3 * nc), "float32")
        self.fine_correction = dev.empty(n, "float32")
        a = _lib.MasArgs()
        a.restrict_offsets = dev.ptr(self.restrict_offsets)
        a.restrict_items = dev.ptr(self.restrict_items)
        a.num_coarse_nodes = nc
        a.num_coarse_levels = layout.num_coarse_levels
        for i, v in enumerate(layout.level_offsets):
            a.level_offsets[i] = int(v)
        a.prolong_maps = dev.ptr(self.prolong)
        a.inverses = dev.ptr(self.inverses)
        a.domain_nodes = dev.ptr(self.domain_nodes)
        a.domain_node_offsets = dev.ptr(self.domain_node_offsets)
        a.num_domains = layout.num_domains
        a.level0_domains = layout.level0_domains
        a.weights = dev.ptr(self.weights)
        a.coarse_residual = dev.ptr(self.coarse_residual)
        a.fine_correction = dev.ptr(self.fine_correction)
        a.coarse_correction = dev.ptr(self.coarse_correction)
        a.inverses_bf16 = dev.ptr(self.inverses_bf16)
        a.bf16_first_domain = self.bf16_first
        a.domain_scale = dev.ptr(self.domain_scale)
        a.domain_size = D
        self.args = a

    def set_fine_weight(self, w0: float):
        """Level-0 weight of the apply (the adaptive fine weight changes with the dynamic
        block count)."""
        if self.weights_host[0] != np.float32(w0):
            self.weights_host[0] = np.float32(w0)
            self.weights.copy_(dev.to_device(self.weights_host, "float32"))

    def update(self, full):
        """Coarse operators, domain matrices and their inverses from the full block-CSR's
        current FP64 values (numeric_assembly.py:22-83, local_inverse.py)."""
        lib = _lib.load()
        s = dev.stream()
        self.domains.zero_()
        self.status.zero_()
        blocks = self.layout.blocks
        if self.merged0:                 # repeated (row, col) entries summed first
            off, items = self.coarsen[0]
            _lib.check(lib.ysp_mas_coarsen(dev.ptr(full.v64), 1, full.nnz, dev.ptr(off), dev.ptr(items),
                                           blocks[0].nnz, dev.ptr(self.values[0]), s), "ysp_mas_coarsen")
            _lib.check(lib.ysp_mas_domain_scatter(dev.ptr(self.values[0]), 0, blocks[0].nnz,
                                                  dev.ptr(self.dest[0]), dev.ptr(self.domains),
                                                  self.D, s),
                       "ysp_mas_domain_scatter")
        else:
            _lib.check(lib.ysp_mas_domain_scatter(dev.ptr(full.v64), 1, full.nnz, dev.ptr(self.dest[0]),
                                                  dev.ptr(self.domains), self.D, s),
                       "ysp_mas_domain_scatter")
        for l in range(1, len(blocks)):
            from_full = l == 1 and not self.merged0
            prev_vals = full.v64 if from_full else self.values[l - 1]
            prev_nnz = full.nnz if from_full else blocks[l - 1].nnz
            off, items = self.coarsen[l]
            _lib.check(lib.ysp_mas_coarsen(dev.ptr(prev_vals), int(from_full), prev_nnz, dev.ptr(off),
                                           dev.ptr(items), blocks[l].nnz, dev.ptr(self.values[l]), s),
                       "ysp_mas_coarsen")
            _lib.check(lib.ysp_mas_domain_scatter(dev.ptr(self.values[l]), 0, blocks[l].nnz,
                                                  dev.ptr(self.dest[l]), dev.ptr(self.domains),
                                                  self.D, s),
                       "ysp_mas_domain_scatter")
        _lib.check(lib.ysp_mas_invert_domains(dev.ptr(self.domains), dev.ptr(self.domain_sizes),
                                              self.layout.num_domains, dev.ptr(self.inverses),
                                              dev.ptr(self.inverses_bf16), self.bf16_first,
                                              dev.ptr(self.status), self.D, s),
                   "ysp_mas_invert_domains")

    def apply(self, r, z):
        _lib.check(_lib.load().ysp_mas_apply(C.byref(self.args), dev.ptr(r), r.numel(), dev.ptr(z),
                                             dev.stream()), "ysp_mas_apply")


class masSolver:
    """MAS-preconditioned CG on the device (``solver/masSolver.py:11-133``).

    Takes the options of the reference's ``MASSolver`` (``solver/mas/solver.py:89-120``)
    by name, with its validation:

    * hierarchy: ``max_domain_dofs`` (at most 48 scalars on this device path; 32, the
      configurations' leaf size, by default, where the reference defaults to 48),
      ``max_levels``, ``target_nodes_per_partition``, ``minimum_domains_for_next_level``,
      ``global_coarse_domain``, ``domain_dof_schedule``, ``target_node_schedule``;
    * weights: ``coarsest_level_weight``, ``level_weights``, ``duplicate_level_weight`` (a
      domain that repeats a previous-level domain), ``adaptive_fine_level_weight`` (default:
      on unless ``level_weights`` is given), ``dense_collision_fine_level_weight`` and
      ``dense_collision_block_ratio`` (the fine weight is raised to the former when the
      dynamic blocks per fine node reach the latter);
    * numeric updates: ``numeric_rebuild_interval`` / ``dynamic_numeric_rebuild_interval``
      (the domain inverses are rebuilt every k-th solve of an unchanged structure, lagged
      in between as ``solver/mas/solver.py:526-568``);
    * ``mixed_spmv=True`` selects the mixed-precision solve below;
    * the CUDA launch tuning (``cuda_threads_per_block``, ``cuda_fixed_inverse_bucket_size``,
      ``cuda_enable_inverse_fallback``) and ``fine_correction_iterations`` (used only by the
      reduced-then-fine mode) are accepted and have no effect on the device kernels here;
    * ``solve_mode="reduced_then_fine"``, ``dynamic_edge_domains``,
      ``collision_aware_reorder``, ``collision_merge_across_types``, another
      ``inverse_backend`` and ``allow_cpu_fallback`` raise NotImplementedError.

    Extensions: ``threshold_mode`` as for ``jacobianPCGSolver``, ``precision="mixed"`` (FP32
    block values in the Krylov loop, FP64 iterative refinement on the true relative
    residual) and ``coarse_bank="bf16"`` (the coarse levels' domain inverses stored BF16,
    level 0 FP32). Returns 0 when converged and -4 otherwise, as the reference.
    """

    def __init__(self, max_domain_dofs=None, max_levels=None, target_nodes_per_partition=None,
                 fine_correction_iterations: int = 8, solve_mode: str = "mas_pcg",
                 inverse_backend: str = "cooperative_gauss_jordan", *,
                 minimum_domains_for_next_level: int = 0, cuda_threads_per_block: int = 96,
                 cuda_fixed_inverse_bucket_size=None, dynamic_edge_domains: bool = False,
                 dynamic_edge_block_ratio: float = 0.0, collision_aware_reorder: bool = False,
                 collision_merge_across_types: bool = False,
                 duplicate_level_weight: float = 0.5, adaptive_fine_level_weight=None,
                 dense_collision_fine_level_weight: float = 3.0,
                 dense_collision_block_ratio: float = 0.1,
                 coarsest_level_weight: float = 2.5, level_weights=None,
                 cuda_enable_inverse_fallback: bool = True, mixed_spmv: bool = False,
                 mixed_spmv_tolerance_factor: float = 1.0, global_coarse_domain: bool = True,
                 domain_dof_schedule=None, target_node_schedule=None,
                 numeric_rebuild_interval: int = 1, dynamic_numeric_rebuild_interval: int = 1,
                 allow_cpu_fallback: bool = False, threshold_mode: str = "reference",
                 precision: str = "fp64", coarse_bank: str = "fp32"):
        if (any(v is not None and v <= 0 for v in (max_domain_dofs, max_levels,
                                                   target_nodes_per_partition))
                or fine_correction_iterations < 0):
            raise ValueError("domain capacity/levels must be positive and correction iterations non-negative")
        if int(minimum_domains_for_next_level) < 0:
            raise ValueError("minimum_domains_for_next_level must be non-negative")
        if cuda_fixed_inverse_bucket_size is not None and cuda_fixed_inverse_bucket_size <= 0:
            raise ValueError("cuda_fixed_inverse_bucket_size must be positive")
        if solve_mode not in ("mas_pcg", "reduced_then_fine"):
            raise ValueError("solve_mode must be 'mas_pcg' or 'reduced_then_fine'")
        if not np.isfinite(dynamic_edge_block_ratio) or dynamic_edge_block_ratio < 0.0:
            raise ValueError("dynamic_edge_block_ratio must be finite and non-negative")
        if not np.isfinite(duplicate_level_weight) or not 0.0 <= duplicate_level_weight <= 1.0:
            raise ValueError("duplicate_level_weight must be between zero and one")
        if not np.isfinite(dense_collision_fine_level_weight) or dense_collision_fine_level_weight <= 0:
            raise ValueError("dense collision fine-level weight must be positive")
        if not np.isfinite(dense_collision_block_ratio) or dense_collision_block_ratio <= 0:
            raise ValueError("dense collision block ratio must be positive")
        if not np.isfinite(coarsest_level_weight) or coarsest_level_weight <= 0:
            raise ValueError("coarsest_level_weight must be finite and positive")
        if level_weights is not None and (not level_weights or any(
                not np.isfinite(w) or w <= 0 for w in level_weights)):
            raise ValueError("level_weights must contain finite positive values")
        if not np.isfinite(mixed_spmv_tolerance_factor) or not 0.0 < mixed_spmv_tolerance_factor <= 1.0:
            raise ValueError("mixed_spmv_tolerance_factor must be in (0, 1]")
        for sched, what in ((domain_dof_schedule, "domain DOF"), (target_node_schedule, "target-node")):
            if sched is not None and (not sched or any(int(v) <= 0 for v in sched)):
                raise ValueError(f"{what} schedule must contain positive values")
        if int(numeric_rebuild_interval) <= 0:
            raise ValueError("numeric rebuild interval must be positive")
        if int(dynamic_numeric_rebuild_interval) <= 0:
            raise ValueError("dynamic numeric rebuild interval must be positive")
        if threshold_mode not in _MODES:
            raise ValueError(f"threshold_mode must be one of {sorted(_MODES)}")
        if precision not in ("fp64", "mixed"):
            raise ValueError("precision must be 'fp64' or 'mixed'")
        if coarse_bank not in ("fp32", "bf16"):
            raise ValueError("coarse_bank must be 'fp32' or 'bf16'")
        # device-path limits, named
        if max_domain_dofs is not None and max_domain_dofs > MAX_DOMAIN_SCALARS:
            raise ValueError(f"max_domain_dofs is limited to {MAX_DOMAIN_SCALARS} scalars on this device path")
        if domain_dof_schedule is not None and max(int(v) for v in domain_dof_schedule) > MAX_DOMAIN_SCALARS:
            raise ValueError(f"domain_dof_schedule entries are limited to {MAX_DOMAIN_SCALARS} scalars")
        if max_levels is not None and max_levels > _lib.MAS_MAX_COARSE_LEVELS + 1:
            raise ValueError(f"max_levels is limited to {_lib.MAS_MAX_COARSE_LEVELS + 1}")
        unsupported = [name for name, bad in (
            ("solve_mode='reduced_then_fine'", solve_mode != "mas_pcg"),
            ("dynamic_edge_domains", bool(dynamic_edge_domains)),
            ("collision_aware_reorder", bool(collision_aware_reorder)),
            ("collision_merge_across_types", bool(collision_merge_across_types)),
            (f"inverse_backend={inverse_backend!r}", inverse_backend != "cooperative_gauss_jordan"),
            ("allow_cpu_fallback (there is no CPU fallback)", bool(allow_cpu_fallback))) if bad]
        if unsupported:
            raise NotImplementedError("masSolver: not on this device path: " + ", ".join(unsupported))
        self.max_domain_dofs = int(max_domain_dofs or DOMAIN_SCALARS)
        self.max_levels = int(max_levels or 6)
        self.target_nodes_per_partition = int(target_nodes_per_partition or 8)
        self.minimum_domains_for_next_level = int(minimum_domains_for_next_level)
        self.domain_dof_schedule = None if domain_dof_schedule is None else tuple(int(v) for v in domain_dof_schedule)
        self.target_node_schedule = None if target_node_schedule is None else tuple(int(v) for v in target_node_schedule)
        self.coarsest_level_weight = float(coarsest_level_weight)
        self.level_weights = None if level_weights is None else tuple(float(w) for w in level_weights)
        self.duplicate_level_weight = float(duplicate_level_weight)
        self.adaptive_fine_level_weight = (level_weights is None if adaptive_fine_level_weight is None
                                           else bool(adaptive_fine_level_weight))
        self.dense_collision_fine_level_weight = float(dense_collision_fine_level_weight)
        self.dense_collision_block_ratio = float(dense_collision_block_ratio)
        self.global_coarse_domain = bool(global_coarse_domain)
        self.numeric_rebuild_interval = int(numeric_rebuild_interval)
        self.dynamic_numeric_rebuild_interval = int(dynamic_numeric_rebuild_interval)
        self.fine_correction_iterations = int(fine_correction_iterations)
        self.mixed_spmv_tolerance_factor = float(mixed_spmv_tolerance_factor)
        self.threshold_mode = threshold_mode
        self.precision = "mixed" if mixed_spmv else precision
        self.coarse_bank = coarse_bank
        self._refine = RefinementBuffers()
        self.reset()

    def reset(self):
        self._hierarchy = None
        self._hierarchy_key = None
        self._device = None
        self._device_key = None
        self._solution = None
        self._work = None
        self._scalars = None
        self._statistics = {}
        self.hierarchy_build_count = 0
        self.iterations = 0
        self._numeric_age = None
        self._numeric_dyn = 0

    @property
    def solution(self):
        return self._solution

    @property
    def statistics(self) -> dict:
        return dict(self._statistics)

    @property
    def hierarchy(self) -> Hierarchy:
        return self._hierarchy

    @property
    def device_tables(self) -> _DeviceMas:
        return self._device

    def _ensure(self, h, n):
        # every block 3x3, or the split 3x3 sub-blocks of an affine-body layout; the
        # hierarchy's nodes are the targets' instances (a 9-DoF affine matrix spans three
        # block rows), and the device tables take it over 3-scalar slots
        f = h.full if h.full is not None else h.split
        if f is None:
            raise NotImplementedError("masSolver: the device MAS path needs block dimensions "
                                      "that are multiples of 3")
        skey = f.key[:2]
        rebuilt = False
        if (self._hierarchy is None or self._hierarchy_key != skey
                or self._device_key != (id(f), f.version)):
            row_ptr = dev.to_host(f.row_ptr, "uint32").astype(np.int64)
            col = dev.to_host(f.col[:f.nnz], "uint32").astype(np.int64)
            if self._hierarchy is None or self._hierarchy_key != skey:
                split = f is h.split
                src = dev.to_host(f.src[:f.nnz], "int64" if split else "uint32")
                t0 = time.perf_counter()
                dims, types, node_of_row = target_nodes(h, n)
                pairs = node_of_row[static_pairs(row_ptr, col, src, 43 if split else 30)]
                self._hierarchy = build_hierarchy(
                    dims, pairs, self.max_domain_dofs,
                    self.max_levels, self.target_nodes_per_partition, self.global_coarse_domain,
                    minimum_domains_for_next_level=self.minimum_domains_for_next_level,
                    domain_dof_schedule=self.domain_dof_schedule,
                    target_node_schedule=self.target_node_schedule, type_ids=types)
                self._hierarchy_key = skey
                self.hierarchy_build_count += 1
                self._hierarchy_seconds = time.perf_counter() - t0
            slots = expand_to_slots(self._hierarchy)
            widest = max(3 * len(d) for lv in slots.levels for d in lv.domains)
            layout = build_layout(slots, row_ptr, col, padded_domain_size(widest))
            w = level_weights(self._hierarchy.num_levels, self.coarsest_level_weight, self.level_weights)
            layout.weights_base = w.copy()
            self._device = _DeviceMas(layout, w, n, self.coarse_bank, self.duplicate_level_weight)
            self._device_key = (id(f), f.version)
            self._numeric_age = None          # the numeric update must run
            rebuilt = True
        if self._solution is None or self._solution.numel() != n:
            self._solution = dev.zeros(n, "float64")
            self._work = dev.zeros(5 * n, "float64")
            self._scalars = dev.zeros(12288, "float64")     # YSP_PCG3_SCALARS
        return rebuilt

    def computeSolution(self, active_hessian, right_hand_side, initial_guess=None, tolerance=1e-3,
                        maxIterations=20000, zero_initial_guess=False):
        t0 = time.perf_counter()
        rhs = right_hand_side.value if isinstance(right_hand_side, vector) else right_hand_side
        if rhs is None or not hasattr(rhs, "is_cuda") or not rhs.is_cuda:
            raise TypeError("masSolver requires a device-resident right-hand side")
        n = int(rhs.numel())
        if active_hessian is None:
            self._solution = dev.zeros(n, "float64")
            self._statistics = {"solver": "mas", "converged": True, "iterations": 0,
                                "solve_seconds": 0.0}
            return 0
        if n != int(active_hessian.rows):
            raise ValueError("masSolver right-hand side size must match the matrix size")
        guess = None
        if not zero_initial_guess and initial_guess is not None:
            guess = initial_guess.value if hasattr(initial_guess, "value") else initial_guess
            if not hasattr(guess, "is_cuda") or int(guess.numel()) != n:
                raise ValueError("masSolver initial guess must be a matching device array")
        h = active_hessian
        rebuilt = self._ensure(h, n)
        t1 = time.perf_counter()
        m = self._device
        f = h.full if h.full is not None else h.split
        dyn = int(sum(h.block_counts_dynamic))
        # adaptive fine weight (solver/mas/cuda_runtime.py:2716-2731)
        w0 = float(m.layout.weights_base[0]) if m.layout.weights_base is not None else 1.0
        if self.adaptive_fine_level_weight and dyn / max(1, f.nbr) >= self.dense_collision_block_ratio:
            w0 = max(w0, self.dense_collision_fine_level_weight)
        m.set_fine_weight(w0)
        # lagged numeric updates (solver/mas/solver.py:520-568): the inverses of an earlier
        # SPD operator stay a valid preconditioner; rebuild when contact turns on or off or
        # the interval expires
        interval = self.dynamic_numeric_rebuild_interval if dyn else self.numeric_rebuild_interval
        transition = (dyn > 0) != (self._numeric_dyn > 0)
        if self._numeric_age is None or transition or self._numeric_age + 1 >= interval:
            m.update(f)
            self._numeric_age = 0
            self._numeric_dyn = dyn
            numeric_rebuilt = True
        else:
            self._numeric_age += 1
            numeric_rebuilt = False
        t2 = time.perf_counter()
        lib = _lib.load()
        stats = _lib.PcgStats()
        if self.precision == "mixed":
            return self._solve_mixed(h, rhs, n, tolerance, maxIterations, rebuilt, t0, t1, t2)
        vals, fmt = h.solve_values()
        rc = lib.ysp_pcg3_mas_solve(
            int(maxIterations), float(tolerance), _MODES[self.threshold_mode],
            dev.ptr(f.row_ptr), dev.ptr(f.col), vals, fmt, f.nnz, C.byref(m.args),
            dev.ptr(rhs), n, dev.ptr(self._solution), dev.ptr(guess), int(guess is None),
            dev.ptr(self._work), dev.ptr(self._scalars), C.byref(stats), dev.stream())
        _lib.check(rc, "ysp_pcg3_mas_solve")
        inverse_status = int(m.status.item())
        self._check_inverses(inverse_status)
        t3 = time.perf_counter()
        self.iterations = max(0, int(stats.iterations))
        lay = m.layout
        converged = rc >= 0
        self._statistics = {
            "solver": "mas", "converged": converged, "iterations": int(stats.iterations),
            "result": 0 if converged else -4, "pcg_result": int(rc),
            "delta0": stats.delta0, "delta": stats.delta, "rr0": stats.rr0, "rr": stats.rr,
            "hierarchy_build_count": self.hierarchy_build_count,
            "hierarchy_build_seconds": getattr(self, "_hierarchy_seconds", 0.0),
            "layout_rebuilt": rebuilt, "numeric_rebuilt": numeric_rebuilt,
            "fine_level_weight": float(m.weights_host[0]),
            "duplicate_domains": int(lay.duplicate.sum()) if lay.duplicate is not None else 0,
            "nodes_per_level": [lv.num_nodes for lv in self._hierarchy.levels],
            "domains_per_level": [len(lv.domains) for lv in self._hierarchy.levels],
            "inverse_status": inverse_status,
            "numeric_update_seconds": t2 - t1, "pcg_seconds": t3 - t2,
            "solve_seconds": t3 - t0, "matrix_size": n, "tolerance": float(tolerance),
        }
        return 0 if converged else -4

    def _check_inverses(self, inverse_status):
        """Raise when the last numeric update flagged a domain; the next solve rebuilds."""
        if inverse_status & 1:
            self._numeric_age = None
            raise LocalInverseError(
                "MAS local inversion failed: a domain matrix has a pivot at or below 1e-12 of "
                "its largest diagonal entry (the operator is singular or indefinite)")

    def _solve_mixed(self, h, rhs, n, tolerance, max_iterations, rebuilt, t0, t1, t2):
        lib = _lib.load()
        f = h.operator
        v32 = h.full_values_f32()
        m = self._device
        stats = _lib.PcgStats()
        s = dev.stream()

        def inner(r, d, tol_in, max_it):
            rc = lib.ysp_pcg3_mas_solve(int(max_it), float(tol_in), 1, dev.ptr(f.row_ptr), dev.ptr(f.col),
                                        dev.ptr(v32), 1, f.nnz, C.byref(m.args), dev.ptr(r), n,
                                        dev.ptr(d), None, 1, dev.ptr(self._work), dev.ptr(self._scalars),
                                        C.byref(stats), s)
            _lib.check(rc, "ysp_pcg3_mas_solve")
            return rc, int(stats.iterations)

        out = refine_solve(h, rhs, self._solution, float(tolerance), int(max_iterations), inner,
                           self._refine)
        inverse_status = int(m.status.item())
        self._check_inverses(inverse_status)
        t3 = time.perf_counter()
        self.iterations = out["iterations"]
        lay = m.layout
        self._statistics = {
            "solver": "mas", "precision": "mixed", "converged": out["converged"],
            "iterations": out["iterations"], "outer_iterations": out["outer_iterations"],
            "relative_residual": out["relative_residual"], "outer_history": out["outer_history"],
            "result": 0 if out["converged"] else -4,
            "hierarchy_build_count": self.hierarchy_build_count, "layout_rebuilt": rebuilt,
            "nodes_per_level": [lv.num_nodes for lv in self._hierarchy.levels],
            "domains_per_level": [len(lv.domains) for lv in self._hierarchy.levels],
            "inverse_status": inverse_status, "numeric_update_seconds": t2 - t1,
            "pcg_seconds": t3 - t2, "solve_seconds": t3 - t0, "matrix_size": n,
            "tolerance": float(tolerance),
        }
        return 0 if out["converged"] else -4
