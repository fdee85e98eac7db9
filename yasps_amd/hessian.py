"""Assembled Hessian in upper block storage, its gradient, and device vectors.

Orchestration follows ``matrixAndVector/matrix/hessian.pyx``: the gradient layout is the
minimisation-target order (``:110-127``); static terms get their sparsity layout once
(``getSparseIndices`` ``:783-808``), dynamic terms are re-indexed on every ``compute``
into an independent layout (``:846-880``); ``compute`` zeroes gradient, diagonal blocks
and block storage, then runs every term's element kernel (``:1159-1207``). Every index
stage runs on the device through the C ABI.
"""
from __future__ import annotations

import itertools

import ctypes as C
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import _lib, device as dev
from .program import encode


def check_deferred():
    """Synchronise the current stream and raise if a queued kernel flag was set."""
    _lib.check(_lib.load().ysp_check_deferred(dev.stream()), "element evaluation")


class vector:
    """A device FP64 vector (``matrixAndVector/vector``)."""

    def __init__(self, size: int = 0, value=None):
        self.value = dev.zeros(size, "float64") if value is None else value

    @property
    def size(self) -> int:
        return int(self.value.numel())

    def updateValue(self, values):
        t = dev.torch()
        if isinstance(values, t.Tensor):
            self.value = values.reshape(-1).to(dev.device(), t.float64)
        else:
            self.value = dev.to_device(np.asarray(values, np.float64).reshape(-1), "float64")
        return self

    def get(self) -> np.ndarray:
        return dev.to_host(self.value)


@dataclass
class TermIndex:
    """Device index buffers of one term (gradientIndicesKernel outputs)."""
    K: int
    n: int
    indices: object = None
    sizes: object = None
    permutations: object = None
    total_sizes: object = None
    coord_outer: object = None
    unique_sizes: object = None
    grouped: object = None
    grouped_outer: object = None
    num_unique: object = None
    coords: object = None
    dims: object = None
    num_coords: int = 0
    lookups: object = None
    g_local: int = 0


@dataclass
class Part:
    """Block storage of one (static or dynamic) part."""
    values: object = None
    positions: object = None
    starts: List[int] = field(default_factory=list)
    counts: List[int] = field(default_factory=list)
    dims: List[int] = field(default_factory=list)
    total: int = 0
    num_blocks: int = 0
    ordinals: object = None
    staging: object = None          # 9 doubles per coordinate (staged assembly)
    contrib_offsets: object = None
    contrib_items: object = None
    coord_bases: List[int] = field(default_factory=list)
    _ct: tuple = ()

    def ctype(self) -> Optional[_lib.BlockMatrix]:
        if not self.counts:
            return None
        starts = _lib.u32_array(self.starts[:-1] if len(self.starts) > len(self.counts) else self.starts)
        counts = _lib.u32_array(self.counts)
        dims = _lib.u32_array(self.dims)
        m = _lib.BlockMatrix(dev.ptr(self.values), dev.ptr(self.positions), len(self.counts),
                             starts, counts, dims)
        self._ct = (starts, counts, dims, m)
        return m


def build_term_index(program, n: int, wrt_start_dev, stream) -> TermIndex:
    """get_indices -> compress_indices -> computeCoordinates for one term."""
    lib = _lib.load()
    conns, unions = [], []
    code = encode(program, conns, unions)
    K = code[1]
    ti = TermIndex(K=K, n=n, g_local=code[2])
    if n == 0:
        return ti
    code_dev = dev.to_device(np.asarray(code, np.int32), "int32")
    conn_ptrs = (C.c_void_p * max(1, len(conns)))(*[dev.ptr(c.value) for c in conns])
    union_ptrs = (C.c_void_p * max(1, len(unions)))(*[dev.ptr(u.counts_device) for u in unions])
    ti.indices = dev.zeros(n * K, "uint32")
    ti.sizes = dev.zeros(n * K, "uint16")
    _lib.check(lib.ysp_get_indices(dev.ptr(code_dev), len(code), conn_ptrs, len(conns), union_ptrs,
                                   len(unions), dev.ptr(wrt_start_dev), dev.ptr(ti.indices),
                                   dev.ptr(ti.sizes), n, K, stream), "ysp_get_indices")
    ti.permutations = dev.zeros(n * K, "int16")
    ti.total_sizes = dev.empty(n, "uint16")
    ti.coord_outer = dev.zeros(n + 1, "uint32")
    ti.unique_sizes = dev.zeros(n + 1, "uint16")
    ti.grouped = dev.empty(n, "uint32")
    ti.grouped_outer = dev.zeros(n + 2, "uint32")
    ti.num_unique = dev.zeros(1, "uint16")
    _lib.check(lib.ysp_compress_indices(
        dev.ptr(ti.indices), dev.ptr(ti.sizes), dev.ptr(ti.permutations), dev.ptr(ti.total_sizes),
        dev.ptr(ti.coord_outer), dev.ptr(ti.unique_sizes), dev.ptr(ti.grouped),
        dev.ptr(ti.grouped_outer), dev.ptr(ti.num_unique), n, K, 1, stream), "ysp_compress_indices")
    ti.num_coords = int(dev.to_host(ti.coord_outer[n:n + 1], "uint32")[0])
    ti.coords = dev.empty(max(1, 2 * ti.num_coords), "uint32")
    ti.dims = dev.empty(max(1, 2 * ti.num_coords), "uint16")
    if ti.num_coords:
        _lib.check(lib.ysp_compute_coordinates(
            dev.ptr(ti.indices), dev.ptr(ti.permutations), dev.ptr(ti.sizes), dev.ptr(ti.coords),
            dev.ptr(ti.dims), dev.ptr(ti.coord_outer), K, n, stream), "ysp_compute_coordinates")
    return ti


def build_part(terms: List[TermIndex], stream) -> Part:
    """get_unique_coords + compress_unique_coords over the concatenated coordinates
    (``coordinateCompressionKernel.pyx:692-726``)."""
    lib = _lib.load()
    t = dev.torch()
    total = sum(ti.num_coords for ti in terms)
    part = Part()
    if total == 0:
        for ti in terms:
            ti.lookups = dev.empty(0, "uint32")
        return part
    coords = t.cat([ti.coords[:2 * ti.num_coords] for ti in terms if ti.num_coords])
    dims = t.cat([ti.dims[:2 * ti.num_coords] for ti in terms if ti.num_coords])
    tmp = dev.empty(4 * total, "uint32")   # 16-byte records
    lookup = dev.empty(total, "uint32")
    nu = C.c_uint32(0)
    _lib.check(lib.ysp_get_unique_coords(dev.ptr(coords), dev.ptr(dims), total, dev.ptr(tmp),
                                         dev.ptr(lookup), C.byref(nu), stream), "ysp_get_unique_coords")
    nu = nu.value
    part.ordinals = lookup.clone()          # coordinate -> unique block ordinal
    positions = dev.empty(2 * nu, "uint32")
    max_cat = 64
    udims = dev.zeros(2 * max_cat, "uint16")
    outer = dev.zeros(max_cat + 1, "uint32")
    counts = dev.zeros(max_cat, "uint32")
    nd = C.c_uint32(0)
    _lib.check(lib.ysp_compress_unique_coords(dev.ptr(tmp), nu, total, dev.ptr(positions),
                                              dev.ptr(udims), dev.ptr(outer), dev.ptr(counts),
                                              dev.ptr(lookup), C.byref(nd), stream),
               "ysp_compress_unique_coords")
    nd = nd.value
    part.positions = positions
    part.dims = [int(v) for v in dev.to_host(udims[:2 * nd], "uint16")]
    part.starts = [int(v) for v in dev.to_host(outer[:nd + 1], "uint32")]
    part.counts = [int(v) for v in dev.to_host(counts[:nd], "uint32")]
    part.total = part.starts[-1]
    part.num_blocks = nu
    part.values = dev.zeros(part.total, "float64")
    off = 0
    for ti in terms:
        ti.lookups = lookup[off:off + ti.num_coords]
        off += ti.num_coords
    return part


_STRUCTURE_VERSIONS = itertools.count(1)


@dataclass
class FullBsr:
    """Mirrored full 3x3 block-CSR of the current layout (the solve loop's operator)."""
    key: tuple
    nbr: int
    row_ptr: object
    col: object
    src: object
    nnz: int
    v64: object = None
    v32: object = None
    v32_valid: bool = False      # v32 holds the current values
    # changes only when the structure does (a rebuild onto the same pairs keeps it)
    version: int = field(default_factory=lambda: next(_STRUCTURE_VERSIONS))


def _same_structure(old: Optional[FullBsr], new: FullBsr) -> bool:
    """The rebuilt block-CSR equals the previous one (a dynamic term re-indexed onto the
    same pairs): the previous object, and with it its structure version, is kept, so
    structures derived from it (the MAS device layout) are not rebuilt."""
    if old is None or old.nnz != new.nnz or old.nbr != new.nbr:
        return False
    t = dev.torch()
    k = new.nnz
    return bool(t.equal(old.row_ptr, new.row_ptr) and t.equal(old.col[:k], new.col[:k])
                and t.equal(old.src[:k], new.src[:k]))


class hessian:
    """Sum of energy terms w.r.t. a list of targets, in symmetric upper block storage."""

    def __init__(self, wrt, terms, dynamic_terms):
        self.wrt = list(wrt)
        self.terms = list(terms)                 # [(expr, projection_method)]
        self.dynamic_terms = list(dynamic_terms)
        sizes = [w.size * w.numInstances for w in self.wrt]
        self.gradient_segments_start_cpu = [0]
        self.diagonal_blocks_start_cpu = [0]
        for w, s in zip(self.wrt, sizes):
            self.gradient_segments_start_cpu.append(self.gradient_segments_start_cpu[-1] + s)
            self.diagonal_blocks_start_cpu.append(self.diagonal_blocks_start_cpu[-1]
                                                  + w.size * w.size * w.numInstances)
        self.wrt_start_indices = self.gradient_segments_start_cpu[:-1]
        self.rows = self.cols = self.gradient_segments_start_cpu[-1]
        if self.rows <= 0:
            raise ValueError("hessian: total size must be positive")
        self._wrt_pos = {w: i for i, w in enumerate(self.wrt)}
        self._setup = False
        self.static_index: List[TermIndex] = []
        self.dynamic_index: List[TermIndex] = []
        self.static = Part()
        self.dynamic = Part()
        self.full: Optional[FullBsr] = None
        self.split: Optional[FullBsr] = None     # 3x3 sub-blocks of 9x9 / 9x3 / 3x9 layouts
        self.want_f32 = False
        # where the solve loop reads FP64 operator values: "auto" (the upper storage when
        # the mirrored copy would not stay in the Infinity Cache), "mirrored" or "upper"
        self.value_source = "auto"
        self._upper = None
        self.use_staging = True
        self._grad_offsets = None

    def __add__(self, other: "hessian") -> "hessian":
        """Sum of two Hessians over the same targets (the adjoint Hessian of
        ``bunny_material_optimization.py:581-603``, a sum of ``diff2`` terms): a new
        Hessian that assembles the terms of both."""
        if not isinstance(other, hessian):
            raise TypeError(f"hessian.__add__: unsupported operand type(s) for +: 'hessian' and "
                            f"'{type(other).__name__}'")
        if len(self.wrt) != len(other.wrt) or any(a is not b for a, b in zip(self.wrt, other.wrt)):
            raise ValueError("hessian.__add__: wrt mismatch.")
        return hessian(self.wrt, self.terms + other.terms, self.dynamic_terms + other.dynamic_terms)

    # --------------------------------------------------------------- layout
    def _ensure_setup(self):
        if self._setup:
            return
        s = dev.stream()
        self._wrt_start_dev = dev.to_device(np.asarray(self.gradient_segments_start_cpu, np.uint32),
                                            "uint32")
        self._diag_start_dev = dev.to_device(np.asarray(self.diagonal_blocks_start_cpu, np.uint32),
                                             "uint32")
        self.gradient = vector(self.rows)
        self.diagonal_blocks = dev.zeros(self.diagonal_blocks_start_cpu[-1], "float64")
        self.diagonal_blocks_inverse = dev.zeros(self.diagonal_blocks_start_cpu[-1], "float64")
        self.static_index = [build_term_index(e.program(self._wrt_pos), e.num_instances,
                                              self._wrt_start_dev, s) for e, _ in self.terms]
        self.static = build_part(self.static_index, s)
        self._setup = True
        self._prepare_staging(self.static, self.static_index)
        if not self.dynamic_terms:
            self._prepare_gradient_staging()

    def _reindex_dynamic(self):
        if not self.dynamic_terms:
            return
        s = dev.stream()
        self.dynamic_index = [build_term_index(e.program(self._wrt_pos), e.num_instances,
                                               self._wrt_start_dev, s) for e, _ in self.dynamic_terms]
        self.dynamic = build_part(self.dynamic_index, s)
        self._prepare_staging(self.dynamic, self.dynamic_index)
        self._prepare_gradient_staging()

    # --------------------------------------------------------------- staged assembly
    @property
    def staged(self) -> bool:
        """Atomic-free assembly (3x3 blocks and 3-vector gradient segments only)."""
        return self.use_staging and self.all_3x3 and self._grad_offsets is not None

    def _prepare_staging(self, part: Part, index: List[TermIndex]):
        part.staging = None
        if not (self.use_staging and self.all_3x3) or part.num_blocks == 0:
            return
        lib = _lib.load()
        total = sum(ti.num_coords for ti in index)
        bases, acc = [], 0
        for ti in index:
            bases.append(acc)
            acc += ti.num_coords
        part.coord_bases = bases
        part.staging = dev.empty(9 * total, "float64")
        part.contrib_offsets = dev.empty(part.num_blocks + 1, "uint32")
        part.contrib_items = dev.empty(max(1, total), "uint32")
        _lib.check(lib.ysp_inverse_map(dev.ptr(part.ordinals), None, total, part.num_blocks,
                                       dev.ptr(part.contrib_offsets), dev.ptr(part.contrib_items),
                                       dev.stream()), "ysp_inverse_map")

    def _prepare_gradient_staging(self):
        self._grad_offsets = None
        if not (self.use_staging and self.all_3x3):
            return
        lib = _lib.load()
        t = dev.torch()
        s = dev.stream()
        index = self.static_index + self.dynamic_index
        keys, vals, bases, acc = [], [], [], 0
        try:
            for ti in index:
                bases.append(acc)
                if ti.n == 0:
                    continue
                k = dev.empty(ti.n * ti.K, "uint32")
                v = dev.empty(ti.n * ti.K, "uint32")
                _lib.check(lib.ysp_gradient_contributions(dev.ptr(ti.indices), dev.ptr(ti.sizes),
                                                          ti.n, ti.K, acc, ti.g_local, dev.ptr(k),
                                                          dev.ptr(v), s), "ysp_gradient_contributions")
                keys.append(k)
                vals.append(v)
                acc += ti.n * ti.g_local
        except _lib.YspError:
            return      # segments that are not 3-vectors: keep the atomic path
        self._grad_bases = bases
        self._grad_staging = dev.empty(max(1, acc), "float64")
        nslots = sum(k.numel() for k in keys)
        self._grad_offsets = dev.empty(self.rows // 3 + 1, "uint32")
        self._grad_items = dev.empty(max(1, nslots), "uint32")
        if nslots:
            kk, vv = t.cat(keys), t.cat(vals)
            _lib.check(lib.ysp_inverse_map(dev.ptr(kk), dev.ptr(vv), nslots, self.rows // 3,
                                           dev.ptr(self._grad_offsets), dev.ptr(self._grad_items), s),
                       "ysp_inverse_map")
        else:
            self._grad_offsets.zero_()

    # --------------------------------------------------------------- storage views
    @property
    def blocks_flattened(self):
        return self.static.values

    @property
    def block_positions(self):
        return self.static.positions

    @property
    def blocks_start_indices(self):
        return self.static.starts

    @property
    def block_counts(self):
        return self.static.counts

    @property
    def block_dimensions(self):
        return self.static.dims

    @property
    def blocks_flattened_dynamic(self):
        return self.dynamic.values

    @property
    def block_positions_dynamic(self):
        return self.dynamic.positions

    @property
    def blocks_start_indices_dynamic(self):
        return self.dynamic.starts

    @property
    def block_counts_dynamic(self):
        return self.dynamic.counts

    @property
    def block_dimensions_dynamic(self):
        return self.dynamic.dims

    @property
    def diagonal(self):
        """Scalar diagonal of the assembled Hessian (``hessian.pyx:188``), taken from the
        diagonal blocks every assembly fills."""
        if not self._setup:
            return dev.zeros(0, "float64")
        out = dev.zeros(self.rows, "float64")
        lib = _lib.load()
        sizes = [w.size for w in self.wrt]
        counts = [w.numInstances for w in self.wrt]
        _lib.check(lib.ysp_diagonal_from_blocks(
            dev.ptr(self.diagonal_blocks), dev.ptr(out),
            _lib.u32_array(self.diagonal_blocks_start_cpu[:-1]), _lib.u32_array(counts),
            _lib.u32_array(sizes), _lib.u32_array(self.gradient_segments_start_cpu[:-1]),
            len(self.wrt), dev.stream()), "ysp_diagonal_from_blocks")
        return out

    # --------------------------------------------------------------- numeric
    def _scatter_args(self, ti: TermIndex, part: Part) -> _lib.ScatterArgs:
        return _lib.ScatterArgs(dev.ptr(ti.indices), dev.ptr(ti.sizes), dev.ptr(ti.permutations),
                                dev.ptr(ti.lookups), dev.ptr(ti.coord_outer), ti.K, len(self.wrt),
                                dev.ptr(self._diag_start_dev), dev.ptr(self._wrt_start_dev),
                                dev.ptr(self.gradient.value), dev.ptr(part.values),
                                dev.ptr(self.diagonal_blocks))

    def compute(self, local_gradient=None, check: bool = True):
        """Evaluate, project and assemble every term. The element kernels' flags (slot
        layouts they do not implement) are examined at the end (check=True: synchronises
        the stream) or by the caller's later ``check_deferred`` (a Newton step checks
        after its solve, so that assembly and solve run without a host round trip)."""
        self._ensure_setup()
        if local_gradient is not None and local_gradient is not self.gradient:
            self.gradient = local_gradient
        self._reindex_dynamic()
        if self.staged:
            self._compute_staged()
        else:
            self._compute_atomic()
        self._update_full()
        if check:
            check_deferred()
        return self

    def _compute_atomic(self):
        self.gradient.value.zero_()
        self.diagonal_blocks.zero_()
        if self.static.values is not None:
            self.static.values.zero_()
        if self.dynamic.values is not None:
            self.dynamic.values.zero_()
        for (expr, method), ti in zip(self.terms, self.static_index):
            if ti.n:
                expr.evaluate(self._scatter_args(ti, self.static), method)
        for (expr, method), ti in zip(self.dynamic_terms, self.dynamic_index):
            if ti.n:
                expr.evaluate(self._scatter_args(ti, self.dynamic), method)

    def _compute_staged(self):
        lib = _lib.load()
        s = dev.stream()
        self.diagonal_blocks.zero_()
        gptr = dev.ptr(self._grad_staging)
        parts = ((self.terms, self.static_index, self.static),
                 (self.dynamic_terms, self.dynamic_index, self.dynamic))
        k = 0
        for terms, index, part in parts:
            for (expr, method), ti, base in zip(terms, index, part.coord_bases or [0] * len(index)):
                if ti.n:
                    sa = self._scatter_args(ti, part)
                    sa.block_staging = dev.ptr(part.staging) + 8 * 9 * base
                    sa.gradient_staging = gptr + 8 * self._grad_bases[k]
                    sa.gradient_local_size = ti.g_local
                    expr.evaluate(sa, method)
                k += 1
        for i, (_, index, part) in enumerate(parts):
            if part.num_blocks:
                _lib.check(lib.ysp_gather_blocks3(
                    dev.ptr(part.staging), dev.ptr(part.contrib_offsets), dev.ptr(part.contrib_items),
                    part.num_blocks, dev.ptr(part.positions), dev.ptr(part.values),
                    dev.ptr(self.diagonal_blocks), dev.ptr(self._diag_start_dev),
                    dev.ptr(self._wrt_start_dev), len(self.wrt), i, s), "ysp_gather_blocks3")
        _lib.check(lib.ysp_gather_gradient3(gptr, dev.ptr(self._grad_offsets), dev.ptr(self._grad_items),
                                            self.rows, dev.ptr(self.gradient.value), s),
                   "ysp_gather_gradient3")

    # --------------------------------------------------------------- full block-CSR
    @property
    def all_3x3(self) -> bool:
        return all(w.size == 3 for w in self.wrt) and all(
            d == 3 for part in (self.static, self.dynamic) for d in part.dims)

    @property
    def splittable(self) -> bool:
        """Every target segment and block dimension is a multiple of 3 (at most 12)."""
        return all(w.size % 3 == 0 for w in self.wrt) and all(
            d % 3 == 0 and 0 < d <= 12 for part in (self.static, self.dynamic) for d in part.dims)

    def _update_split(self):
        """The mirrored full 3x3 block-CSR of a layout with affine-body segments (every
        block split into 3x3 sub-blocks); the solve loop's operator there."""
        if not self.splittable:
            self.split = None
            return
        lib = _lib.load()
        s = dev.stream()
        key = (id(self.static.positions), self.static.num_blocks,
               id(self.dynamic.positions), self.dynamic.num_blocks)
        if self.split is None or self.split.key != key:
            cap = 0
            for part in (self.static, self.dynamic):
                for c, cnt in enumerate(part.counts):
                    cap += 2 * cnt * (part.dims[2 * c] // 3) * (part.dims[2 * c + 1] // 3)
            cap = max(1, cap)
            nbr = self.rows // 3
            row_ptr = dev.empty(nbr + 1, "uint32")
            col = dev.empty(cap, "uint32")
            src = dev.empty(cap, "int64")
            nnz = C.c_uint32(0)
            a, b = self.static.ctype(), self.dynamic.ctype()
            _lib.check(lib.ysp_bsr3_build_split(C.byref(a) if a else None, C.byref(b) if b else None,
                                                self.rows, dev.ptr(row_ptr), dev.ptr(col),
                                                dev.ptr(src), cap, C.byref(nnz), s),
                       "ysp_bsr3_build_split")
            new = FullBsr(key, nbr, row_ptr, col, src, nnz.value)
            if _same_structure(self.split, new):
                self.split.key = key
            else:
                self.split = new
        f = self.split
        stride = (f.nnz + 63) // 64 * 64
        if f.v64 is None or f.v64.numel() != 9 * stride:
            f.v64 = dev.empty(9 * stride, "float64")
        _lib.check(lib.ysp_bsr3_gather_split(dev.ptr(self.static.values), dev.ptr(self.dynamic.values),
                                             dev.ptr(f.src), f.nnz, dev.ptr(f.v64), s),
                   "ysp_bsr3_gather_split")
        f.v32_valid = False                 # FP32 values are cast from these on demand

    def _update_full(self):
        if not self.all_3x3:
            self.full = None
            self._update_split()
            return
        self.split = None
        lib = _lib.load()
        s = dev.stream()
        key = (id(self.static.positions), self.static.num_blocks,
               id(self.dynamic.positions), self.dynamic.num_blocks)
        if self.full is None or self.full.key != key:
            cap = max(1, 2 * (self.static.num_blocks + self.dynamic.num_blocks))
            nbr = self.rows // 3
            row_ptr = dev.empty(nbr + 1, "uint32")
            col = dev.empty(cap, "uint32")
            src = dev.empty(cap, "uint32")
            nnz = C.c_uint32(0)
            _lib.check(lib.ysp_bsr3_build(dev.ptr(self.static.positions), self.static.num_blocks,
                                          dev.ptr(self.dynamic.positions), self.dynamic.num_blocks,
                                          self.rows, dev.ptr(row_ptr), dev.ptr(col), dev.ptr(src),
                                          cap, C.byref(nnz), s), "ysp_bsr3_build")
            new = FullBsr(key, nbr, row_ptr, col, src, nnz.value)
            if _same_structure(self.full, new):
                self.full.key = key
            else:
                self.full = new
        f = self.full
        stride = (f.nnz + 63) // 64 * 64
        if f.v64 is None or f.v64.numel() != 9 * stride:
            f.v64 = dev.empty(9 * stride, "float64")
        if self.want_f32 and (f.v32 is None or f.v32.numel() != 9 * stride):
            f.v32 = dev.empty(9 * stride, "float32")
        _lib.check(lib.ysp_bsr3_gather_values(dev.ptr(self.static.values), dev.ptr(self.dynamic.values),
                                              dev.ptr(f.src), f.nnz, dev.ptr(f.v64),
                                              dev.ptr(f.v32) if self.want_f32 else None, s),
                   "ysp_bsr3_gather_values")
        f.v32_valid = bool(self.want_f32)

    @property
    def operator(self) -> Optional[FullBsr]:
        """The solve loop's mirrored full 3x3 block-CSR: ``full`` for 3x3 layouts, ``split``
        for layouts with affine-body segments (None for other layouts)."""
        return self.full if self.full is not None else self.split

    # mirrored FP64 copies above this many bytes leave too little of the 256 MB Infinity
    # Cache for the rest of the loop's working set
    UPPER_VALUE_BYTES = 160 << 20

    def solve_values(self):
        """(pointer, format) of the FP64 operator values the solve loop reads: the
        component-major mirrored copy (``YSP_VALUES_F64``) or the upper block storage read
        through each entry's source word (``YSP_VALUES_UPPER``, a ``ysp_upper_values`` held
        by this object), chosen by ``value_source``. Both give the same products."""
        f = self.operator
        if f is None:
            raise NotImplementedError("the solve loop's value sources need 3 x 3 blocks or "
                                      "affine-body segments of multiples of 3")
        if self.value_source not in ("auto", "mirrored", "upper"):
            raise ValueError("value_source must be 'auto', 'mirrored' or 'upper'")
        # "auto": only 3x3 layouts without a dynamic part, whose upper blocks of a block row
        # lie next to each other (the split layout's categories and the dynamic part scatter
        # a row's sources, and its reads then cost more than the mirrored copy's)
        upper = self.value_source == "upper" or (
            self.value_source == "auto" and self.full is not None and self.dynamic.values is None
            and 72 * f.nnz > self.UPPER_VALUE_BYTES)
        if not upper:
            return dev.ptr(f.v64), _lib.VALUES_F64
        sv = self.static.values if self.static.values is not None else self.dynamic.values
        dv = self.dynamic.values
        self._upper = _lib.UpperValues(dev.ptr(f.src), int(self.full is None), dev.ptr(sv),
                                       dev.ptr(dv) if dv is not None else None)
        return C.addressof(self._upper), _lib.VALUES_UPPER

    def full_values_f32(self):
        """FP32 copy of the full (or split) block-CSR values (gathered with the FP64 values
        when ``want_f32`` is set, else cast on demand)."""
        f = self.operator
        if f is None:
            raise NotImplementedError("FP32 block values need 3 x 3 blocks or affine-body "
                                      "segments of multiples of 3")
        stride = (f.nnz + 63) // 64 * 64
        if f.v32 is None or f.v32.numel() != 9 * stride:
            f.v32 = dev.empty(9 * stride, "float32")
            f.v32_valid = False
        if not f.v32_valid:
            _lib.check(_lib.load().ysp_cast_f32(dev.ptr(f.v64), dev.ptr(f.v32), 9 * stride, dev.stream()),
                       "ysp_cast_f32")
            f.v32_valid = True
        return f.v32

    def spmv(self, vec: vector, transpose: bool = False, use_f32: bool = False) -> vector:
        self._ensure_setup()
        out = vector(self.rows)
        if (self.full is not None or (self.split is not None and not use_f32)) and not transpose:
            f = self.full if self.full is not None else self.split
            vals = f.v32 if use_f32 else f.v64
            _lib.check(_lib.load().ysp_bsr3_spmv(dev.ptr(f.row_ptr), dev.ptr(f.col), dev.ptr(vals),
                                                 int(use_f32), f.nnz, self.rows, dev.ptr(vec.value),
                                                 dev.ptr(out.value), dev.stream()), "ysp_bsr3_spmv")
            return out
        a, b = self.static.ctype(), self.dynamic.ctype()
        _lib.check(_lib.load().ysp_block_spmv(C.byref(a) if a else None, C.byref(b) if b else None,
                                              dev.ptr(vec.value), dev.ptr(out.value), self.rows,
                                              int(transpose), 0, dev.stream()), "ysp_block_spmv")
        return out

    # --------------------------------------------------------------- host views
    def dense(self) -> np.ndarray:
        """Dense symmetric matrix (small problems; tests)."""
        n = self.rows
        out = np.zeros((n, n))
        for part in (self.static, self.dynamic):
            if not part.counts:
                continue
            pos = dev.to_host(part.positions, "uint32").reshape(-1, 2)
            vals = dev.to_host(part.values)
            k = 0
            for c, cnt in enumerate(part.counts):
                h, w = part.dims[2 * c], part.dims[2 * c + 1]
                for b in range(cnt):
                    r, cc = map(int, pos[k])
                    o = part.starts[c] + b * h * w
                    blk = vals[o:o + h * w].reshape(h, w)
                    out[r:r + h, cc:cc + w] += blk
                    if r != cc:
                        out[cc:cc + w, r:r + h] += blk.T
                    k += 1
        return out
