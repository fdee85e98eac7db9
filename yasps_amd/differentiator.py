"""Second derivatives of energies: ``differentiator().diff2`` and the rectangular
``secondOrderJacobian`` (``differentiator/differentiator.pyx:321-489``,
``matrixAndVector/matrix/secondOrderJacobian.pyx:17-559``), as the reference's adjoint
examples use them (``examples/bunny_material_optimization/bunny_material_optimization.py:
588-634,1119-1128``):

* ``diff2([E...], wrt, wrt, projection_method=..., dynamic_instances=...)`` with equal
  target lists is the (projected) Hessian of the energies: a ``hessian`` whose terms are
  those energies; ``+`` sums such Hessians, ``compute()`` assembles them and the PCG of
  ``solver().computeSolution`` solves with them.
* ``diff2([E...], rows, columns)`` with different lists is the unprojected mixed second
  derivative d (d E / d rows) / d columns, ``rows x cols`` with the targets' layouts
  (segments in list order, ``size x instances`` scalars each). ``compute()`` evaluates it
  and ``spmv(v, transpose=...)`` multiplies by it; ``+`` adds terms.

Columns are attributes the energy's parameters are made of: its own scalar parameters
(Lamé coefficients, stiffnesses, barrier distance, time step, mass) and the attributes of
expressions that define them (``young`` and ``poisson`` through
``mu = young / (2 (1 + poisson))``, ``last_position`` through the inertia's previous
position and ``velocity = (last_position - last_last_position) / dt``; ``yasps_amd.expr``).
A column target has one instance (shared by every element) or one per element of the
energy. The block of an element is (d g / d p) (d p / d theta) + (d g / d s)
(d s / d theta): generated derivatives of the element gradient w.r.t. the energy's
parameters (``csrc/generated_mixed.hpp``) times the expressions' derivatives, formed by
the HIP kernel ``ysp_element_param_jacobian`` (for a run-time generated user energy, the
same kernel in its module, ``yasps_amd/usercode.py``). Not supported (NotImplementedError): energies
on affine-body vertices, ``quadratic_form``, columns that are DoF of the energy or
rest-shape inputs, and ``local_targets``."""
from __future__ import annotations

import ctypes as C
import functools
from typing import List, Optional

import numpy as np

from . import _lib, device as dev
from .core import attribute, COMPUTED
from .expr import derivative as expr_derivative
from .hessian import hessian, vector
from .program import encode


def _validate_targets(targets, name: str, who: str) -> List[attribute]:
    """``secondOrderJacobian.pyx:55-82``: a non-empty list of distinct, non-dynamic
    attributes."""
    if not isinstance(targets, list) or len(targets) == 0:
        raise ValueError(f"{who}: {name} must be a non-empty list.")
    out, seen = [], set()
    for t in targets:
        if not isinstance(t, attribute):
            raise TypeError(f"{who}: every {name} item must be an attribute.")
        if t.isDynamic:
            raise ValueError(f"{who}: {name} cannot contain dynamic attributes.")
        if t.hash in seen:
            raise ValueError(f"{who}: {name} contains a duplicate target.")
        seen.add(t.hash)
        out.append(t)
    return out


def _starts(targets) -> List[int]:
    s = [0]
    for t in targets:
        s.append(s[-1] + t.size * t.numInstances)
    return s


def _same_targets(a, b) -> bool:
    a = a if isinstance(a, list) else [a]
    b = b if isinstance(b, list) else [b]
    return len(a) == len(b) and all(x is y for x, y in zip(a, b))


def _energy(src) -> object:
    if not isinstance(src, attribute) or src.operator != COMPUTED or \
            not hasattr(src.expression, "program"):
        raise TypeError("differentiator.diff2: sources must be energy attributes built from "
                        "yasps_amd.energies")
    if src.size != 1:
        raise ValueError("differentiator.diff2: source must be a scalar attribute.")
    return src.expression


class _ColumnBlocks:
    """Element blocks of one energy for one column target (``keep``: the inputs of their
    kernel, held until the next compute)."""

    def __init__(self, blocks, w: int, shared: bool, col_start: int, keep):
        self.blocks, self.w, self.shared, self.col_start = blocks, w, shared, col_start
        self.keep = keep


class _Term:
    """One energy of a secondOrderJacobian: its row map and its blocks per column target."""

    def __init__(self, source: attribute, owner: "secondOrderJacobian"):
        self.source = source
        self.expr = _energy(source)
        self.owner = owner
        self.rows = None
        self.n = 0
        self.dofs = 0
        self.width = 3                  # DoF per row block (a vertex; 9 for a body matrix)
        self.columns: List[_ColumnBlocks] = []

    def _row_map(self, info) -> object:
        """Global row of every DoF of every element (get_indices on the energy's program
        with the row targets, then ysp_jacobian_rows)."""
        lib = _lib.load()
        s = dev.stream()
        o = self.owner
        prog = self.expr.program({w: i for i, w in enumerate(o.row_wrt)})
        conns, unions = [], []
        code = encode(prog, conns, unions)
        K = code[1]
        n = info.n
        code_dev = dev.to_device(np.asarray(code, np.int32), "int32")
        conn_ptrs = (C.c_void_p * max(1, len(conns)))(*[dev.ptr(c.value) for c in conns])
        union_ptrs = (C.c_void_p * max(1, len(unions)))(*[dev.ptr(u.counts_device) for u in unions])
        idx = dev.zeros(n * K, "uint32")
        sizes = dev.zeros(n * K, "uint16")
        _lib.check(lib.ysp_get_indices(dev.ptr(code_dev), len(code), conn_ptrs, len(conns),
                                       union_ptrs, len(unions), dev.ptr(o._row_start_dev),
                                       dev.ptr(idx), dev.ptr(sizes), n, K, s), "ysp_get_indices")
        rows = dev.empty(n * info.dofs, "int32")
        _lib.check(lib.ysp_jacobian_rows(dev.ptr(idx), dev.ptr(sizes), K, n, info.dofs,
                                         dev.ptr(rows), s), "ysp_jacobian_rows")
        return rows

    def compute(self):
        info = self.expr.mixed_info()
        self.n, self.dofs, self.width = info.n, info.dofs, int(info.positions.width)
        self.columns = []
        if info.n == 0:
            return
        o = self.owner
        torch = dev.torch()
        lib = _lib.load()
        s = dev.stream()
        self.rows = self._row_map(info)
        n = info.n
        dof_ids = {id(a) for a in info.dof_leaves}
        rest_ids = {id(a) for a in info.rest_leaves}
        for j, theta in enumerate(o.column_wrt):
            w = theta.size
            dp = [expr_derivative(a, theta, n, r) for a, r in info.p_args]
            ds = [None if a is None else expr_derivative(a, theta, n, 1) for a in info.s_args]
            if id(theta) in dof_ids:
                raise NotImplementedError(
                    f"secondOrderJacobian: column target {theta.fullName} is a DoF of "
                    f"{self.source.fullName}; diff2 with equal targets gives its Hessian")
            if id(theta) in rest_ids:
                raise NotImplementedError(
                    f"secondOrderJacobian: column target {theta.fullName} is a rest-shape input "
                    f"of {self.source.fullName}")
            if all(d is None for d in dp) and all(d is None for d in ds):
                continue                                # structurally zero
            if theta.numInstances == 1:
                shared = True
            elif theta.numInstances == n:
                shared = False
            else:
                raise NotImplementedError(
                    f"secondOrderJacobian: column target {theta.fullName} has "
                    f"{theta.numInstances} instances; {self.source.fullName} has {n} elements "
                    "(a column target has one instance or one per element)")
            if n > 1 and shared is False and w > 9:
                raise NotImplementedError("secondOrderJacobian: column targets of at most 9 values")

            def stack(parts, rows_list):
                if all(d is None for d in parts):
                    return None
                cols = [torch.zeros((n, r, w), dtype=torch.float64, device=dev.device())
                        if d is None else d for d, r in zip(parts, rows_list)]
                return torch.cat(cols, dim=1).contiguous()   # [n, slots, w]

            cp = stack(dp, [r for _, r in info.p_args]) if info.p_args else None
            cs = stack(ds, [1] * len(info.s_args))
            blocks = dev.empty(n * info.dofs * w, "float64")
            ep = info.ep
            kernel = (info.param_jacobian if info.param_jacobian is not None
                      else functools.partial(lib.ysp_element_param_jacobian, info.kind))
            _lib.check(kernel(
                C.byref(info.positions), dev.ptr(info.elem2v), n,
                dev.ptr(ep) if ep is not None else None, dev.ptr(info.se), info.se_stride,
                dev.ptr(cp) if cp is not None else None, cp.shape[1] * w if cp is not None else 0,
                dev.ptr(cs) if cs is not None else None, cs.shape[1] * w if cs is not None else 0,
                w, dev.ptr(blocks), s), "ysp_element_param_jacobian")
            self.columns.append(_ColumnBlocks(blocks, w, shared, o.column_start_indices[j],
                                              (info, cp, cs)))

    def spmv(self, x, y, transpose: bool):
        lib = _lib.load()
        s = dev.stream()
        for c in self.columns:
            if transpose:
                xin = x
                yout = y[c.col_start:]
            else:
                xin = x[c.col_start:]
                yout = y
            _lib.check(lib.ysp_jacobian_spmv(dev.ptr(c.blocks), self.n, self.dofs, c.w,
                                             dev.ptr(self.rows), int(c.shared), dev.ptr(xin),
                                             dev.ptr(yout), int(bool(transpose)), s),
                       "ysp_jacobian_spmv")


class secondOrderJacobian:
    """Rectangular mixed second derivative d (d E / d rows) / d columns
    (``matrixAndVector/matrix/secondOrderJacobian.pyx:17-559``)."""

    def __init__(self, row_wrt: List[attribute], column_wrt: List[attribute]):
        who = "secondOrderJacobian.__init__"
        self._row_wrt = _validate_targets(row_wrt, "row_wrt", who)
        self._column_wrt = _validate_targets(column_wrt, "column_wrt", who)
        self._row_starts = _starts(self._row_wrt)
        self._col_starts = _starts(self._column_wrt)
        self.rows = self._row_starts[-1]
        self.cols = self._col_starts[-1]
        self._terms: List[_Term] = []
        self._terms_dynamic: List[_Term] = []
        self._row_start_dev = None
        self._blocks_summary = None

    # ------------------------------------------------------------ layout
    @property
    def row_wrt(self) -> List[attribute]:
        return self._row_wrt

    @property
    def column_wrt(self) -> List[attribute]:
        return self._column_wrt

    @property
    def row_start_indices(self) -> List[int]:
        return self._row_starts

    @property
    def column_start_indices(self) -> List[int]:
        return self._col_starts

    @property
    def sources(self) -> List[attribute]:
        return [t.source for t in self._terms]

    @property
    def sources_dynamic(self) -> List[attribute]:
        return [t.source for t in self._terms_dynamic]

    def _add_source(self, source: attribute, dynamic: bool):
        (self._terms_dynamic if dynamic else self._terms).append(_Term(source, self))

    def __add__(self, other: "secondOrderJacobian") -> "secondOrderJacobian":
        """``secondOrderJacobian.pyx:293-326``: same row and column targets; the sum
        holds both operands' terms."""
        if not isinstance(other, secondOrderJacobian):
            raise TypeError("secondOrderJacobian.__add__: unsupported operand type(s) for +: "
                            f"'secondOrderJacobian' and '{type(other).__name__}'")
        if not _same_targets(self._row_wrt, other.row_wrt):
            raise ValueError("secondOrderJacobian.__add__: row_wrt mismatch.")
        if not _same_targets(self._column_wrt, other.column_wrt):
            raise ValueError("secondOrderJacobian.__add__: column_wrt mismatch.")
        out = secondOrderJacobian(self._row_wrt, self._column_wrt)
        for t in self._terms + other._terms:
            out._add_source(t.source, False)
        for t in self._terms_dynamic + other._terms_dynamic:
            out._add_source(t.source, True)
        return out

    # ------------------------------------------------------------ values
    def compute(self) -> "secondOrderJacobian":
        """Evaluate every term's element blocks at the current attribute values (dynamic
        terms with their current instances)."""
        if self._row_start_dev is None:
            self._row_start_dev = dev.to_device(np.asarray(self._row_starts, np.uint32), "uint32")
        for t in self._terms + self._terms_dynamic:
            t.compute()
        self._blocks_summary = None
        return self

    def _summary(self):
        """Distinct (row block, column block) pairs per block shape, as the reference's
        compressed storage holds them (``secondOrderJacobian.pyx:357-385``); read back
        from the device on first use after a compute."""
        if self._blocks_summary is None:
            pairs = {}
            for t in self._terms + self._terms_dynamic:
                if not t.columns:
                    continue
                r = dev.to_host(t.rows).reshape(t.n, t.dofs // t.width, t.width)[:, :, 0]
                for c in t.columns:
                    e = np.broadcast_to(np.arange(t.n)[:, None], r.shape)
                    ok = r >= 0
                    col = np.zeros_like(e) if c.shared else e
                    key = (t.width, c.w)
                    keys = r[ok].astype(np.int64) * (1 << 32) + (c.col_start + col[ok] * c.w)
                    pairs.setdefault(key, []).append(keys)
            self._blocks_summary = {k: int(np.unique(np.concatenate(v)).size)
                                    for k, v in pairs.items()}
        return self._blocks_summary

    @property
    def block_dimensions(self) -> List[List[int]]:
        return [list(k) for k in self._summary()]

    @property
    def block_counts(self) -> List[int]:
        return list(self._summary().values())

    def spmv(self, vec, transpose: bool = False) -> vector:
        """J v (``transpose=False``: v over the columns) or J^T v (v over the rows), a new
        vector (``matrix.spmv``, ``matrixAndVector/matrix/matrix.pyx:21-86``)."""
        x = vec.value if hasattr(vec, "value") else vec
        x = x.reshape(-1)
        n_in, n_out = (self.rows, self.cols) if transpose else (self.cols, self.rows)
        if x.numel() != n_in:
            raise ValueError(f"secondOrderJacobian.spmv: vector of {x.numel()} entries, expected "
                             f"{n_in}")
        if x.dtype != dev.torch().float64:
            x = x.to(dev.torch().float64)
        x = x.contiguous()
        out = vector(n_out)
        y = out.value
        y.zero_()
        for t in self._terms + self._terms_dynamic:
            t.spmv(x, y, transpose)
        return out


class differentiator:
    """``differentiator/differentiator.pyx``: ``diff2`` over the energies of
    ``yasps_amd.energies``."""

    def diff2(self, source, target1: List[attribute], target2: List[attribute],
              local_targets: List[attribute] = [], projection_method=1, save_intermediate=False,
              separate_hessian_jacobian=False, dynamic_instances=False, grouped_add=False,
              lto=False):
        """``differentiator.pyx:321-337``: equal target lists give the Hessian of the
        sources (projected with ``projection_method``), different ones the rectangular
        ``secondOrderJacobian`` (no projection). ``save_intermediate``,
        ``separate_hessian_jacobian``, ``grouped_add`` and ``lto`` select reference code
        generation variants and do not change the result here."""
        sources = source if isinstance(source, list) else [source]
        if len(sources) == 0:
            raise ValueError("differentiator.diff2: source can not be empty.")
        if local_targets:
            raise NotImplementedError("differentiator.diff2: local_targets are not supported")
        exprs = [_energy(s) for s in sources]
        if _same_targets(target1, target2):
            if not isinstance(grouped_add, bool):
                raise TypeError("differentiator.diff2: grouped_add must be bool.")
            if not isinstance(lto, bool):
                raise TypeError("differentiator.diff2: lto must be bool.")
            if projection_method not in (-1, 0, 1, 2):
                raise ValueError("differentiator.diff2: projection_method must be -1, 0, 1 or 2")
            wrt = target1 if isinstance(target1, list) else [target1]
            terms = [(e, projection_method) for e in exprs]
            dyn = dynamic_instances or any(s.isDynamic for s in sources)
            return hessian(wrt, [] if dyn else terms, terms if dyn else [])
        total: Optional[secondOrderJacobian] = None
        for s in sources:
            j = secondOrderJacobian(target1, target2)
            j._add_source(s, bool(dynamic_instances or s.isDynamic))
            total = j if total is None else total + j
        return total
