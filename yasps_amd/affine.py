"""Affine-body vertex maps: ``x = t + A r`` for the vertices of a body
(examples/one_bunny_partial_abd/one_bunny_partial_abd.py:127-144).

``translation + matrix * rest_position`` on JOIN attributes (the body's t and A through a
vertex -> body connectivity of dimension 1) and a per-vertex rest position builds an
``AffineMap``; bound with ``addAttribute(name, computed_attribute=...)`` it is the
vertices' current position. Energies that reach it (through a union or directly) get
the body's A (9) and t (3) segments as their slots and use the chain rule on the device.
"""
from __future__ import annotations

import numpy as np

from . import _lib, device as dev
from .core import attribute, COMPUTED, CONSTANT, DATA, JOIN


class MatrixVector:
    def __init__(self, matrix, vector):
        self.matrix = matrix
        self.vector = vector

    def __radd__(self, other):
        return AffineMap(other, self.matrix, self.vector)


class AffineMap:
    kind = "affine_map"

    def __init__(self, translation, matrix, rest):
        for a, nm, rows, cols in ((translation, "translation", 3, 1), (matrix, "matrix", 3, 3)):
            if not isinstance(a, attribute) or a.operator != JOIN or a.through.dimension != 1 or \
                    a.rows * a.cols != rows * cols:
                raise TypeError(f"affine map: {nm} must be a {rows}x{cols} JOIN through a "
                                "vertex -> body connectivity of dimension 1")
        if translation.through is not matrix.through:
            raise ValueError("affine map: translation and matrix must share the connectivity")
        if not isinstance(rest, attribute) or rest.size != 3 or rest.operator not in (DATA, CONSTANT):
            raise TypeError("affine map: rest position must be a 3-vector DATA/CONSTANT attribute")
        if rest.correspondance is not translation.correspondance:
            raise ValueError("affine map: rest position must live on the vertex primitive")
        if matrix.source.size != 9 or translation.source.size != 3:
            raise TypeError("affine map: the body attributes must hold 9 and 3 values")
        self.t, self.A, self.r = translation, matrix, rest
        self._buf = None

    def bind(self, name: str, owner) -> attribute:
        if owner is not self.t.correspondance:
            raise ValueError("affine map: bind the position to the vertex primitive")
        att = attribute(name, owner, 3, 1, COMPUTED, expression=self)
        self.attribute = att
        return att

    @property
    def conn(self):
        return self.t.through

    @property
    def body_matrix(self) -> attribute:
        return self.A.source

    @property
    def body_translation(self) -> attribute:
        return self.t.source

    def values(self, att=None):
        """Current positions t_b + A_b r_v on the device."""
        n = self.t.correspondance.numInstances
        if self._buf is None or self._buf.numel() != 3 * n:
            self._buf = dev.empty(max(1, 3 * n), "float64")
        if n:
            _lib.check(_lib.load().ysp_affine_positions(
                dev.ptr(self.body_matrix.value), dev.ptr(self.body_translation.value),
                dev.ptr(self.conn.value), dev.ptr(self.r.value), n, dev.ptr(self._buf),
                dev.stream()), "ysp_affine_positions")
        return self._buf[:3 * n] if n else self._buf[:0]


def is_affine(att) -> bool:
    return isinstance(att, attribute) and att.operator == COMPUTED and \
        isinstance(att.expression, AffineMap)
