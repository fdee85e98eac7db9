"""Elementwise expressions of attributes, the computed attributes the reference examples
build with attribute arithmetic before handing them to an energy
(``examples/bunny_material_optimization/bunny_material_optimization.py:340,355-359``):

    mu = young / (2.0 * (1.0 + poisson))
    lam = young * poisson / ((1.0 + poisson) * (1.0 - 2.0 * poisson))
    velocity = (last_position - last_last_position) / dt

An expression combines attributes of one primitive (one value block per instance) and
one-instance attributes (broadcast) with +, -, *, / and numbers. Instances are
independent, so an expression has a value per instance and, for a leaf attribute theta,
a per-instance derivative d value / d theta (rows x theta.size). Energies take
expressions where the reference helpers take computed attributes (SNH mu and lam, the
inertia's previous position, velocity and mass); the mixed second derivatives of
``yasps_amd.differentiator`` use the derivatives to carry d g / d (mu, lam, ...) over to
the attribute the caller differentiates by (the chain rule the reference's symbolic
differentiator applies, ``differentiator/differentiator.pyx:367-489``).

Values are evaluated elementwise on the device with the tensors of the leaves; a leaf's
tensor comes from ``value_of`` (the attribute's device value by default)."""
from __future__ import annotations

from typing import Callable, Optional

import numpy as np

from . import device as dev


def _torch():
    return dev.torch()


def _default_value(att):
    return att.value


class Expr:
    """Base node. ``rows``: values per instance (1 or 3)."""
    rows: int = 1

    # ------------------------------------------------------------ structure
    def leaves(self) -> list:
        raise NotImplementedError

    def depends_on(self, theta) -> bool:
        return any(l is theta for l in self.leaves())

    def instances(self) -> int:
        """Number of instances (1 when every leaf is a one-instance attribute); leaves of
        several instances must agree."""
        n = 1
        for l in self.leaves():
            m = l.numInstances
            if m != 1:
                if n != 1 and m != n:
                    raise ValueError(f"expression mixes {n} and {m} instances ({l.fullName})")
                n = m
        return n

    def owner(self):
        """The primitive of the leaves with more than one instance (None if all are
        one-instance)."""
        for l in self.leaves():
            if l.numInstances != 1:
                return l.correspondance
        return None

    # ------------------------------------------------------------ evaluation
    def evaluate(self, n: int, value_of: Callable = _default_value, device=None):
        """[n, rows] float64 values (on ``device``, the framework's device by default)."""
        raise NotImplementedError

    def partial(self, theta, n: int, value_of: Callable = _default_value, device=None):
        """[n, rows, theta.size] derivative w.r.t. the leaf ``theta`` (instance by
        instance: the value of instance i depends on instance i of theta, or on its only
        instance when theta has one), or None when the expression does not depend on it."""
        raise NotImplementedError

    # ------------------------------------------------------------ algebra
    def __add__(self, o):
        return Binary("+", self, as_expr(o))

    def __radd__(self, o):
        return Binary("+", as_expr(o), self)

    def __sub__(self, o):
        return Binary("-", self, as_expr(o))

    def __rsub__(self, o):
        return Binary("-", as_expr(o), self)

    def __mul__(self, o):
        return Binary("*", self, as_expr(o))

    def __rmul__(self, o):
        return Binary("*", as_expr(o), self)

    def __truediv__(self, o):
        return Binary("/", self, as_expr(o))

    def __rtruediv__(self, o):
        return Binary("/", as_expr(o), self)

    def __neg__(self):
        return Binary("*", Const(-1.0), self)


class Const(Expr):
    def __init__(self, value):
        a = np.asarray(value, dtype=np.float64).reshape(-1)
        if a.size not in (1, 3):
            raise ValueError("expression constants are scalars or 3-vectors")
        self.value = a
        self.rows = a.size

    def leaves(self):
        return []

    def evaluate(self, n, value_of=_default_value, device=None):
        t = _torch()
        return t.as_tensor(self.value, dtype=t.float64,
                           device=device or dev.device()).reshape(1, -1).expand(n, -1)

    def partial(self, theta, n, value_of=_default_value, device=None):
        return None


class Leaf(Expr):
    def __init__(self, att):
        if att.size not in (1, 3):
            raise ValueError(f"{att.fullName}: expression leaves are scalars or 3-vectors")
        self.att = att
        self.rows = att.size

    def leaves(self):
        return [self.att]

    def evaluate(self, n, value_of=_default_value, device=None):
        v = value_of(self.att).reshape(-1, self.rows)
        if v.shape[0] == n:
            return v
        if v.shape[0] == 1:
            return v.expand(n, -1)
        raise ValueError(f"{self.att.fullName}: {v.shape[0]} instances, expected 1 or {n}")

    def partial(self, theta, n, value_of=_default_value, device=None):
        if theta is not self.att:
            return None
        t = _torch()
        eye = t.eye(self.rows, dtype=t.float64, device=device or dev.device())
        return eye.reshape(1, self.rows, self.rows).expand(n, -1, -1)


def _bcast(a, rows):
    """[n, r] -> [n, rows] (r == rows or r == 1)."""
    return a if a.shape[1] == rows else a.expand(-1, rows)


def _bcast_d(d, rows):
    """[n, r, w] -> [n, rows, w]."""
    return d if d.shape[1] == rows else d.expand(-1, rows, -1)


class Binary(Expr):
    def __init__(self, op: str, a: Expr, b: Expr):
        if op in "+-" and a.rows != b.rows:
            raise ValueError(f"expression {op}: operands of {a.rows} and {b.rows} rows")
        if op == "*" and a.rows != b.rows and 1 not in (a.rows, b.rows):
            raise ValueError("expression *: elementwise products need equal rows or a scalar")
        if op == "/" and b.rows != 1:
            raise ValueError("expression /: the divisor must be a scalar")
        self.op, self.a, self.b = op, a, b
        self.rows = max(a.rows, b.rows)

    def leaves(self):
        seen, out = set(), []
        for l in self.a.leaves() + self.b.leaves():
            if id(l) not in seen:
                seen.add(id(l))
                out.append(l)
        return out

    def evaluate(self, n, value_of=_default_value, device=None):
        a = _bcast(self.a.evaluate(n, value_of, device), self.rows)
        b = _bcast(self.b.evaluate(n, value_of, device), self.rows)
        if self.op == "+":
            return a + b
        if self.op == "-":
            return a - b
        if self.op == "*":
            return a * b
        return a / b

    def partial(self, theta, n, value_of=_default_value, device=None):
        da = self.a.partial(theta, n, value_of, device)
        db = self.b.partial(theta, n, value_of, device)
        if da is None and db is None:
            return None
        r, w = self.rows, theta.size
        t = _torch()
        zero = None

        def z():
            nonlocal zero
            if zero is None:
                zero = t.zeros((n, r, w), dtype=t.float64, device=device or dev.device())
            return zero

        da = z() if da is None else _bcast_d(da, r)
        db = z() if db is None else _bcast_d(db, r)
        if self.op == "+":
            return da + db
        if self.op == "-":
            return da - db
        a = _bcast(self.a.evaluate(n, value_of, device), r).unsqueeze(-1)
        b = _bcast(self.b.evaluate(n, value_of, device), r).unsqueeze(-1)
        if self.op == "*":
            return da * b + a * db
        return da / b - a * db / (b * b)


def as_expr(o) -> Expr:
    from .core import attribute
    if isinstance(o, Expr):
        return o
    if isinstance(o, attribute):
        return Leaf(o)
    if isinstance(o, (int, float, np.floating, np.integer)):
        return Const(float(o))
    if isinstance(o, (list, tuple, np.ndarray)):
        return Const(o)
    raise TypeError(f"expression operand of type {type(o).__name__}")


def is_expr(o) -> bool:
    return isinstance(o, Expr)


def values(o, n: int, rows: int, value_of: Callable = _default_value, device=None):
    """Values of an energy argument (number, attribute or expression) as a contiguous
    [n, rows] tensor, or [1, rows] when it has one instance (broadcast by the caller)."""
    e = as_expr(o)
    if e.rows != rows:
        raise ValueError(f"expected {rows} values per instance, got {e.rows}")
    m = e.instances()
    if m not in (1, n):
        raise ValueError(f"argument has {m} instances, expected 1 or {n}")
    return e.evaluate(m, value_of, device).contiguous()


def derivative(o, theta, n: int, rows: int, value_of: Callable = _default_value,
               device=None) -> Optional[object]:
    """[n, rows, theta.size] derivative of an energy argument w.r.t. the attribute theta
    (None when the argument does not depend on it)."""
    if isinstance(o, (int, float)):
        return None
    e = as_expr(o)
    if not e.depends_on(theta):
        return None
    d = e.partial(theta, n, value_of, device)
    return None if d is None else _bcast_d(d, rows).contiguous()
