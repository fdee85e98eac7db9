"""Energy helpers with the signatures of the reference example helpers.

``stable_neo_hookean`` and ``inertia`` take the same arguments, in the same order, as
``examples/one_bunny/helpers.py:42-71``; instead of a symbolic expression they return an
energy description that ``primitive.addAttribute(name, computed_attribute=...)`` binds to
a primitive and ``scene.addEnergy`` evaluates with the matching HIP kernel.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import _lib, device as dev
from .core import attribute, COMPUTED, DATA, CONSTANT, JOIN, UNION, primitiveUnion
from .program import PLeaf, PJoin, PUnion, PConcat


def _leaf(att: attribute, wrt_pos: dict) -> Optional[PLeaf]:
    """Leaf of a DATA attribute (CONSTANT attributes have no path and no slot)."""
    if att.operator == CONSTANT:
        return None
    if att.operator != DATA:
        raise NotImplementedError(f"{att.fullName}: only DATA leaves are supported below a JOIN/UNION")
    return PLeaf(att.size, wrt_pos.get(att), att.isPrimitiveLevel)


def source_node(src: attribute, wrt_pos: dict):
    """Program node of the attribute a JOIN gathers from."""
    if src.operator == UNION:
        members = []
        for child in src.children:
            leaf = _leaf(child, wrt_pos)
            if leaf is None:
                raise NotImplementedError(
                    f"{child.fullName}: a CONSTANT union child would leave zero rows in the "
                    "local Hessian; declare it with addAttribute (DATA, not a target) instead")
            members.append([leaf])
        return PUnion(src.correspondance, members)
    leaf = _leaf(src, wrt_pos)
    if leaf is None:
        raise NotImplementedError(f"{src.fullName}: the variable of an energy cannot be a constant")
    return leaf


def union_view(src: attribute) -> _lib.UnionF64:
    """ysp_union_f64 of a position source (a union attribute or a plain attribute)."""
    u = _lib.UnionF64()
    if src.operator == UNION:
        kids = src.children
        if len(kids) > _lib.MAX_UNION_CHILDREN:
            raise NotImplementedError("union with more than 8 children")
        u.num_children = len(kids)
        u.counts = dev.ptr(src.correspondance.counts_device)
        for i, k in enumerate(kids):
            u.data[i] = dev.ptr(k.value)
        u._keep = [src.correspondance.counts_device] + [k.value for k in kids]
    else:
        u.num_children = 1
        counts = dev.to_device(np.array([src.numInstances], np.uint32), "uint32")
        u.counts = dev.ptr(counts)
        u.data[0] = dev.ptr(src.value)
        u._keep = [counts, src.value]
    u.width = src.size
    return u


def _scalar_or_array(att, n_instances: int):
    """(device pointer tensor, stride) of a per-element parameter: a one-instance
    constant is broadcast (stride 0), a per-instance attribute is read at stride 1."""
    if isinstance(att, (int, float)):
        return dev.to_device(np.array([float(att)]), "float64"), 0
    if att.size != 1:
        raise ValueError(f"{att.fullName}: expected a scalar parameter")
    if att.numInstances == 1:
        return att.value, 0
    if att.numInstances != n_instances:
        raise ValueError(f"{att.fullName}: has {att.numInstances} instances, expected 1 or {n_instances}")
    return att.value, 1


def _host_scalar(att) -> float:
    if isinstance(att, (int, float)):
        return float(att)
    v = att.host_value()
    if v.size != 1:
        raise ValueError(f"{att.fullName}: expected one scalar value")
    return float(v[0])


class EnergyExpr:
    """Base of the hand-written energies. ``bind`` turns it into a COMPUTED attribute."""
    kind = "energy"

    def bind(self, name: str, owner) -> attribute:
        self.check_owner(owner)
        att = attribute(name, owner, 1, 1, COMPUTED, expression=self)
        self.energy_attribute = att
        return att

    def check_owner(self, owner):
        pass

    @property
    def num_instances(self) -> int:
        return self.energy_attribute.numInstances


class StableNeoHookean(EnergyExpr):
    """``examples/one_bunny/helpers.py:42-67`` on a tet primitive."""
    kind = "stable_neo_hookean"

    def __init__(self, tet_position_rest, tet_position, mu, lam, dt):
        for a, nm in ((tet_position_rest, "tet_position_rest"), (tet_position, "tet_position")):
            if not isinstance(a, attribute) or a.operator != JOIN or a.through.dimension != 4 \
                    or a.source.size != 3:
                raise TypeError(f"stable_neo_hookean: {nm} must be a JOIN of a 3-vector through a "
                                "4-vertex connectivity")
        if tet_position_rest.through is not tet_position.through:
            raise ValueError("stable_neo_hookean: rest and current positions must share the connectivity")
        self.rest = tet_position_rest
        self.pos = tet_position
        self.mu, self.lam, self.dt = mu, lam, dt
        self._rest_key = None
        self._dm_inv = None
        self._vol = None

    def check_owner(self, owner):
        if owner is not self.pos.correspondance:
            raise ValueError("stable_neo_hookean: bind the energy to the tet primitive")

    @property
    def conn(self):
        return self.pos.through

    def program(self, wrt_pos: dict):
        return PConcat([PJoin(self.conn, 4, [source_node(self.pos.source, wrt_pos)])])

    def rest_data(self):
        src = self.rest.source
        key = (self.conn.version, tuple(c.version for c in (src.children or [src])),
               self.num_instances)
        if key != self._rest_key:
            n = self.num_instances
            self._dm_inv = dev.empty(9 * n, "float64")
            self._vol = dev.empty(n, "float64")
            u = union_view(src)
            _lib.check(_lib.load().ysp_snh_rest_data(C.byref(u), dev.ptr(self.conn.value), n,
                                                     dev.ptr(self._dm_inv), dev.ptr(self._vol),
                                                     dev.stream()), "ysp_snh_rest_data")
            self._rest_key = key
        return self._dm_inv, self._vol

    def _params(self):
        n = self.num_instances
        mu, mus = _scalar_or_array(self.mu, n)
        lam, lams = _scalar_or_array(self.lam, n)
        return mu, mus, lam, lams, _host_scalar(self.dt)

    def evaluate(self, scatter: _lib.ScatterArgs, projection_method: int):
        n = self.num_instances
        if n == 0:
            return
        dm_inv, vol = self.rest_data()
        mu, mus, lam, lams, dt = self._params()
        u = union_view(self.pos.source)
        method = max(0, projection_method)
        _lib.check(_lib.load().ysp_eval_snh_tet(
            C.byref(u), dev.ptr(self.conn.value), n, dev.ptr(dm_inv), dev.ptr(vol),
            dev.ptr(mu), mus, dev.ptr(lam), lams, dt, method, C.byref(scatter), dev.stream()),
            "ysp_eval_snh_tet")

    def energy_values(self, att):
        n = self.num_instances
        out = dev.empty(n, "float64")
        if n == 0:
            return out
        dm_inv, vol = self.rest_data()
        mu, mus, lam, lams, dt = self._params()
        u = union_view(self.pos.source)
        _lib.check(_lib.load().ysp_energy_snh_tet(
            C.byref(u), dev.ptr(self.conn.value), n, dev.ptr(dm_inv), dev.ptr(vol), dev.ptr(mu),
            mus, dev.ptr(lam), lams, dt, dev.ptr(out), dev.stream()), "ysp_energy_snh_tet")
        return out


class Inertia(EnergyExpr):
    """``examples/one_bunny/helpers.py:69-71`` on a vertex primitive."""
    kind = "inertia"

    def __init__(self, x_before, vel, dt, x, mass):
        for a, nm in ((x_before, "x_before"), (vel, "vel"), (x, "x")):
            if not isinstance(a, attribute) or a.size != 3:
                raise TypeError(f"inertia: {nm} must be a 3-vector attribute")
        if not (x_before.correspondance is vel.correspondance is x.correspondance):
            raise ValueError("inertia: x_before, vel and x must live on one primitive")
        if x.operator != DATA:
            raise TypeError("inertia: x must be a DATA attribute")
        self.xp, self.v, self.x, self.mass, self.dt = x_before, vel, x, mass, dt

    def check_owner(self, owner):
        if owner is not self.x.correspondance:
            raise ValueError("inertia: bind the energy to the primitive of x")

    def program(self, wrt_pos: dict):
        return PConcat([PLeaf(3, wrt_pos.get(self.x), True)])

    def _mass(self):
        n = self.num_instances
        if isinstance(self.mass, (int, float)):
            return dev.to_device(np.full(n, float(self.mass)), "float64")
        if self.mass.numInstances == 1 and n != 1:
            return self.mass.value.expand(n).contiguous()
        return self.mass.value

    def evaluate(self, scatter: _lib.ScatterArgs, projection_method: int):
        n = self.num_instances
        if n == 0:
            return
        m = self._mass()
        _lib.check(_lib.load().ysp_eval_inertia(
            dev.ptr(self.x.value), dev.ptr(self.xp.value), dev.ptr(self.v.value), dev.ptr(m),
            _host_scalar(self.dt), n, max(0, projection_method), C.byref(scatter), dev.stream()),
            "ysp_eval_inertia")

    def energy_values(self, att):
        n = self.num_instances
        out = dev.empty(n, "float64")
        if n == 0:
            return out
        m = self._mass()
        _lib.check(_lib.load().ysp_energy_inertia(
            dev.ptr(self.x.value), dev.ptr(self.xp.value), dev.ptr(self.v.value), dev.ptr(m),
            _host_scalar(self.dt), n, dev.ptr(out), dev.stream()), "ysp_energy_inertia")
        return out


def stable_neo_hookean(tet_position_rest, tet_position, mu, lam, dt) -> StableNeoHookean:
    return StableNeoHookean(tet_position_rest, tet_position, mu, lam, dt)


def inertia(x_before, vel, dt, x, mass) -> Inertia:
    return Inertia(x_before, vel, dt, x, mass)
