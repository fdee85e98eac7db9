"""Energy helpers with the signatures of the reference example helpers.

``stable_neo_hookean`` and ``inertia`` take the same arguments, in the same order, as
``examples/one_bunny/helpers.py:42-71``; instead of a symbolic expression they return an
energy description that ``primitive.addAttribute(name, computed_attribute=...)`` binds to
a primitive and ``scene.addEnergy`` evaluates with the matching HIP kernel.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import _lib, device as dev
from .core import attribute, COMPUTED, DATA, CONSTANT, JOIN, UNION, primitiveUnion
from .program import PLeaf, PJoin, PUnion, PConcat
from .affine import is_affine
from .expr import is_expr, values as expr_values


def _leaf(att: attribute, wrt_pos: dict) -> Optional[PLeaf]:
    """Leaf of a DATA attribute (CONSTANT attributes have no path and no slot)."""
    if att.operator == CONSTANT:
        return None
    if att.operator != DATA:
        raise NotImplementedError(f"{att.fullName}: only DATA leaves are supported below a JOIN/UNION")
    return PLeaf(att.size, wrt_pos.get(att), att.isPrimitiveLevel)


def _affine_node(att: attribute, wrt_pos: dict) -> PJoin:
    """Slots of an affine-body vertex: the body's A (9) then t (3) through the vertex ->
    body connectivity."""
    m = att.expression
    return PJoin(m.conn, 1, [PLeaf(9, wrt_pos.get(m.body_matrix), True),
                             PLeaf(3, wrt_pos.get(m.body_translation), True)])


def source_node(src: attribute, wrt_pos: dict):
    """Program node of the attribute a JOIN gathers from."""
    if is_affine(src):
        return _affine_node(src, wrt_pos)
    if src.operator == UNION:
        members = []
        for child in src.children:
            if is_affine(child):
                members.append([_affine_node(child, wrt_pos)])
                continue
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
        vals = [k.expression.values() if is_affine(k) else k.value for k in kids]
        for i, v in enumerate(vals):
            u.data[i] = dev.ptr(v)
        u._keep = [src.correspondance.counts_device] + vals
    elif is_affine(src):
        u.num_children = 1
        counts = dev.constant([src.numInstances], "uint32")
        v = src.expression.values()
        u.counts = dev.ptr(counts)
        u.data[0] = dev.ptr(v)
        u._keep = [counts, v]
    else:
        u.num_children = 1
        counts = dev.constant([src.numInstances], "uint32")
        u.counts = dev.ptr(counts)
        u.data[0] = dev.ptr(src.value)
        u._keep = [counts, src.value]
    u.width = src.size
    return u


def affine_children(src: attribute):
    """ysp_affine_children of a position source, or None when no child is affine."""
    kids = src.children if src.operator == UNION else [src]
    if not any(is_affine(k) for k in kids):
        return None
    a = _lib.AffineChildren()
    keep = []
    for i, k in enumerate(kids):
        if is_affine(k):
            m = k.expression
            a.mask |= 1 << i
            a.vertex_body[i] = dev.ptr(m.conn.value)
            a.rest_position[i] = dev.ptr(m.r.value)
            keep += [m.conn.value, m.r.value]
    a._keep = keep
    return a


def _scalar_or_array(att, n_instances: int):
    """(device pointer tensor, stride) of a per-element parameter: a one-instance
    constant is broadcast (stride 0), a per-instance attribute is read at stride 1; an
    expression (yasps_amd.expr) is evaluated per instance."""
    if isinstance(att, (int, float)):
        return dev.constant([float(att)], "float64"), 0
    if is_expr(att):
        v = expr_values(att, n_instances, 1)
        return v.reshape(-1), (1 if v.shape[0] == n_instances and n_instances != 1 else 0)
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
    if is_expr(att):
        if att.instances() != 1:
            raise ValueError("expected a scalar parameter (an expression of one-instance "
                             "attributes)")
        return float(dev.to_host(expr_values(att, 1, 1)).reshape(-1)[0])
    v = att.host_value()
    if v.size != 1:
        raise ValueError(f"{att.fullName}: expected one scalar value")
    return float(v[0])


@dataclass
class MixedInfo:
    """Inputs of the mixed-derivative kernel (ysp_element_param_jacobian) for one energy:
    the generated kind, its vertex data and element map, element parameters ep [n, P]
    (rest-shape data, or the energy's own arguments for the kinds whose d g / d p is
    generated), scene parameters se (se_stride doubles per element, 0 = shared), the
    argument each parameter slot comes from (p_args: (argument, rows) in slot order, for
    the kinds with d g / d p; s_args: one argument per scene slot, None for a slot the
    kernel fills itself), and the attributes that are DoF sources or rest-shape inputs
    (a column target among them has no mixed Jacobian here)."""
    kind: int
    positions: object
    elem2v: object
    n: int
    dofs: int
    ep: object
    se: object
    se_stride: int
    p_args: list
    s_args: list
    dof_leaves: list
    rest_leaves: list = field(default_factory=list)
    param_jacobian: object = None    # another entry point with the ABI of
                                     # ysp_element_param_jacobian minus the kind (user energies)


def attribute_leaves(src) -> list:
    """An attribute and, through unions and affine maps, the attributes it is made of."""
    if src is None:
        return []
    out = [src]
    for k in getattr(src, "children", []) or []:
        out += attribute_leaves(k)
    if is_affine(src):
        m = src.expression
        out += [m.t.source, m.A.source, m.r]
    return out


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

    def _generic(self):
        """Element/scene parameters of the generated SNH (vertices driven by affine
        bodies take the generic chain-rule kernel)."""
        n = self.num_instances
        dm_inv, vol = self.rest_data()
        t = dev.torch()
        ep = t.cat([dm_inv.reshape(n, 9), vol.reshape(n, 1)], dim=1).contiguous()
        sp = dev.constant([_host_scalar(self.mu), _host_scalar(self.lam), _host_scalar(self.dt)],
                          "float64")
        return ep, sp

    def evaluate(self, scatter: _lib.ScatterArgs, projection_method: int):
        n = self.num_instances
        if n == 0:
            return
        method = max(0, projection_method)
        aff = affine_children(self.pos.source)
        if aff is not None:
            ep, sp = self._generic()
            u = union_view(self.pos.source)
            _lib.check(_lib.load().ysp_eval_element_energy(
                KIND_IDS["snh"], C.byref(u), C.byref(aff), dev.ptr(self.conn.value), n, dev.ptr(ep),
                dev.ptr(sp), method, C.byref(scatter), dev.stream()), "ysp_eval_element_energy")
            return
        dm_inv, vol = self.rest_data()
        mu, mus, lam, lams, dt = self._params()
        u = union_view(self.pos.source)
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


    def mixed_info(self) -> MixedInfo:
        """Generated SNH: element parameters [Dm^-1 (9), vol] (rest shape), scene
        parameters [mu, lam, dt] per element (Lamé coefficients may be per-tet attributes
        or expressions)."""
        if affine_children(self.pos.source) is not None:
            raise NotImplementedError("mixed Jacobians of SNH on affine-body vertices are not "
                                      "supported")
        n = self.num_instances
        t = dev.torch()
        dm_inv, vol = self.rest_data()
        ep = t.cat([dm_inv.reshape(n, 9), vol.reshape(n, 1)], dim=1).contiguous()
        cols = [expr_values(a, n, 1).expand(n, 1) for a in (self.mu, self.lam, self.dt)]
        se = t.cat(cols, dim=1).contiguous()
        return MixedInfo(KIND_IDS["snh"], union_view(self.pos.source), self.conn.value, n, 12,
                         ep, se, 3, [], [self.mu, self.lam, self.dt],
                         attribute_leaves(self.pos.source), attribute_leaves(self.rest.source))


class Inertia(EnergyExpr):
    """``examples/one_bunny/helpers.py:69-71`` on a vertex primitive."""
    kind = "inertia"

    def __init__(self, x_before, vel, dt, x, mass):
        if not isinstance(x, attribute) or x.size != 3:
            raise TypeError("inertia: x must be a 3-vector attribute")
        for a, nm in ((x_before, "x_before"), (vel, "vel")):
            if is_expr(a):
                if a.rows != 3:
                    raise TypeError(f"inertia: {nm} must be a 3-vector expression")
                if a.owner() not in (None, x.correspondance):
                    raise ValueError(f"inertia: the attributes of {nm} must live on x's primitive")
            elif not isinstance(a, attribute) or a.size != 3:
                raise TypeError(f"inertia: {nm} must be a 3-vector attribute or expression")
            elif a.correspondance is not x.correspondance:
                raise ValueError("inertia: x_before, vel and x must live on one primitive")
        if x.operator != DATA and not is_affine(x):
            raise TypeError("inertia: x must be a DATA attribute or an affine-body vertex map")
        self.xp, self.v, self.x, self.mass, self.dt = x_before, vel, x, mass, dt
        self._iota = None

    def _vec(self, a):
        """[n, 3] contiguous values of x_before / vel (attribute or expression)."""
        n = self.num_instances
        return expr_values(a, n, 3).expand(n, 3).contiguous()

    def check_owner(self, owner):
        if owner is not self.x.correspondance:
            raise ValueError("inertia: bind the energy to the primitive of x")

    def program(self, wrt_pos: dict):
        if is_affine(self.x):
            return PConcat([_affine_node(self.x, wrt_pos)])
        return PConcat([PLeaf(3, wrt_pos.get(self.x), True)])

    def _generic_args(self):
        n = self.num_instances
        t = dev.torch()
        ep = t.cat([self._vec(self.xp), self._vec(self.v), self._mass().reshape(n, 1)],
                   dim=1).contiguous()
        sp = dev.constant([_host_scalar(self.dt)], "float64")
        if self._iota is None or self._iota.numel() != n:
            self._iota = dev.to_device(np.arange(n, dtype=np.uint32), "uint32")
        return ep, sp, union_view(self.x), affine_children(self.x)

    def _mass(self):
        n = self.num_instances
        if isinstance(self.mass, (int, float)):
            return dev.constant_full(n, float(self.mass))
        if is_expr(self.mass):
            return expr_values(self.mass, n, 1).expand(n, 1).reshape(n).contiguous()
        if self.mass.numInstances == 1 and n != 1:
            return self.mass.value.expand(n).contiguous()
        return self.mass.value

    def evaluate(self, scatter: _lib.ScatterArgs, projection_method: int):
        n = self.num_instances
        if n == 0:
            return
        if is_affine(self.x):
            ep, sp, u, aff = self._generic_args()
            _lib.check(_lib.load().ysp_eval_element_energy(
                KIND_IDS["inertia"], C.byref(u), C.byref(aff), dev.ptr(self._iota), n, dev.ptr(ep),
                dev.ptr(sp), max(0, projection_method), C.byref(scatter), dev.stream()),
                "ysp_eval_element_energy")
            return
        m = self._mass()
        xp, v = self._vec(self.xp), self._vec(self.v)
        _lib.check(_lib.load().ysp_eval_inertia(
            dev.ptr(self.x.value), dev.ptr(xp), dev.ptr(v), dev.ptr(m),
            _host_scalar(self.dt), n, max(0, projection_method), C.byref(scatter), dev.stream()),
            "ysp_eval_inertia")

    def energy_values(self, att):
        n = self.num_instances
        out = dev.empty(n, "float64")
        if n == 0:
            return out
        if is_affine(self.x):
            ep, sp, u, _ = self._generic_args()
            _lib.check(_lib.load().ysp_element_energy_values(
                KIND_IDS["inertia"], C.byref(u), dev.ptr(self._iota), n, dev.ptr(ep), dev.ptr(sp),
                dev.ptr(out), dev.stream()), "ysp_element_energy_values")
            return out
        m = self._mass()
        xp, v = self._vec(self.xp), self._vec(self.v)
        _lib.check(_lib.load().ysp_energy_inertia(
            dev.ptr(self.x.value), dev.ptr(xp), dev.ptr(v), dev.ptr(m),
            _host_scalar(self.dt), n, dev.ptr(out), dev.stream()), "ysp_energy_inertia")
        return out


    def mixed_info(self) -> MixedInfo:
        """Generated inertia: element parameters [x_prev (3), v (3), m] are the energy's own
        arguments (d g / d p generated), scene parameter [dt]."""
        if is_affine(self.x):
            raise NotImplementedError("mixed Jacobians of the inertia of affine-body vertices "
                                      "are not supported")
        ep, sp, u, _ = self._generic_args()
        return MixedInfo(KIND_IDS["inertia"], u, self._iota, self.num_instances, 3, ep, sp, 0,
                         [(self.xp, 3), (self.v, 3), (self.mass, 1)], [self.dt], [self.x])


def stable_neo_hookean(tet_position_rest, tet_position, mu, lam, dt) -> StableNeoHookean:
    return StableNeoHookean(tet_position_rest, tet_position, mu, lam, dt)


def inertia(x_before, vel, dt, x, mass) -> Inertia:
    return Inertia(x_before, vel, dt, x, mass)


# ---------------------------------------------------------------- generated energies
KIND_IDS = {"point_point": 0, "point_edge": 1, "point_triangle": 2, "edge_edge": 3,
            "bending": 4, "baraff_witkin": 5, "stvk_membrane": 6, "affine_orthogonality": 7,
            "snh": 8, "inertia": 9, "quadratic_form": 10}


class ElementEnergy(EnergyExpr):
    """An energy of ``csrc/generated_energies.hpp`` on the elements of one primitive whose
    vertex data is a JOIN (``position``) through a fixed-arity connectivity; bending and
    membrane energies also take the rest JOIN (``rest``) for their element parameters."""

    def __init__(self, kind: str, position, scene_params, rest=None, thickness: float = 1.0):
        self.kind = kind
        self.kind_id = KIND_IDS[kind]
        arity, width, npar, nsc = (C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_uint32())
        _lib.check(_lib.load().ysp_energy_arity(self.kind_id, C.byref(arity), C.byref(width),
                                                C.byref(npar), C.byref(nsc)), "ysp_energy_arity")
        self.arity, self.width, self.npar, self.nsc = arity.value, width.value, npar.value, nsc.value
        for a, nm in ((position, "position"), (rest, "rest")):
            if a is None:
                continue
            if not isinstance(a, attribute) or a.operator != JOIN or \
                    a.through.dimension != self.arity or a.source.size != self.width:
                raise TypeError(f"{kind}: {nm} must be a JOIN of a {self.width}-vector through a "
                                f"{self.arity}-vertex connectivity")
        if (rest is None) != (self.npar == 0):
            raise ValueError(f"{kind}: rest positions are {'required' if self.npar else 'not used'}")
        if rest is not None and rest.through is not position.through:
            raise ValueError(f"{kind}: rest and current positions must share the connectivity")
        if len(scene_params) != self.nsc:
            raise ValueError(f"{kind}: expected {self.nsc} scene parameters")
        self.pos, self.rest = position, rest
        self.scene_params = list(scene_params)
        self.thickness = float(thickness)
        self._rest_key = None
        self._ep = None

    def check_owner(self, owner):
        if owner is not self.pos.correspondance:
            raise ValueError(f"{self.kind}: bind the energy to the primitive of its connectivity")

    @property
    def conn(self):
        return self.pos.through

    def program(self, wrt_pos: dict):
        return PConcat([PJoin(self.conn, self.arity, [source_node(self.pos.source, wrt_pos)])])

    def element_params(self):
        n = self.num_instances
        if self.npar == 0:
            return None
        src = self.rest.source
        key = (self.conn.version, tuple(c.version for c in (src.children or [src])), n)
        if key != self._rest_key:
            self._ep = dev.empty(max(1, self.npar * n), "float64")
            u = union_view(src)
            _lib.check(_lib.load().ysp_element_rest_data(
                self.kind_id, C.byref(u), dev.ptr(self.conn.value), n, self.thickness,
                dev.ptr(self._ep), dev.stream()), "ysp_element_rest_data")
            self._rest_key = key
        return self._ep

    def _scene(self):
        vals = [_host_scalar(v) for v in self.scene_params]
        return dev.constant(vals, "float64")

    def elem2v(self):
        return self.conn.value

    def positions(self):
        return union_view(self.pos.source)

    def affine(self):
        return affine_children(self.pos.source)

    def evaluate(self, scatter: _lib.ScatterArgs, projection_method: int):
        n = self.num_instances
        if n == 0:
            return
        ep = self.element_params()
        sp = self._scene()
        u = self.positions()
        aff = self.affine()
        _lib.check(_lib.load().ysp_eval_element_energy(
            self.kind_id, C.byref(u), C.byref(aff) if aff is not None else None,
            dev.ptr(self.elem2v()), n, dev.ptr(ep), dev.ptr(sp), max(0, projection_method),
            C.byref(scatter), dev.stream()), "ysp_eval_element_energy")

    def energy_values(self, att):
        n = self.num_instances
        out = dev.empty(n, "float64")
        if n == 0:
            return out
        ep = self.element_params()
        sp = self._scene()
        u = self.positions()
        _lib.check(_lib.load().ysp_element_energy_values(
            self.kind_id, C.byref(u), dev.ptr(self.elem2v()), n, dev.ptr(ep), dev.ptr(sp),
            dev.ptr(out), dev.stream()), "ysp_element_energy_values")
        return out


    def mixed_info(self) -> MixedInfo:
        """Generated kinds: element parameters from the rest shape, scene parameters
        shared by every element (Baraff-Witkin's stretch coefficients are chosen per
        element by the kernel)."""
        if self.kind == "quadratic_form":
            raise NotImplementedError("quadratic_form has no mixed Jacobian")
        if self.affine() is not None:
            raise NotImplementedError(f"mixed Jacobians of {self.kind} on affine-body vertices "
                                      "are not supported")
        s_args = list(self.scene_params)
        if self.kind == "baraff_witkin":
            s_args[3] = s_args[4] = None
        return MixedInfo(self.kind_id, self.positions(), self.elem2v(), self.num_instances,
                         self.arity * self.width, self.element_params(), self._scene(), 0, [],
                         s_args, attribute_leaves(self.pos.source),
                         attribute_leaves(self.rest.source if self.rest is not None else None))


class AffineOrthogonality(ElementEnergy):
    """``examples/one_bunny_partial_abd/helpers.py:132-139`` on the 9-vector affine matrix
    of every body (row-major A), times ``stiffness``."""

    def __init__(self, rotation, stiffness=1.0):
        if not isinstance(rotation, attribute) or rotation.size != 9 or rotation.operator != DATA:
            raise TypeError("affine_energy: rotation must be a 9-entry DATA attribute (3x3)")
        self.kind = "affine_orthogonality"
        self.kind_id = KIND_IDS[self.kind]
        self.arity, self.width, self.npar, self.nsc = 1, 9, 0, 1
        self.A = rotation
        self.scene_params = [stiffness]
        self.pos = self.rest = None
        self._iota = None

    def check_owner(self, owner):
        if owner is not self.A.correspondance and owner is not getattr(self.A.correspondance, "parent", None):
            raise ValueError("affine_energy: bind the energy to the body primitive or its mesh")

    @property
    def num_instances(self) -> int:
        return self.A.numInstances

    def program(self, wrt_pos: dict):
        return PConcat([PLeaf(9, wrt_pos.get(self.A), True)])

    def element_params(self):
        return None

    def elem2v(self):
        n = self.num_instances
        if self._iota is None or self._iota.numel() != n:
            self._iota = dev.to_device(np.arange(n, dtype=np.uint32), "uint32")
        return self._iota

    def affine(self):
        return None

    def positions(self):
        u = _lib.UnionF64()
        u.num_children = 1
        counts = dev.constant([self.A.numInstances], "uint32")
        u.counts = dev.ptr(counts)
        u.data[0] = dev.ptr(self.A.value)
        u.width = 9
        u._keep = [counts, self.A.value]
        return u

    def mixed_info(self) -> MixedInfo:
        return MixedInfo(self.kind_id, self.positions(), self.elem2v(), self.num_instances, 9,
                         None, self._scene(), 0, [], list(self.scene_params), [self.A])


def point_point(position, dHat, kappa) -> ElementEnergy:
    """``examples/one_bunny/helpers.py:74-82``."""
    return ElementEnergy("point_point", position, [dHat, kappa])


def point_edge(position, dHat, kappa) -> ElementEnergy:
    """``helpers.py:84-94`` (squared distance of the point to the edge's line)."""
    return ElementEnergy("point_edge", position, [dHat, kappa])


def point_triangle(position, dHat, kappa) -> ElementEnergy:
    """``helpers.py:96-107`` (squared distance of the point to the triangle's plane)."""
    return ElementEnergy("point_triangle", position, [dHat, kappa])


def edge_edge(position, dHat, kappa) -> ElementEnergy:
    """``helpers.py:109-120`` (squared distance between the two edges' lines)."""
    return ElementEnergy("edge_edge", position, [dHat, kappa])


def bending(x, x_init, bendStiff) -> ElementEnergy:
    """``examples/many_bunnies_one_cloth/helpers.py:193-212`` on hinges (4 vertices: the
    shared edge x0-x1 and the two opposite vertices)."""
    return ElementEnergy("bending", x, [bendStiff], rest=x_init)


def baraff_witkin(x_init, x, stretchS, shearS, thickness, dt) -> ElementEnergy:
    """``examples/many_bunnies_one_cloth/helpers.py:239-259``."""
    return ElementEnergy("baraff_witkin", x, [stretchS, shearS, dt, 0.0, 0.0], rest=x_init,
                         thickness=_host_scalar(thickness))


def stvk_membrane(x_init, x, mu, lam, thickness, dt) -> ElementEnergy:
    """StVK membrane area dt^2 (mu |E|^2 + lam/2 tr(E)^2), E = (F^T F - I)/2 with F in the
    rest frame of ``baraff_witkin`` (no reference counterpart)."""
    return ElementEnergy("stvk_membrane", x, [mu, lam, dt], rest=x_init,
                         thickness=_host_scalar(thickness))


def affine_energy(rotation, stiffness=1.0) -> AffineOrthogonality:
    """``examples/one_bunny_partial_abd/helpers.py:132-139``."""
    return AffineOrthogonality(rotation, stiffness)


QUADRATIC_FORM_ROWS = 6


def quadratic_form(position, jacobian, inner_hessian, linear=None, offset=None) -> ElementEnergy:
    """E = y^T Q y / 2 + l^T y on y = J x + c, x the 12 coordinates of a 4-vertex element:
    the energy the reference's sparsity tests build symbolically from a linear map and a
    quadratic (``tests/differentiator/test_hessian_sparsity_permutation_gpu.py:17-38``).
    The projection acts on Q (the map is linear: ``differentiator.pyx:795-801``), so the
    element Hessian is J^T spd(Q) J. J is m x 12 with 1 <= m <= 6. Q must be symmetric to
    rounding (a product such as B diag(d) B^T); its symmetric part (Q + Q^T) / 2, the
    Hessian of y^T Q y / 2, is used."""
    J = np.asarray(jacobian, dtype=np.float64)
    if J.ndim != 2 or J.shape[1] != 12 or not 1 <= J.shape[0] <= QUADRATIC_FORM_ROWS:
        raise ValueError("quadratic_form: jacobian must be m x 12 with 1 <= m <= 6")
    m = J.shape[0]
    Q = np.asarray(inner_hessian, dtype=np.float64)
    if Q.shape != (m, m) or not np.all(np.isfinite(Q)) or \
            np.abs(Q - Q.T).max() > 1e-12 * max(np.abs(Q).max(), 1e-300):
        raise ValueError("quadratic_form: inner_hessian must be a symmetric m x m matrix")
    Q = 0.5 * (Q + Q.T)
    lin = np.zeros(m) if linear is None else np.asarray(linear, dtype=np.float64).reshape(-1)
    off = np.zeros(m) if offset is None else np.asarray(offset, dtype=np.float64).reshape(-1)
    if lin.size != m or off.size != m:
        raise ValueError("quadratic_form: linear and offset must have m entries")
    R = QUADRATIC_FORM_ROWS
    Jp = np.zeros((R, 12))
    Jp[:m] = J
    Qp = np.eye(R)
    Qp[:m, :m] = Q
    lp, cp = np.zeros(R), np.zeros(R)
    lp[:m], cp[:m] = lin, off
    params = np.concatenate([Jp.ravel(), Qp.ravel(), lp, cp])
    return ElementEnergy("quadratic_form", position, [float(v) for v in params])

