"""Scene graph of the drop-in API: scene -> mesh -> primitive / primitiveUnion,
attributes and connectivities.

Mirrors the reference's user-facing classes (``yasps/yasps/scene/scene.pyx``,
``mesh/mesh.pyx``, ``primitive/primitive.pyx``, ``primitive/primitiveUnion.pyx``,
``connectivity/connectivity.pyx``, ``attribute/attribute.pyx``) for the parts the Newton
linear-solve path uses: DATA / CONSTANT attributes with per-instance values, JOIN
attributes (``addAttribute(name, through=conn, source=attr)``), UNION attributes over
stacked child primitives, and computed energy attributes built from the energy helpers
in ``yasps_amd.energies``. There is no symbolic differentiator: energies are hand-written
kernels selected by the helper that built them. Arithmetic on scalar and 3-vector
attributes builds elementwise expressions (``yasps_amd.expr``) that energies take as
parameters (Lamé coefficients from Young's modulus and Poisson's ratio, a velocity from
two previous positions).

Attribute values live on the device (torch tensors, flat float64 [instances * size]);
a value set before any device work is kept on the host and uploaded on first use.
"""
from __future__ import annotations

from typing import Dict, List, Optional
import itertools
import numpy as np

from . import device as dev

DATA = "DATA"
CONSTANT = "CONSTANT"
JOIN = "JOIN"
UNION = "UNION"
COMPUTED = "COMPUTED"

_uid = itertools.count()


class attribute:
    """One attribute of a scene, mesh or primitive (``attribute/attribute.pyx``)."""

    def __init__(self, name: str, owner, rows: int, cols: int, operator: str,
                 through=None, source=None, children=None, expression=None):
        if rows <= 0 or cols <= 0:
            raise ValueError(f"attribute {name}: rows and cols must be positive")
        self.name = name
        self.correspondance = owner
        self.rows = int(rows)
        self.cols = int(cols)
        self.operator = operator
        self.through = through
        self.source = source
        self.children = list(children or [])
        self.expression = expression
        self._host: Optional[np.ndarray] = None
        self._dev = None
        self.version = 0
        self._uid = next(_uid)

    # ------------------------------------------------------------ identity
    @property
    def fullName(self) -> str:
        return f"{self.correspondance.fullName}_{self.name}"

    @property
    def size(self) -> int:
        return self.rows * self.cols

    @property
    def hash(self) -> int:
        return hash((self.fullName, self._uid))

    def __hash__(self):
        return self._uid

    def __eq__(self, other):
        return self is other

    @property
    def numInstances(self) -> int:
        return self.correspondance.numInstances

    @property
    def isDynamic(self) -> bool:
        return bool(getattr(self.correspondance, "isDynamic", False))

    @property
    def isPrimitiveLevel(self) -> bool:
        return isinstance(self.correspondance, (primitive, primitiveUnion))

    # ------------------------------------------------------------ algebra
    def resize(self, rows: int, cols: int) -> "attribute":
        """Reinterpret the per-instance values as rows x cols (same size), as
        ``attribute.resize`` does for a JOINed 9-vector used as a 3 x 3 matrix."""
        if rows * cols != self.size:
            raise ValueError(f"attribute {self.fullName}: cannot resize {self.size} values to {rows}x{cols}")
        self.rows, self.cols = int(rows), int(cols)
        return self

    def __mul__(self, other):
        # a 3 x 3 matrix times a 3-vector attribute: the affine-body vertex map; scalars
        # and 3-vectors otherwise combine elementwise (yasps_amd.expr)
        if self.size == 9 and isinstance(other, attribute) and other.size == 3:
            from .affine import MatrixVector
            return MatrixVector(self, other)
        from .expr import as_expr
        return as_expr(self) * other

    def __add__(self, other):
        from .affine import AffineMap, MatrixVector
        if isinstance(other, MatrixVector):
            return AffineMap(self, other.matrix, other.vector)
        from .expr import as_expr
        return as_expr(self) + other

    def __radd__(self, other):
        from .expr import as_expr
        return as_expr(other) + self

    def __sub__(self, other):
        from .expr import as_expr
        return as_expr(self) - other

    def __rsub__(self, other):
        from .expr import as_expr
        return as_expr(other) - self

    def __rmul__(self, other):
        from .expr import as_expr
        return as_expr(other) * self

    def __truediv__(self, other):
        from .expr import as_expr
        return as_expr(self) / other

    def __rtruediv__(self, other):
        from .expr import as_expr
        return as_expr(other) / self

    def __neg__(self):
        from .expr import as_expr
        return -as_expr(self)

    # ------------------------------------------------------------ values
    def updateValue(self, values, deepCopy: bool = False):
        """Set the per-instance values (``attribute.pyx:417-436``). Accepts numpy
        arrays / lists (host) or device tensors (kept by reference unless deepCopy)."""
        if self.operator not in (DATA, CONSTANT):
            raise ValueError(f"attribute {self.fullName}: only DATA/CONSTANT values can be updated")
        t = dev.torch()
        expected = self.size * self.numInstances
        if isinstance(values, t.Tensor):
            v = values.reshape(-1)
            if v.dtype != t.float64:
                v = v.to(t.float64)
            if v.numel() != expected:
                raise ValueError(f"attribute {self.fullName}: expected {expected} values, got {v.numel()}")
            if not v.is_cuda:
                v = v.to(dev.device())
            elif deepCopy:
                v = v.clone()
            self._dev = v
            self._host = None
        else:
            a = np.asarray(values, dtype=np.float64).reshape(-1)
            if a.size != expected:
                raise ValueError(f"attribute {self.fullName}: expected {expected} values, got {a.size}")
            self._host = a.copy()
            self._dev = None
        self.version += 1
        return self

    @property
    def value(self):
        """Device tensor of the values (DATA/CONSTANT) or of ``compute()``."""
        if self.operator in (DATA, CONSTANT):
            if self._dev is None:
                if self._host is None:
                    self._host = np.zeros(self.size * self.numInstances)
                self._dev = dev.to_device(self._host, "float64")
            return self._dev
        if self._dev is None:
            self.compute()
        return self._dev

    def host_value(self) -> np.ndarray:
        """Host copy of the values. DATA/CONSTANT values set from the host keep that copy
        (the device copy is made from it and replaced, not modified, by updateValue), so
        reading a scalar parameter does not synchronise the stream."""
        if self.operator in (DATA, CONSTANT):
            if self._host is not None:
                return self._host.copy()
            if self._dev is None:
                return np.zeros(self.size * self.numInstances)
        return dev.to_host(self.value)

    def compute(self):
        """Materialise a UNION, JOIN or energy attribute on the device."""
        t = dev.torch()
        if self.operator in (DATA, CONSTANT):
            return self
        if self.operator == UNION:
            self._dev = t.cat([c.value.reshape(-1) for c in self.children]) if self.children else \
                dev.zeros(0, "float64")
        elif self.operator == JOIN:
            src = self.source.value.reshape(-1, self.source.size)
            idx = self.through.value.long().reshape(-1)
            self._dev = src[idx].reshape(-1)
        elif self.operator == COMPUTED:
            values = getattr(self.expression, "values", None)
            self._dev = values(self) if values is not None else self.expression.energy_values(self)
        return self

    def __repr__(self):
        return f"attribute({self.fullName}, {self.operator}, {self.rows}x{self.cols})"


class connectivity:
    """Fixed-arity uint32 connectivity from a primitive to a primitive or union
    (``connectivity/connectivity.pyx:13-118``)."""

    def __init__(self, name: str, owner, target, data, dimension: int):
        if dimension <= 0:
            raise ValueError("connectivity: CSR connectivities (dimension 0) are not supported")
        self.name = name
        self.correspondance = owner
        self.target = target
        self.dimension = int(dimension)
        self._host = np.zeros((0, self.dimension), dtype=np.uint32)
        self._dev = None
        self.version = 0
        self.updateConnectivity(data)

    @property
    def fullName(self):
        return f"{self.correspondance.fullName}_{self.name}"

    def updateConnectivity(self, data):
        t = dev.torch()
        if isinstance(data, t.Tensor):
            v = data.reshape(-1)
            if v.dtype not in (t.int32, t.uint32):
                v = v.to(t.int32)
            if v.dtype == t.uint32:
                v = v.view(t.int32)
            self._dev = v if v.is_cuda else v.to(dev.device())
            self._host = None
        else:
            a = np.asarray(data, dtype=np.int64).reshape(-1)
            if a.size % self.dimension:
                raise ValueError(f"connectivity {self.fullName}: size not a multiple of {self.dimension}")
            if a.size and (a.min() < 0 or a.max() >= max(1, self.target.numInstances)):
                raise ValueError(f"connectivity {self.fullName}: index out of range of {self.target.fullName}")
            self._host = a.astype(np.uint32).reshape(-1, self.dimension)
            self._dev = None
        self.version += 1

    @property
    def value(self):
        if self._dev is None:
            self._dev = dev.to_device(self._host.reshape(-1), "uint32")
        return self._dev

    def host_value(self) -> np.ndarray:
        if self._host is None:
            self._host = dev.to_host(self._dev, "uint32").reshape(-1, self.dimension)
        return self._host

    @property
    def numEntries(self):
        if self._host is not None:
            return self._host.shape[0]
        return self._dev.numel() // self.dimension


class _AttributeOwner:
    def __init__(self, name: str, parent):
        self.name = name
        self.parent = parent
        self.attributes: Dict[str, attribute] = {}

    @property
    def fullName(self) -> str:
        return self.name if self.parent is None else f"{self.parent.fullName}_{self.name}"

    def __getitem__(self, name: str) -> attribute:
        return self.attributes[name]

    def __contains__(self, name):
        return name in self.attributes

    def _register(self, att: attribute) -> attribute:
        if att.name in self.attributes:
            raise ValueError(f"{self.fullName}: attribute {att.name} already exists")
        self.attributes[att.name] = att
        return att

    def addConstant(self, name: str, rows: int = 1, cols: int = 1) -> attribute:
        return self._register(attribute(name, self, rows, cols, CONSTANT))

    def addAttribute(self, name: str, rows: int = 0, cols: int = 0, computed_attribute=None):
        """Scene- or mesh-level attribute: a DATA value or a computed energy (e.g. the
        affine orthogonality of every body, one_bunny_partial_abd.py:229-230)."""
        if computed_attribute is not None:
            return self._register(computed_attribute.bind(name, self))
        return self._register(attribute(name, self, rows, cols, DATA))


class primitive(_AttributeOwner):
    """A primitive type with an instance count (``primitive/primitive.pyx``)."""

    def __init__(self, name: str, mesh_, numInstances: int, isDynamic: bool = False):
        super().__init__(name, mesh_)
        if numInstances < 0:
            raise ValueError("primitive: numInstances must be non-negative")
        self.numInstances = int(numInstances)
        self.isDynamic = bool(isDynamic)
        self.connectivities: Dict[str, connectivity] = {}
        self.unions: List["primitiveUnion"] = []

    def addAttribute(self, name: str, rows: int = 0, cols: int = 0, computed_attribute=None,
                     through: Optional[connectivity] = None, source: Optional[attribute] = None):
        if through is not None or source is not None:
            if through is None or source is None:
                raise ValueError("addAttribute: through and source go together")
            if through.correspondance is not self:
                raise ValueError("addAttribute: the connectivity must start at this primitive")
            if source.correspondance is not through.target:
                raise ValueError("addAttribute: source must live on the connectivity's target")
            return self._register(attribute(name, self, through.dimension * source.rows,
                                            source.cols, JOIN, through=through, source=source))
        if computed_attribute is not None:
            return self._register(computed_attribute.bind(name, self))
        return self._register(attribute(name, self, rows, cols, DATA))

    def addConnectivity(self, name: str, to, data, dimension: int) -> connectivity:
        c = connectivity(name, self, to, data, dimension)
        self.connectivities[name] = c
        return c

    def updateNumInstances(self, n: int):
        if not self.isDynamic:
            raise ValueError(f"{self.fullName}: only dynamic primitives change their instance count")
        self.numInstances = int(n)
        for u in self.unions:
            u._counts_changed()


class primitiveUnion(_AttributeOwner):
    """Stacked child primitives seen as one (``primitive/primitiveUnion.pyx``)."""

    def __init__(self, name: str, mesh_, children: List[primitive]):
        super().__init__(name, mesh_)
        if not children:
            raise ValueError("primitiveUnion: needs at least one child")
        self.children = list(children)
        for c in self.children:
            c.unions.append(self)
        self._counts_dev = None

    @property
    def numInstances(self) -> int:
        return sum(c.numInstances for c in self.children)

    @property
    def isDynamic(self) -> bool:
        return any(c.isDynamic for c in self.children)

    @property
    def children_counts(self) -> List[int]:
        return [c.numInstances for c in self.children]

    def _counts_changed(self):
        self._counts_dev = None

    @property
    def counts_device(self):
        if self._counts_dev is None:
            self._counts_dev = dev.to_device(np.asarray(self.children_counts, np.uint32), "uint32")
        return self._counts_dev

    def addAttribute(self, name: str, rows: int = 0, cols: int = 0, computed_attribute=None,
                     through=None, source=None):
        """Union attribute over the children's attributes named ``name``."""
        if computed_attribute is not None:
            return self._register(computed_attribute.bind(name, self))
        kids = []
        for c in self.children:
            if name not in c.attributes:
                raise ValueError(f"primitiveUnion {self.fullName}: child {c.fullName} has no {name}")
            kids.append(c.attributes[name])
        shapes = {(k.rows, k.cols) for k in kids}
        if len(shapes) != 1:
            raise ValueError(f"primitiveUnion {self.fullName}: children disagree on the shape of {name}")
        r, cc = shapes.pop()
        return self._register(attribute(name, self, r, cc, UNION, children=kids))


class mesh(_AttributeOwner):
    numInstances = 1

    def __init__(self, name: str, scene_):
        super().__init__(name, scene_)
        self.primitives: Dict[str, object] = {}

    def addPrimitive(self, name: str, numInstances: int = 0, isDynamic: bool = False) -> primitive:
        if name in self.primitives:
            raise ValueError(f"mesh {self.fullName}: primitive {name} exists")
        p = primitive(name, self, numInstances, isDynamic)
        self.primitives[name] = p
        setattr(self, name, p)
        return p

    def addPrimitiveUnion(self, name: str, children: List[primitive]) -> primitiveUnion:
        if name in self.primitives:
            raise ValueError(f"mesh {self.fullName}: primitive {name} exists")
        u = primitiveUnion(name, self, children)
        self.primitives[name] = u
        setattr(self, name, u)
        return u


class scene(_AttributeOwner):
    """Top-level container and Newton-step facade (``scene/scene.pyx``)."""
    numInstances = 1

    def __init__(self, name: str):
        super().__init__(name, None)
        self.meshes: Dict[str, mesh] = {}
        self._pre_targets = set()
        from .minimizer import minimizer
        self.minimizer = minimizer()

    def addMesh(self, name: str) -> mesh:
        m = mesh(name, self)
        self.meshes[name] = m
        setattr(self, name, m)
        return m

    def addEnergy(self, e: attribute, targets=(), projection_method: int = 1,
                  save_intermediate: bool = False, gradient_only: bool = False,
                  dynamic_instances: bool = False, separate_hessian_jacobian: bool = False,
                  grouped_add: bool = False, lto: bool = False):
        """``scene.pyx:155`` (same argument order). ``save_intermediate``,
        ``separate_hessian_jacobian``, ``grouped_add`` and ``lto`` select reference
        code-generation variants; every energy here is a hand-written kernel, so they are
        accepted and do not change the result. ``gradient_only=True`` raises
        NotImplementedError, as the reference minimizer does (``minimizer.pyx:188-189``).
        Every attribute in ``targets`` must later be passed to ``addMinimizeTarget``
        (``scene.pyx:165-166,190-198``)."""
        if e.name == "":
            raise ValueError("scene.addEnergy: energy attribute must have a name.")
        if not isinstance(grouped_add, bool):
            raise TypeError("scene.addEnergy: grouped_add must be bool.")
        if not isinstance(lto, bool):
            raise TypeError("scene.addEnergy: lto must be bool.")
        if projection_method not in (-1, 0, 1, 2):
            raise ValueError("addEnergy: projection_method must be -1, 0, 1 or 2")
        if e.operator != COMPUTED:
            raise TypeError("addEnergy: expects an energy attribute built from yasps_amd.energies")
        if gradient_only:
            raise NotImplementedError("minimizer.addEnergy: gradient_only is not supported in the "
                                      "Hessian-based minimizer yet.")
        for t in targets:
            self._pre_targets.add(t.fullName)
        self.minimizer.addEnergy(e, list(targets), projection_method,
                                 dynamic_instances or e.isDynamic)

    def addMinimizeTarget(self, wrt: List[attribute]):
        """``scene.pyx:190-198``. Every target's size must be a multiple of 3: the device
        energies act on 3-vector vertices and on the 9 + 3 segments of affine bodies, so an
        element never has a 1-scalar compressed Hessian and the reference's N == 1
        projection rule (``hessianKernelHeader.pyx:149-152``) has no kernel here."""
        for t in wrt:
            if t.size % 3:
                raise NotImplementedError(
                    f"scene.addMinimizeTarget: {t.fullName} has {t.size} scalars per instance; "
                    "targets of 3-vectors (vertices) and 9 / 3 (affine bodies) are supported")
        names = {t.fullName for t in wrt}
        if len(names) != len(wrt):
            raise ValueError("scene.addMinimizeTarget: target contains duplicate attributes.")
        if not self._pre_targets.issubset(names):
            missing = self._pre_targets - names
            raise ValueError(f"scene.addMinimizeTarget: target is missing attributes {missing} "
                             "that are required by the energies added.")
        self.minimizer.addWrt(wrt)

    def minimizeEnergy(self, tolerance: float = 1e-3, maxIterations: int = 20000):
        """Assemble gradient and Hessian, solve H x = g from a zero initial guess and
        return the solution as one device slice per target (``scene.pyx:171-176``)."""
        return self.minimizer.computeSolution(tolerance=tolerance, maxIterations=maxIterations)

    @property
    def gradient(self):
        return self.minimizer.gradient

    @property
    def gradientSegments(self):
        return self.minimizer.gradientSegments

    @property
    def diagonal(self):
        """Scalar diagonal of the assembled Hessian (``scene.pyx:186-188``; the reference
        fills it in its FullProject element kernels, ``hessianKernelFullProject.pyx:187-192``)."""
        return self.minimizer.diagonal

    @property
    def hessian(self):
        return self.minimizer.activeHessian
