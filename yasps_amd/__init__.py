"""yasps_amd: MI355X-native Newton linear-solve path with the YASPS Python API.

The scene/energy/solver interface mirrors the reference (``scene``, ``mesh.addPrimitive``,
``primitive.addAttribute/addConnectivity``, ``scene.addEnergy``, ``addMinimizeTarget``,
``minimizeEnergy``, ``solver``); the numerics run in hand-written HIP kernels for gfx950
behind the C ABI of ``include/yasps_amd.h``.
"""
from .core import scene, mesh, primitive, primitiveUnion, connectivity, attribute  # noqa: F401
from .energies import stable_neo_hookean, inertia  # noqa: F401
from .hessian import hessian, vector  # noqa: F401
from .solver import solver  # noqa: F401
from .minimizer import minimizer  # noqa: F401

__all__ = ["scene", "mesh", "primitive", "primitiveUnion", "connectivity", "attribute",
           "stable_neo_hookean", "inertia", "hessian", "vector", "solver", "minimizer"]
