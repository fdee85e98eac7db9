"""Write tests/golden/element_energies.npz: seeded element inputs of every generated
energy kind (current positions X, element parameters P, scene parameters s and, for
bending and membranes, the rest positions R that P comes from, thickness 0.003) with their
FP64 energy, gradient and (unprojected) Hessian from the oracle's SymPy restatement
(oracle/element_energies.py). Run from the repository root:

    python tests/golden/make_energy_golden.py
"""
from pathlib import Path
import sys

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
from oracle import element_energies as ee  # noqa: E402

OUT = Path(__file__).with_name("element_energies.npz")
COUNT = 24
THICKNESS = 0.003


def inputs(kind, rng):
    arity, width, npar, nsc = ee.SHAPES[kind]
    if kind in ("point_point", "point_edge", "point_triangle", "edge_edge"):
        # element-sized primitives near each other: squared distances O(dhat)
        base = rng.uniform(-1, 1, (COUNT, 1, 3))
        X = base + 0.04 * rng.standard_normal((COUNT, arity, 3))
        return X.reshape(COUNT, -1), np.zeros((COUNT, 0)), np.array([4e-3, 1e3]), None
    if kind == "bending":
        q0 = rng.standard_normal((COUNT, 3))
        q1 = q0 + np.array([1.0, 0.0, 0.0]) + 0.1 * rng.standard_normal((COUNT, 3))
        q2 = q0 + np.array([0.5, 0.8, 0.0]) + 0.1 * rng.standard_normal((COUNT, 3))
        q3 = q0 + np.array([0.5, -0.8, 0.2]) + 0.1 * rng.standard_normal((COUNT, 3))
        rest = [q0, q1, q2, q3]
        P = ee.bending_rest(*rest)
        X = np.concatenate([q + 0.05 * rng.standard_normal(q.shape) for q in rest], axis=1)
        return X, P, np.array([2.5]), np.concatenate(rest, axis=1)
    if kind in ("baraff_witkin", "stvk_membrane"):
        v0 = rng.standard_normal((COUNT, 3))
        v1 = v0 + np.array([1.0, 0.0, 0.1]) + 0.1 * rng.standard_normal((COUNT, 3))
        v2 = v0 + np.array([0.2, 0.1, 1.0]) + 0.1 * rng.standard_normal((COUNT, 3))
        P = ee.triangle_rest(v0, v1, v2, THICKNESS)
        X = np.concatenate([v + 0.08 * rng.standard_normal(v.shape) for v in (v0, v1, v2)], axis=1)
        s = np.array([1e3, 1e2, 1e-2, 0.0, 0.0]) if kind == "baraff_witkin" else np.array([2e3, 5e3, 1e-2])
        return X, P, s, np.concatenate([v0, v1, v2], axis=1)
    if kind == "affine_orthogonality":
        Q = np.linalg.qr(rng.standard_normal((COUNT, 3, 3)))[0]
        X = (Q + 0.1 * rng.standard_normal((COUNT, 3, 3))).reshape(COUNT, 9)
        return X, np.zeros((COUNT, 0)), np.array([1e4]), None
    raise KeyError(kind)


def main():
    rng = np.random.default_rng(20261202)
    data = {}
    for kind in ee.KINDS:
        X, P, s, R = inputs(kind, rng)
        E, g, H = ee.evaluate(kind, X, P, s)
        data[f"{kind}_X"], data[f"{kind}_P"], data[f"{kind}_s"] = X, P, s
        if R is not None:
            data[f"{kind}_R"] = R
        data[f"{kind}_E"], data[f"{kind}_g"], data[f"{kind}_H"] = E, g, H
        print(kind, X.shape, float(np.abs(H).max()), flush=True)
    np.savez_compressed(OUT, **data)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
