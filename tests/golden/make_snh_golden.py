"""Write tests/golden/snh_cube6.json: Stable Neo-Hookean energy, gradient and Hessian of
six tets evaluated exactly with SymPy (a computer-algebra system, not the reference).

Inputs are the fixture literals of the reference script
``tests/energy/check_stable_neo_hookean.py:60-104`` (signed volumes, inverse rest
matrices IB, deformed tet positions, tet-to-vertex indices, mu = 2000, lam = 1000). The
energy follows ``examples/one_bunny/helpers.py:42-67`` with dt^2 = 1:
    FI = F^T IB with F rows = (x1-x0, x2-x0, x3-x0), I_C = tr(FI^T FI), J = det FI,
    E = vol * (mu/2 (I_C - 3) - mu/2 log(I_C + 1) + lam/2 (J - (1 + 3/4 mu/lam))^2).
The script also assembles the 24x24 Hessian and the gradient over the 8 vertices.
Run: python tests/golden/make_snh_golden.py
"""
import json
from pathlib import Path

import numpy as np
import sympy as sp

VOLS = [10.66666667, 10.66666667, -10.66666667, -10.66666667, -10.66666667, 10.66666667]
IBS = [0.25, 0.0, 0.0, 0.0, 0.25, -0.25, 0.0, 0.0, 0.25,
       0.25, 0.25, 0.0, -0.25, 0.0, -0.25, 0.0, -0.0, 0.25,
       0.25, -0.0, -0.0, -0.0, -0.25, 0.25, -0.0, 0.25, -0.0,
       0.25, -0.0, 0.25, -0.25, -0.25, 0.0, -0.0, 0.25, -0.0,
       -0.0, 0.25, -0.25, 0.25, -0.0, 0.25, -0.25, -0.0, -0.0,
       0.0, -0.25, 0.25, 0.25, 0.25, -0.0, -0.25, 0.0, 0.0]
TET_POSITIONS = [0., 0., 0., 2., 0., 0., 0., 2., 0., 0., 2., 2.,
                 2., 0., 0., 2., 1.87938524, 0.68404029, 0., 2., 0., 0., 2., 2.,
                 0., 0., 0., 2., 0., 0., 0., 0., 2., 0., 2., 2.,
                 2., 0., 0., 2., -0.68404029, 1.87938524, 0., 0., 2., 0., 2., 2.,
                 2., 0., 0., 2., 1.87938524, 0.68404029, 2., 1.19534495, 2.56342553, 0., 2., 2.,
                 2., 0., 0., 2., -0.68404029, 1.87938524, 2., 1.19534495, 2.56342553, 0., 2., 2.]
INDICES = [[0, 1, 3, 7], [1, 2, 3, 7], [0, 1, 4, 7], [1, 5, 4, 7], [1, 2, 6, 7], [1, 5, 6, 7]]
MU, LAM = 2000.0, 1000.0


def main():
    X = sp.Matrix(4, 3, lambda a, i: sp.Symbol(f"x{a}{i}"))
    IB = sp.Matrix(3, 3, lambda i, j: sp.Symbol(f"ib{i}{j}"))
    vol, mu, lam = sp.symbols("vol mu lam")
    Frows = sp.Matrix.vstack(*[(X.row(k + 1) - X.row(0)) for k in range(3)])
    FI = Frows.T * IB
    ic = (FI.T * FI).trace()
    J = FI.det()
    energy = vol * (mu / 2 * (ic - 3) - mu / 2 * sp.log(ic + 1)
                    + lam / 2 * (J - (1 + sp.Rational(3, 4) * mu / lam)) ** 2)
    xs = list(X)
    grad = [sp.diff(energy, v) for v in xs]
    hess = [[sp.diff(g, v) for v in xs] for g in grad]
    out = {"mu": MU, "lam": LAM, "vols": VOLS, "IB": IBS, "tet_positions": TET_POSITIONS,
           "indices": INDICES, "energy": [], "gradient": [], "hessian": []}
    gsum = np.zeros(24)
    hsum = np.zeros((24, 24))
    esum = 0.0
    for t in range(6):
        subs = {vol: sp.Float(VOLS[t], 30), mu: sp.Float(MU, 30), lam: sp.Float(LAM, 30)}
        for a in range(4):
            for i in range(3):
                subs[X[a, i]] = sp.Float(TET_POSITIONS[12 * t + 3 * a + i], 30)
        for i in range(3):
            for j in range(3):
                subs[IB[i, j]] = sp.Float(IBS[9 * t + 3 * i + j], 30)
        e = float(energy.evalf(30, subs=subs))
        g = np.array([float(v.evalf(30, subs=subs)) for v in grad])
        H = np.array([[float(v.evalf(30, subs=subs)) for v in row] for row in hess])
        out["energy"].append(e)
        out["gradient"].append(g.tolist())
        out["hessian"].append(H.reshape(-1).tolist())
        esum += e
        for a, va in enumerate(INDICES[t]):
            gsum[3 * va:3 * va + 3] += g[3 * a:3 * a + 3]
            for b, vb in enumerate(INDICES[t]):
                hsum[3 * va:3 * va + 3, 3 * vb:3 * vb + 3] += H[3 * a:3 * a + 3, 3 * b:3 * b + 3]
    out["energy_sum"] = esum
    out["gradient_sum"] = gsum.tolist()
    out["hessian_sum"] = hsum.reshape(-1).tolist()
    path = Path(__file__).with_name("snh_cube6.json")
    path.write_text(json.dumps(out))
    print(f"wrote {path}")


if __name__ == "__main__":
    main()
