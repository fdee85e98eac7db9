"""Write tests/golden/contact_small.npz: the element gradients and (unprojected, upper
packed) Hessians of the cloth and barrier terms of the small contact scene used by
tests/test_gpu_contact.py, from the oracle's SymPy restatement
(oracle/element_energies.py; a few minutes of SymPy differentiation, which is why the
values are stored). The scene is contact_scene(**SCENE) from yasps_amd/scenes.py
(seeded); a checksum of its positions and connectivity guards against drift. Run from the
repository root:

    python tests/golden/make_contact_golden.py
"""
from pathlib import Path
import sys

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
from oracle import contact  # noqa: E402
from yasps_amd.scenes import contact_scene  # noqa: E402

OUT = Path(__file__).with_name("contact_small.npz")
SCENE = dict(bodies_x=2, bodies_z=1, cells=4, radius=0.12, cloth_m=9, max_pairs=10 ** 9,
             seed=0)


def checksum(cs):
    parts = [cs.positions(), cs.rest, cs.tets, cs.triangles, cs.hinges, cs.pt, cs.ee]
    return np.array([float(np.asarray(p, dtype=np.float64).sum()) for p in parts])


def pack(H):
    n = H.shape[1]
    iu = np.triu_indices(n)
    return H[:, iu[0], iu[1]]


def main():
    cs = contact_scene(**SCENE)
    terms = contact.element_terms(cs)
    out = {"checksum": checksum(cs)}
    for kind, (g, H) in terms.items():
        out[f"{kind}_g"] = g
        out[f"{kind}_H"] = pack(H)
    np.savez_compressed(OUT, **out)
    print("wrote", OUT, {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
