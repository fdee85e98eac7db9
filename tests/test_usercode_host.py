"""Run-time user energies on the CPU: the emitted element function (the code every user
module compiles) built for the host and checked against the SymPy expression it was
generated from (energy, gradient and packed Hessian at random points) and against central
differences; the device module of the spring example built by hipcc, loaded, and its
entry points present; argument validation."""
import ctypes as C
import subprocess

import numpy as np
import pytest
import sympy as sp

from yasps_amd import usercode
from yasps_amd.codegen.gen_energies import emit_energy


def _spring(x, p, s):
    d = sp.Matrix(x[3:6]) - sp.Matrix(x[0:3])
    return s[0] / 2 * (sp.sqrt(d.dot(d)) - p[0]) ** 2


def _area_tri(x, p, s):
    # a 3-vertex energy with two scene parameters and no element parameter
    a, b, c = (sp.Matrix(x[3 * i:3 * i + 3]) for i in range(3))
    n = (b - a).cross(c - a)
    return s[0] * (sp.sqrt(n.dot(n)) / 2 - s[1]) ** 2 + sp.Rational(1, 10) * (a - c).dot(a - c)


CASES = [("spring", _spring, 6, 1, 1), ("area_tri", _area_tri, 9, 0, 2)]


def _host_function(tmp_path, name, fn, n, npar, nsc):
    code, _ = emit_energy(name, fn, n, npar, nsc, qualifier="static inline")
    src = tmp_path / f"{name}.cpp"
    src.write_text("#include <cmath>\n" + code + f"""
extern "C" void eval(const double* x, const double* p, const double* s, double* E, double* g,
                     double* H) {{ gen_{name}(x, p, s, *E, g, H); }}
""")
    so = tmp_path / f"{name}.so"
    subprocess.run(["g++", "-O2", "-shared", "-fPIC", str(src), "-o", str(so)], check=True)
    lib = C.CDLL(str(so))
    lib.eval.restype = None
    lib.eval.argtypes = [C.c_void_p] * 6
    return lib


def _call(lib, x, p, s, n):
    E = np.zeros(1)
    g = np.zeros(n)
    H = np.zeros(n * (n + 1) // 2)
    lib.eval(*(a.ctypes.data for a in (x, p, s, E, g, H)))
    return E[0], g, H


@pytest.mark.parametrize("name,fn,n,npar,nsc", CASES)
def test_emitted_function_matches_its_expression(tmp_path, name, fn, n, npar, nsc):
    lib = _host_function(tmp_path, name, fn, n, npar, nsc)
    xs = sp.symbols(f"x0:{n}", real=True)
    ps = sp.symbols(f"p0:{max(npar, 1)}", real=True)
    ss = sp.symbols(f"s0:{max(nsc, 1)}", real=True)
    E = fn(xs, ps, ss)
    fE = sp.lambdify((xs, ps, ss), E, "numpy")
    fg = sp.lambdify((xs, ps, ss), [sp.diff(E, v) for v in xs], "numpy")
    fH = sp.lambdify((xs, ps, ss), sp.hessian(E, xs), "numpy")
    rng = np.random.default_rng(1)
    iu = np.triu_indices(n)
    for _ in range(4):
        x = rng.standard_normal(n)
        p = rng.uniform(0.5, 1.5, max(npar, 1))
        s = rng.uniform(0.5, 2.0, max(nsc, 1))
        e, g, H = _call(lib, x, p, s, n)
        ref_g = np.asarray(fg(x, p, s), float)
        ref_H = np.asarray(fH(x, p, s), float)
        assert abs(e - fE(x, p, s)) <= 1e-12 * max(1.0, abs(e))
        assert np.abs(g - ref_g).max() <= 1e-12 * max(1.0, np.abs(ref_g).max())
        assert np.abs(H - ref_H[iu]).max() <= 1e-11 * max(1.0, np.abs(ref_H).max())
        # and the derivatives are those of the energy itself: central differences
        h = 1e-6
        fd = np.array([(_call(lib, x + h * np.eye(n)[i], p, s, n)[0]
                        - _call(lib, x - h * np.eye(n)[i], p, s, n)[0]) / (2 * h) for i in range(n)])
        assert np.abs(fd - g).max() <= 1e-6 * max(1.0, np.abs(g).max())
        Hfull = np.zeros((n, n))
        Hfull[iu] = H
        Hfull = Hfull + np.triu(Hfull, 1).T
        fdH = np.array([(_call(lib, x + h * np.eye(n)[i], p, s, n)[1]
                         - _call(lib, x - h * np.eye(n)[i], p, s, n)[1]) / (2 * h) for i in range(n)])
        assert np.abs(fdH - Hfull).max() <= 1e-5 * max(1.0, np.abs(Hfull).max())


def test_device_module_builds_and_exports_its_entry_points():
    """The module exports every entry point include/yasps_amd_user_energy.h declares."""
    import re
    from pathlib import Path
    header = Path(__file__).resolve().parents[1] / "include" / "yasps_amd_user_energy.h"
    names = re.findall(r"^int\s+(ysp_\w+)\(", header.read_text(), re.M)
    assert sorted(names) == ["ysp_user_energy_arity", "ysp_user_energy_eval",
                             "ysp_user_energy_param_jacobian", "ysp_user_energy_values"]
    kind = usercode.example("spring")
    lib = kind.library()
    for name in names:
        assert hasattr(lib, name)
    vals = [C.c_uint32() for _ in range(4)]
    assert lib.ysp_user_energy_arity(*(C.byref(v) for v in vals)) == 0
    assert [v.value for v in vals] == [2, 3, 1, 1]
    # the same source maps to the same module file
    assert usercode.example("spring").library_file == kind.library_file


def test_user_energy_validation():
    with pytest.raises(ValueError):
        usercode.UserEnergyKind("bad name", _spring, 2)
    with pytest.raises(ValueError):
        usercode.UserEnergyKind("big", _spring, arity=5, width=3)      # 15 DoF
    with pytest.raises(ValueError):
        usercode.UserEnergyKind("neg", _spring, 2, element_params=-1)
