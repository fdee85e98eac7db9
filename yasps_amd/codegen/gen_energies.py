"""Generate the element energy functions of ``csrc/generated_energies.hpp`` with SymPy.

Run ``python -m yasps_amd.codegen.gen_energies`` (SymPy needed only here; the generated
headers are committed and the build does not run this); ``--mixed`` writes
``csrc/generated_mixed.hpp``, the derivatives of the gradients w.r.t. the parameters. Each energy is written once as a
SymPy expression of the element's vertex positions x (arity x 3, row-major), per-element
parameters p and scene parameters s; the script differentiates it exactly, shares common
subexpressions (sympy.cse) and emits

    __device__ void gen_<name>(const double* x, const double* p, const double* s,
                               double& E, double* g, double* H)

with g the gradient (3 arity entries) and H the packed upper triangle of the Hessian
(row-major, entry (i, j), i <= j, at i N - i (i - 1) / 2 + j - i).

Energies (reference definitions, restated):
* IPC barriers on squared distances, kappa (d - dhat)^2 log^2(d / dhat)
  (examples/one_bunny/helpers.py:74-120): point-point, point-edge (point to line),
  point-triangle (point to plane), edge-edge (line to line); s = [dhat, kappa].
* discrete-shell bending k (theta - theta0)^2 |e0| with the dihedral angle of
  examples/many_bunnies_one_cloth/helpers.py:193-212; p = [theta0, |e0|], s = [k].
* Baraff-Witkin stretch/shear (helpers.py:239-259): F = [x1-x0, x2-x0] Minv,
  area dt^2 (ks (cu (|F a|-1)^2 + cv (|F b|-1)^2) + kh (F a . F b)^2); the branch
  coefficients cu, cv (0.01 below unit stretch, else 1) are piecewise constants chosen
  from the current F by the kernel; p = [Minv (4), area], s = [ks, kh, dt, cu, cv]
  (cu, cv filled per element).
* StVK membrane (no reference counterpart): area dt^2 (mu |E|^2 + lam/2 tr(E)^2),
  E = (F^T F - I) / 2; p = [Minv (4), area], s = [mu, lam, dt].
* affine-body orthogonality 1/2 |A^T A - I|_F^2 (one_bunny_partial_abd/helpers.py:132-139)
  on the 9 entries of A (row-major), times a stiffness s = [k].
* Stable Neo-Hookean tet (examples/one_bunny/helpers.py:42-67) for the generic path
  (vertices mapped through affine bodies): p = [Dm^-1 (9, row-major), vol],
  s = [mu, lam, dt]; F = Ds Dm^-1 with Ds = [x1-x0, x2-x0, x3-x0].
* inertia 1/2 m |x - (x_prev + v dt - (0, 9.8 dt^2, 0))|^2 (helpers.py:69-71) for the
  generic path: p = [x_prev (3), v (3), m], s = [dt].
"""
from __future__ import annotations

import sys
from pathlib import Path

import sympy as sp

OUT = Path(__file__).resolve().parents[1] / "csrc" / "generated_energies.hpp"


def vec(xs, i):
    return sp.Matrix(xs[3 * i:3 * i + 3])


def barrier(d, s):
    dhat, kappa = s[0], s[1]
    return kappa * (d - dhat) ** 2 * sp.log(d / dhat) ** 2


def e_point_point(x, p, s):
    p0, p1 = vec(x, 0), vec(x, 1)
    e = p1 - p0
    return barrier(e.dot(e), s)


def e_point_edge(x, p, s):
    p0, p1, p2 = vec(x, 0), vec(x, 1), vec(x, 2)
    c = (p1 - p0).cross(p2 - p0)
    e = p2 - p1
    return barrier(c.dot(c) / e.dot(e), s)


def e_point_triangle(x, p, s):
    p0, p1, p2, p3 = (vec(x, i) for i in range(4))
    b = (p2 - p1).cross(p3 - p1)
    atb = (p0 - p1).dot(b)
    return barrier(atb * atb / b.dot(b), s)


def e_edge_edge(x, p, s):
    p0, p1, p2, p3 = (vec(x, i) for i in range(4))
    b = (p1 - p0).cross(p3 - p2)
    atb = (p2 - p0).dot(b)
    return barrier(atb * atb / b.dot(b), s)


def edge_theta(q0, q1, q2, q3):
    n0 = (q0 - q2).cross(q1 - q2)
    n1 = (q1 - q3).cross(q0 - q3)
    axis = q1 - q0
    num = n0.cross(n1).dot(axis) / sp.sqrt(axis.dot(axis))
    den = n0.dot(n1) + sp.sqrt(n0.dot(n0)) * sp.sqrt(n1.dot(n1))
    return 2 * sp.atan2(num, den)


def e_bending(x, p, s):
    q = [vec(x, i) for i in range(4)]
    t = edge_theta(*q)
    return s[0] * (t - p[0]) ** 2 * p[1]


def _membrane_f(x, p):
    x0, x1, x2 = vec(x, 0), vec(x, 1), vec(x, 2)
    D = sp.Matrix.hstack(x1 - x0, x2 - x0)           # 3 x 2
    Minv = sp.Matrix([[p[0], p[1]], [p[2], p[3]]])
    return D * Minv


def e_baraff_witkin(x, p, s):
    F = _membrane_f(x, p)
    fu, fv = F[:, 0], F[:, 1]
    i5u = sp.sqrt(fu.dot(fu))
    i5v = sp.sqrt(fv.dot(fv))
    i6 = fu.dot(fv)
    ks, kh, dt, cu, cv = s[0], s[1], s[2], s[3], s[4]
    stretch = cu * (i5u - 1) ** 2 + cv * (i5v - 1) ** 2
    return (ks * stretch + kh * i6 * i6) * p[4] * dt * dt


def e_stvk_membrane(x, p, s):
    F = _membrane_f(x, p)
    E = (F.T * F - sp.eye(2)) / 2
    mu, lam, dt = s[0], s[1], s[2]
    tr = E[0, 0] + E[1, 1]
    e2 = sum(E[i, j] ** 2 for i in range(2) for j in range(2))
    return p[4] * dt * dt * (mu * e2 + lam / 2 * tr * tr)


def e_affine_orthogonality(x, p, s):
    A = sp.Matrix(3, 3, list(x[:9]))
    D = A.T * A - sp.eye(3)
    return s[0] * sp.Rational(1, 2) * sum(D[i, j] ** 2 for i in range(3) for j in range(3))


def e_snh(x, p, s):
    x0, x1, x2, x3 = (vec(x, i) for i in range(4))
    Ds = sp.Matrix.hstack(x1 - x0, x2 - x0, x3 - x0)
    Dmi = sp.Matrix(3, 3, list(p[:9]))
    F = Ds * Dmi
    mu, lam, dt = s[0], s[1], s[2]
    ic = sum(F[i, j] ** 2 for i in range(3) for j in range(3))
    J = F.det()
    alpha = 1 + sp.Rational(3, 4) * mu / lam
    return dt * dt * p[9] * (mu / 2 * (ic - 3) - mu / 2 * sp.log(ic + 1) + lam / 2 * (J - alpha) ** 2)


def e_inertia(x, p, s):
    dt = s[0]
    xv = vec(x, 0)
    xp = sp.Matrix(p[0:3])
    v = sp.Matrix(p[3:6])
    tgt = xp + v * dt - sp.Matrix([0, sp.Rational(98, 10) * dt * dt, 0])
    d = xv - tgt
    return p[6] * d.dot(d) / 2


# name: (function, number of DoF, per-element params, scene params)
ENERGIES = {
    "point_point": (e_point_point, 6, 0, 2),
    "point_edge": (e_point_edge, 9, 0, 2),
    "point_triangle": (e_point_triangle, 12, 0, 2),
    "edge_edge": (e_edge_edge, 12, 0, 2),
    "bending": (e_bending, 12, 2, 1),
    "baraff_witkin": (e_baraff_witkin, 9, 5, 5),
    "stvk_membrane": (e_stvk_membrane, 9, 5, 3),
    "affine_orthogonality": (e_affine_orthogonality, 9, 0, 1),
    "snh": (e_snh, 12, 10, 3),
    "inertia": (e_inertia, 3, 7, 1),
}


def generate(name):
    fn, n, npar, nsc = ENERGIES[name]
    return emit_energy(name, fn, n, npar, nsc)


def energy_symbols(n, npar, nsc):
    """The SymPy symbols x, p, s an energy of n DoF, npar element and nsc scene
    parameters is written in."""
    return (sp.symbols(f"x0:{n}", real=True), sp.symbols(f"p0:{max(npar, 1)}", real=True),
            sp.symbols(f"s0:{max(nsc, 1)}", real=True))


def emit_energy(name, fn, n, npar, nsc, qualifier="__device__ __forceinline__", energy=None):
    """C++ of gen_<name>(x, p, s, E, g, H) for the energy fn(x, p, s) of n DoF, npar
    element and nsc scene parameters (SymPy symbols of ``energy_symbols``; ``energy``: the
    expression already built from them), declared with ``qualifier``; returns (code,
    number of shared subexpressions)."""
    x, p, s = energy_symbols(n, npar, nsc)
    E = sp.sympify(fn(x, p, s) if energy is None else energy)
    g = [sp.diff(E, xi) for xi in x]
    H = [sp.diff(g[i], x[j]) for i in range(n) for j in range(i, n)]
    repl, red = sp.cse([E] + g + H, symbols=sp.numbered_symbols("t"), optimizations="basic")
    subs = {xi: sp.Symbol(f"x[{i}]") for i, xi in enumerate(x)}
    subs.update({pi: sp.Symbol(f"p[{i}]") for i, pi in enumerate(p)})
    subs.update({si: sp.Symbol(f"s[{i}]") for i, si in enumerate(s)})
    lines = [f"// {name}: {n} DoF, {npar} element parameters, {nsc} scene parameters",
             f"constexpr int kGen_{name}_N = {n};",
             f"{qualifier} void gen_{name}(const double* __restrict__ x, "
             "const double* __restrict__ p, const double* __restrict__ s, double& E, "
             "double* __restrict__ g, double* __restrict__ H) {"]
    for sym, expr in repl:
        lines.append(f"  const double {sym} = {sp.ccode(expr.xreplace(subs))};")
    lines.append(f"  E = {sp.ccode(red[0].xreplace(subs))};")
    for i in range(n):
        lines.append(f"  g[{i}] = {sp.ccode(red[1 + i].xreplace(subs))};")
    for k in range(len(H)):
        lines.append(f"  H[{k}] = {sp.ccode(red[1 + n + k].xreplace(subs))};")
    lines.append("}")
    return "\n".join(lines), len(repl)


# kinds whose element parameters p are energy arguments (not rest-shape data): their mixed
# derivatives d g / d p are generated as well as d g / d s
MIXED_P = {"inertia"}
OUT_MIXED = OUT.with_name("generated_mixed.hpp")


def generate_mixed(name):
    """d g / d s (n x S, row-major) of every kind and d g / d p (n x P) of the kinds in
    MIXED_P: the mixed second derivatives of the energy w.r.t. its DoF and its parameters,
    the element blocks of a rectangular Jacobian d g / d theta once multiplied by the
    parameters' own derivatives d (p, s) / d theta."""
    fn, n, npar, nsc = ENERGIES[name]
    return emit_mixed(name, fn, n, npar, nsc, npar if name in MIXED_P else 0)


def emit_mixed(name, fn, n, npar, nsc, wp, qualifier="__device__ __forceinline__",
               energy=None):
    """C++ of mixed_<name>(x, p, s, Gp, Gs): d g / d p for the first ``wp`` element
    parameters (n x wp, row-major) and d g / d s (n x nsc) of the energy fn(x, p, s)
    (``energy``: the expression already built from ``energy_symbols``)."""
    x, p, s = energy_symbols(n, npar, nsc)
    E = sp.sympify(fn(x, p, s) if energy is None else energy)
    s = s[:nsc]
    g = [sp.diff(E, xi) for xi in x]
    Gp = [sp.diff(g[i], p[j]) for i in range(n) for j in range(wp)]
    Gs = [sp.diff(g[i], s[j]) for i in range(n) for j in range(nsc)]
    repl, red = sp.cse(Gp + Gs, symbols=sp.numbered_symbols("t"), optimizations="basic")
    subs = {xi: sp.Symbol(f"x[{i}]") for i, xi in enumerate(x)}
    subs.update({pi: sp.Symbol(f"p[{i}]") for i, pi in enumerate(p)})
    subs.update({si: sp.Symbol(f"s[{i}]") for i, si in enumerate(s)})
    lines = [f"// {name}: d g / d p ({n} x {wp}), d g / d s ({n} x {nsc})",
             f"constexpr int kMixed_{name}_P = {wp};",
             f"{qualifier} void mixed_{name}(const double* __restrict__ x, "
             "const double* __restrict__ p, const double* __restrict__ s, "
             "double* __restrict__ Gp, double* __restrict__ Gs) {"]
    for sym, expr in repl:
        lines.append(f"  const double {sym} = {sp.ccode(expr.xreplace(subs))};")
    for k in range(len(Gp)):
        lines.append(f"  Gp[{k}] = {sp.ccode(red[k].xreplace(subs))};")
    for k in range(len(Gs)):
        lines.append(f"  Gs[{k}] = {sp.ccode(red[len(Gp) + k].xreplace(subs))};")
    if not Gp:
        lines.append("  (void)Gp;")
    lines.append("}")
    return "\n".join(lines), len(repl)


def main_mixed(names=None):
    names = names or list(ENERGIES)
    parts = ["// Generated by yasps_amd/codegen/gen_energies.py (mixed derivatives of the energies",
             "// of generated_energies.hpp w.r.t. their parameters); do not edit by hand.",
             "#pragma once",
             "#include <cmath>",
             "",
             "namespace ysp {",
             "namespace gen {",
             ""]
    for name in names:
        code, nrep = generate_mixed(name)
        print(f"{name} (mixed): {nrep} shared subexpressions", file=sys.stderr)
        parts.append(code)
        parts.append("")
    parts += ["}  // namespace gen", "}  // namespace ysp", ""]
    OUT_MIXED.write_text("\n".join(parts))
    print(f"wrote {OUT_MIXED}", file=sys.stderr)


def main(names=None):
    names = names or list(ENERGIES)
    parts = ["// Generated by yasps_amd/codegen/gen_energies.py from the SymPy energy definitions",
             "// listed there; do not edit by hand.",
             "#pragma once",
             "#include <cmath>",
             "",
             "namespace ysp {",
             "namespace gen {",
             ""]
    for name in names:
        code, nrep = generate(name)
        print(f"{name}: {nrep} shared subexpressions", file=sys.stderr)
        parts.append(code)
        parts.append("")
    parts += ["}  // namespace gen", "}  // namespace ysp", ""]
    OUT.write_text("\n".join(parts))
    print(f"wrote {OUT}", file=sys.stderr)


if __name__ == "__main__":
    if sys.argv[1:2] == ["--mixed"]:
        main_mixed(sys.argv[2:] or None)
    else:
        main(sys.argv[1:] or None)
