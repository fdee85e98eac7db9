"""Newton linear-solve benchmark on one MI355X.

One step = one Newton linear solve of the tet-cube scene (Stable Neo-Hookean + inertia,
bottom face fixed): element gradients/Hessians with PSD projection, scatter into the
3x3 block storage, block-Jacobi setup and PCG to a true relative residual of 1e-6.
Inputs stay resident in HBM; the static sparsity layout is built once before timing.

Prints one JSON line (see the keys below). ``--gpus N`` with N > 1 exits: this build
runs on one GPU.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=48, help="cube resolution (n^3 hexes x 6 tets)")
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--spmv-reps", type=int, default=50)
    ap.add_argument("--no-forms", action="store_true",
                    help="skip the mixed-precision and MAS legs")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving stream")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the material adjoint")
    ap.add_argument("--no-dynamic", action="store_true",
                    help="skip the dynamic contact-term leg")
    ap.add_argument("--no-elements", action="store_true",
                    help="skip the per-energy assembly leg")
    ap.add_argument("--no-contact", action="store_true",
                    help="skip the contact scene with affine bodies")
    ap.add_argument("--contact-pairs", type=int, default=1_500_000)
    ap.add_argument("--contact-steps", type=int, default=3)
    ap.add_argument("--adjoint-n", type=int, default=32)
    ap.add_argument("--adjoint-fd", type=int, default=64)
    ap.add_argument("--serving-requests", type=int, default=512)
    ap.add_argument("--kernels-only", action="store_true",
                    help="one solve, the SpMV roofline launches and a calibration copy of a "
                         "known size, no JSON line (for hardware-counter runs)")
    return ap.parse_args()


# HBM bytes per launch of bsr3_full_spmv_kernel<double> measured with hardware counters
# on the same workload (rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE, over
# `bench.py --kernels-only`, each scaled by the counter's reading for a 256 MiB device
# copy: FETCH x1.9998, WRITE x1.0); summary in profiles/spmv_cube48_fp64_pmc.md.
MEASURED_TRAFFIC = {48: 133998758}


def measured_traffic(n):
    return MEASURED_TRAFFIC.get(n)


def calibration_copy(mib=256):
    """A plain device copy of a known size (the reference point for the counters)."""
    from yasps_amd import device as dev
    torch = dev.torch()
    a = torch.ones(mib * 131072, dtype=torch.float64, device=dev.device())
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()


def spmv_roofline(hs, reps):
    """Average duration of the full-BSR 3x3 SpMV kernel (the operator of the PCG loop),
    timed with HIP events on the stream it is launched on, and its algorithmic bytes
    per launch (symmetric upper storage, each stored block read once)."""
    from yasps_amd import _lib, device as dev
    torch = dev.torch()
    lib = _lib.load()
    f = hs.full
    x = dev.to_device(np.linspace(-1.0, 1.0, hs.rows), "float64")
    y = dev.zeros(hs.rows, "float64")
    stream = torch.cuda.current_stream()

    def launch():
        lib.ysp_bsr3_spmv(dev.ptr(f.row_ptr), dev.ptr(f.col), dev.ptr(f.v64), 0, f.nnz, hs.rows,
                          dev.ptr(x), dev.ptr(y), stream.cuda_stream)
    for _ in range(3):
        launch()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(reps):
        launch()
    end.record(stream)
    end.synchronize()
    ms = start.elapsed_time(end) / reps
    # cold: reading 1 GiB between launches evicts L2 and the 256 MB Infinity Cache (with
    # clean lines: a write here would leave dirty lines whose write-back competes with the
    # next launch), so each timed launch streams its operands from HBM
    flush = torch.ones(1 << 27, dtype=torch.float64, device=dev.device())
    cold = []
    for _ in range(min(reps, 20)):
        flush.sum()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        launch()
        e1.record(stream)
        cold.append((e0, e1))
    torch.cuda.synchronize()
    ms_cold = float(np.mean([a.elapsed_time(b) for a, b in cold]))
    del flush
    nb = hs.static.num_blocks
    n = hs.rows
    # algorithmic: 72 B values + 8 B (row, col) per stored upper block, x read, y written
    bytes_alg = nb * (9 * 8 + 8) + n * 16
    # what the mirrored full block-CSR kernel actually streams
    bytes_full = f.nnz * (9 * 8 + 4) + (n // 3 + 1) * 4 + n * 16
    gbs = bytes_alg / (ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": measured_traffic(hs_n(hs)),
            "kernel": "bsr3_full_spmv_kernel<double>", "avg_us": round(ms * 1e3, 2),
            "bytes_per_launch": int(bytes_alg), "full_bsr_bytes_per_launch": int(bytes_full),
            "full_bsr_GBs": round(bytes_full / (ms * 1e-3) / 1e9, 1),
            "cold": {"avg_us": round(ms_cold * 1e3, 2),
                     "achieved": round(bytes_alg / (ms_cold * 1e-3) / 1e9, 1),
                     "frac": round(bytes_alg / (ms_cold * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     "full_bsr_GBs": round(bytes_full / (ms_cold * 1e-3) / 1e9, 1),
                     "note": "1 GiB read between launches (L2 and MALL evicted); single-launch event pairs"}}


def hs_n(hs):
    """Cube resolution of the benchmark Hessian (rows = 3 (n+1)^2 n)."""
    nv = hs.rows // 3
    for n in range(1, 1024):
        if (n + 1) ** 2 * n == nv:
            return n
    return None


def timed_solves(s0, steps, warmup, tol):
    from yasps_amd import device as dev
    torch = dev.torch()
    for _ in range(warmup):
        s0.minimizeEnergy(tolerance=tol)
    torch.cuda.synchronize()
    lat, iters = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        ts = time.perf_counter()
        s0.minimizeEnergy(tolerance=tol)
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - ts)
        iters.append(s0.minimizer.linearSolver.statistics["iterations"])
    torch.cuda.synchronize()
    return time.perf_counter() - t0, lat, iters


def form_leg(cs, args, x64, ref_value, solver, **options):
    """Mixed-precision or MAS leg on the same scene: solves/s, its ratio to the plain FP64
    path, the true FP64 residual and the error against the FP64 solution."""
    from yasps_amd.scenes import build_yasps_scene
    from yasps_amd import device as dev
    torch = dev.torch()
    s1, _ = build_yasps_scene(cs, solver_name=solver, **options)
    s1.minimizeEnergy(tolerance=args.tol)
    torch.cuda.synchronize()
    elapsed, lat, iters = timed_solves(s1, args.steps, args.warmup, args.tol)
    st = s1.minimizer.linearSolver.statistics
    hs = s1.minimizer.activeHessian
    x = s1.minimizer.linearSolver.solution.clone()
    g = hs.gradient.value
    from yasps_amd.hessian import vector
    r = hs.spmv(vector(hs.rows, x)).value - g
    value = args.steps / elapsed
    return {"solver": solver, "options": {k: v for k, v in options.items()},
            "value": round(value, 3), "ms_per_step": round(1e3 * elapsed / args.steps, 3),
            "vs_fp64_jacobi": round(value / ref_value, 3),
            "p50_ms": round(1e3 * float(np.percentile(lat, 50)), 3),
            "p99_ms": round(1e3 * float(np.percentile(lat, 99)), 3),
            "pcg_iterations": int(np.median(iters)),
            "outer_iterations": st.get("outer_iterations"),
            "true_fp64_rel_residual": float(torch.linalg.norm(r) / torch.linalg.norm(g)),
            "rel_err_vs_fp64": float(torch.linalg.norm(x - x64) / torch.linalg.norm(x64))}


def serving_leg(args):
    """Serving stream: requests with n log-uniform in [6, 48] and E log-uniform in
    [1e4, 1e6], each carrying its positions (pinned host payload), first all at once
    (capacity), then Poisson arrivals at 80% of that capacity; p50/p99 latency from
    arrival to solution. The default server overlaps request loading (a loader thread
    and stream) with the batched PCG chunks; the same stream through the sequential
    server (loads between chunks: larger batches in bursts, longer waits) is reported
    beside it."""
    from yasps_amd.serving import (SolveServer, attach_payloads, serving_stream,
                                   poisson_arrivals, summary)
    base = attach_payloads(serving_stream(count=args.serving_requests, seed=args.seed))

    def fresh(arrivals=False):
        return [type(r)(id=r.id, n=r.n, young=r.young, seed=r.seed, positions=r.positions,
                        arrival=r.arrival if arrivals else 0.0) for r in base]

    def run(overlap):
        srv = SolveServer(max_active=32, passes=8, tolerance=args.tol, overlap_loads=overlap)
        t0 = time.perf_counter()
        srv.warm([r.n for r in base])
        # the first closed-loop pass grows the slot pool to the concurrency it needs (static
        # structures of meshes that are in flight more than once); the second one measures
        srv.serve(fresh())
        warm_s = time.perf_counter() - t0
        cap = summary(srv.serve(fresh()))
        rate = 0.8 * cap["solves_per_s"]
        reqs = poisson_arrivals(fresh(), rate, seed=args.seed + 1)
        srv.batch_sizes.clear()
        st = summary(srv.serve(reqs))
        return {"capacity_solves_per_s": round(cap["solves_per_s"], 2),
                "arrival_rate_per_s": round(rate, 2), "value": round(st["solves_per_s"], 2),
                "p50_ms": round(st["p50_ms"], 3), "p99_ms": round(st["p99_ms"], 3),
                "admission_wait_p50_ms": round(st["wait_p50_ms"], 3),
                "admission_wait_p99_ms": round(st["wait_p99_ms"], 3),
                "converged": st["converged"], "mean_pcg_iterations": round(st["mean_iterations"], 1),
                "mean_batch": round(float(np.mean(srv.batch_sizes)), 2),
                "max_batch": int(max(srv.batch_sizes)), "warm_seconds": round(warm_s, 2),
                "slots": int(sum(len(v) for v in srv.pool.values()))}

    out = {"requests": len(base), **run(True),
           "solver": "batched block-Jacobi PCG, FP64, ||g-Hx||/||g|| <= %g" % args.tol}
    out["sequential_loads"] = run(False)
    return out


def adjoint_leg(args):
    """Material inverse on an n^3 cube (bottom fixed, gravity, theta = log E per tet ~
    N(log 1e5, 0.3)): equilibrium, loss against the equilibrium of a second theta, adjoint
    gradient in FP64 and mixed precision, central differences (h = 1e-6) on random
    entries."""
    from yasps_amd.adjoint import MaterialProblem
    prob = MaterialProblem(n=args.adjoint_n)
    rng = np.random.default_rng(args.seed)
    theta = np.log(1e5) + 0.3 * rng.standard_normal(prob.num_tets)
    theta_t = np.log(1e5) + 0.3 * np.random.default_rng(args.seed + 1).standard_normal(prob.num_tets)
    target = prob.equilibrium(theta_t).reshape(-1, 3)[prob.top].copy()
    t0 = time.perf_counter()
    x = prob.equilibrium(theta, tol=1e-13)
    t1 = time.perf_counter()
    g = prob.vjp(x, prob.loss_grad_x(x, target))
    t2 = time.perf_counter()
    pm = MaterialProblem(n=args.adjoint_n, precision="mixed")
    xm = pm.equilibrium(theta, tol=1e-13)
    t3 = time.perf_counter()
    gm = pm.vjp(xm, pm.loss_grad_x(xm, target))
    t4 = time.perf_counter()
    idx = rng.choice(prob.num_tets, min(args.adjoint_fd, prob.num_tets), replace=False)
    h = 1e-6
    err = []
    for i in idx:
        tp, tm = theta.copy(), theta.copy()
        tp[i] += h
        tm[i] -= h
        fd = (prob.loss(prob.equilibrium(tp, x0=x, tol=1e-13), target)
              - prob.loss(prob.equilibrium(tm, x0=x, tol=1e-13), target)) / (2 * h)
        err.append(abs(fd - g[i]))
    scale = float(np.abs(g).max())
    return {"n": args.adjoint_n, "tets": prob.num_tets, "unknowns": 3 * prob.n_free,
            "newton_iterations": prob.newton_iterations,
            "equilibrium_s": round(t1 - t0, 3), "adjoint_fp64_s": round(t2 - t1, 4),
            "adjoint_mixed_s": round(t4 - t3, 4),
            "fd_entries": len(idx), "fd_max_abs_err_rel_to_max_grad": float(max(err) / scale),
            "mixed_vs_fp64_rel": float(np.linalg.norm(gm - g) / np.linalg.norm(g)),
            "loss": prob.loss(x, target)}


def dynamic_leg(args):
    """The n=48 cube with a dynamic point-point barrier over grid-neighbour pairs of moving
    vertices (axis and face-diagonal neighbours, squared distances below d^2 = 1e-3):
    every Newton solve re-indexes the dynamic term (index generation, local compression,
    coordinates, global dedupe, staging layout), evaluates, assembles the static and
    dynamic parts and solves. Reports solves/s with block-Jacobi and with MAS, and the
    re-index time alone."""
    from yasps_amd import device as dev
    from yasps_amd.energies import point_point
    from yasps_amd.scenes import cube_scene, build_yasps_scene
    torch = dev.torch()
    cs = cube_scene(args.n, seed=args.seed)
    s0, hd = build_yasps_scene(cs, threshold_mode="relative_residual")
    n1 = args.n + 1
    nf = cs.n_fixed
    # moving vertices sit on the (n+1)^2 x n grid above the fixed bottom layer
    ijk = np.round(cs.rest[nf:] * args.n).astype(np.int64)
    key = {tuple(v): nf + i for i, v in enumerate(ijk)}
    pairs = []
    for d in ((1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 0), (1, 0, 1), (0, 1, 1)):
        for i, v in enumerate(ijk):
            w = (v[0] + d[0], v[1] + d[1], v[2] + d[2])
            j = key.get(w)
            if j is not None:
                pairs.append((nf + i, j))
    pairs = np.array(pairs, dtype=np.uint32)
    m = hd["mesh"]
    contacts = m.addPrimitive("contacts", numInstances=len(pairs), isDynamic=True)
    c2v = contacts.addConnectivity("c2v", m.vertices, pairs, 2)
    cp = contacts.addAttribute("positions", through=c2v, source=m.vertices["position"])
    pp = contacts.addAttribute("pp", computed_attribute=point_point(cp, 1e-3, 1e-2))
    s0.addEnergy(pp, projection_method=2, dynamic_instances=True)
    total, lat, iters = timed_solves(s0, max(3, args.steps // 2), 2, args.tol)
    hs = s0.minimizer.activeHessian
    reps = 5
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        hs._reindex_dynamic()
    torch.cuda.synchronize()
    reindex_ms = 1e3 * (time.perf_counter() - t0) / reps
    # the same solves with MAS as the forms run it (4 levels; blocks present in both parts
    # summed into its domains)
    s0.minimizer.setSolver("mas", threshold_mode="relative_residual", max_levels=4,
                           coarsest_level_weight=1.0)
    m_total, m_lat, m_iters = timed_solves(s0, max(3, args.steps // 2), 2, args.tol)
    return {"pairs": int(len(pairs)), "dynamic_blocks": int(hs.dynamic.num_blocks),
            "static_blocks": int(hs.static.num_blocks), "grid": n1,
            "value": round(len(lat) / total, 3), "unit": "Newton solves/s",
            "ms_per_step": round(1e3 * total / len(lat), 3),
            "pcg_iterations": int(np.median(iters)), "reindex_ms": round(reindex_ms, 3),
            "mas": {"value": round(len(m_lat) / m_total, 3),
                    "ms_per_step": round(1e3 * m_total / len(m_lat), 3),
                    "pcg_iterations": int(np.median(m_iters))}}


def contact_leg(args):
    """Contact scene with mixed DoFs on one GPU: 40 voxelised balls
    (16^3 cells, radius 0.045, 13,056 tets each) on a jittered 5 x 8 grid, every other one a
    12-DoF affine body (x = t + A r through JOIN/UNION), the others free vertices, resting
    U(0.2, 2) dhat above a 256 x 256 cloth (StVK membrane + dihedral bending, border
    pinned), dhat = 1e-3 on squared distances; active point-triangle and edge-edge barrier
    pairs (a uniform sample of --contact-pairs when more are active), inertia and affine
    orthogonality. One Newton linear solve per step: re-index of the dynamic pair terms,
    element evaluation, PSD projection, assembly (9x9 / 9x3 / 3x9 / 3x3 blocks), the split
    3x3 operator, preconditioner setup and PCG to a true relative residual of --tol: with
    MAS (32-DoF leaf domains, 4 levels, the affine DoFs as 3-DoF nodes; the leg's value)
    and with block-Jacobi (9x9 blocks for the affine matrices)."""
    from yasps_amd import device as dev
    from yasps_amd.scenes import contact_scene, build_contact_scene
    t0 = time.perf_counter()
    cs = contact_scene(max_pairs=args.contact_pairs, seed=args.seed)
    scene_s = time.perf_counter() - t0
    print(f"contact scene: {cs.tets.shape[0]} tets, {cs.pt.shape[0]} point-triangle and "
          f"{cs.ee.shape[0]} edge-edge pairs ({scene_s:.1f} s)", file=sys.stderr, flush=True)
    s0, _ = build_contact_scene(cs, threshold_mode="relative_residual")
    legs = {}
    for name, solver, opts in (("mas", "mas", dict(max_domain_dofs=32, max_levels=4)),
                               ("block_jacobi", "jacobian", {})):
        s0.minimizer.setSolver(solver, threshold_mode="relative_residual", **opts)
        total, lat, iters = timed_solves(s0, args.contact_steps, 1, args.tol)
        st = s0.minimizer.linearSolver.statistics
        print(f"contact {name}: {len(lat)} solves in {total:.2f} s", file=sys.stderr, flush=True)
        legs[name] = {"value": round(len(lat) / total, 3), "unit": "Newton solves/s",
                      "ms_per_step": round(1e3 * total / len(lat), 3),
                      "pcg_iterations": int(np.median(iters)), "converged": bool(st["converged"]),
                      "relative_residual": (float(np.sqrt(st["rr"] / st["rr0"]))
                                            if st.get("rr0") else None)}
        if name == "mas":
            legs[name]["nodes_per_level"] = st.get("nodes_per_level")
    hs = s0.minimizer.activeHessian
    dev.torch().cuda.synchronize()
    return {"bodies": int(cs.n_bodies + cs.n_free_bodies), "affine_bodies": int(cs.n_bodies),
            "tets": int(cs.tets.shape[0]), "cloth_triangles": int(cs.triangles.shape[0]),
            "hinges": int(cs.hinges.shape[0]), "point_triangle_pairs": int(cs.pt.shape[0]),
            "edge_edge_pairs": int(cs.ee.shape[0]), "unknowns": int(hs.rows),
            "static_blocks": int(hs.static.num_blocks),
            "dynamic_blocks": int(hs.dynamic.num_blocks),
            "split_3x3_entries": int(hs.split.nnz) if hs.split is not None else None,
            "value": legs["mas"]["value"], "unit": "Newton solves/s", "solvers": legs,
            "scene_generation_s": round(scene_s, 2)}


def _energy_compute_ms(s0, reps=10):
    """Average time of one assembly (element evaluation, PSD projection, staged assembly)
    by HIP events on the stream the kernels run on."""
    from yasps_amd import device as dev
    torch = dev.torch()
    hs = s0.minimizer.computeNumericValue()          # static structure, warm-up
    stream = torch.cuda.current_stream()
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(reps):
        hs.compute(check=False)
    end.record(stream)
    end.synchronize()
    hs.compute()                                       # flags of the timed assemblies
    return start.elapsed_time(end) / reps


def element_leg(args):
    """Per-energy assembly on meshes the size of the reference's published per-element
    figures (RTX 4090: SNH 79,935 tets, bending 104,065 hinges, Baraff-Witkin 20,000
    triangles): SNH on the n=24 cube (82,944 tets, bottom fixed), dihedral bending on the
    hinges of a 188 x 188 cloth grid (104,533), Baraff-Witkin on a 101 x 101 grid (20,000
    triangles); every vertex of the cloths free. Projection method 2. The reference's
    figures are gradient and Hessian times; these include projection and assembly."""
    from yasps_amd import scene, energies as en, stable_neo_hookean
    from yasps_amd.scenes import cube_scene, build_yasps_scene, cloth_scene
    out = {}
    cs = cube_scene(24, seed=args.seed)
    s0 = scene("snh")
    dt = s0.addConstant("dt", rows=1, cols=1)
    dt.updateValue([cs.dt])
    m = s0.addMesh("cube")
    nf = cs.n_fixed
    fv = m.addPrimitive("fixed_vertices", numInstances=nf)
    mv = m.addPrimitive("moving_vertices", numInstances=cs.n_moving)
    m.addPrimitiveUnion("vertices", [fv, mv])
    fv.addAttribute("position", rows=3, cols=1).updateValue(cs.position[:nf])
    fv.addConstant("rest_position", rows=3, cols=1).updateValue(cs.rest[:nf])
    mv.addAttribute("position", rows=3, cols=1).updateValue(cs.position[nf:])
    mv.addConstant("rest_position", rows=3, cols=1).updateValue(cs.rest[nf:])
    m.vertices.addAttribute("position")
    m.vertices.addAttribute("rest_position")
    mu = m.addConstant("mu", rows=1, cols=1)
    mu.updateValue([cs.mu])
    lam = m.addConstant("lam", rows=1, cols=1)
    lam.updateValue([cs.lam])
    tets = m.addPrimitive("tets", numInstances=cs.tets.shape[0])
    t2v = tets.addConnectivity("tet2v", m.vertices, cs.tets, 4)
    tp = tets.addAttribute("positions", through=t2v, source=m.vertices["position"])
    trp = tets.addAttribute("rest_positions", through=t2v, source=m.vertices["rest_position"])
    snh = tets.addAttribute("snh", computed_attribute=stable_neo_hookean(trp, tp, mu, lam, dt))
    s0.addEnergy(snh, projection_method=2)
    s0.addMinimizeTarget([mv["position"]])
    ms = _energy_compute_ms(s0)
    out["snh"] = {"elements": int(cs.tets.shape[0]), "ms": round(ms, 4),
                  "ns_per_element": round(1e6 * ms / cs.tets.shape[0], 2),
                  "reference_rtx4090": {"elements": 79935, "gradient_ms": 0.19728,
                                        "hessian_ms": 0.59}}
    for kind, mgrid, ref in (("bending", 188, {"elements": 104065, "gradient_ms": 0.17,
                                                "hessian_ms": 2.81}),
                             ("baraff_witkin", 101, {"elements": 20000, "gradient_ms": 0.04,
                                                     "hessian_ms": 0.12})):
        cl = cloth_scene(mgrid, seed=args.seed)
        s1 = scene(kind)
        mc = s1.addMesh("cloth")
        verts = mc.addPrimitive("vertices", numInstances=mgrid * mgrid)
        verts.addAttribute("position", rows=3, cols=1).updateValue(cl.position)
        verts.addConstant("rest_position", rows=3, cols=1).updateValue(cl.rest)
        conn = cl.hinges if kind == "bending" else cl.triangles
        el = mc.addPrimitive("elements", numInstances=conn.shape[0])
        e2v = el.addConnectivity("e2v", verts, conn, conn.shape[1])
        pos = el.addAttribute("positions", through=e2v, source=verts["position"])
        rp = el.addAttribute("rest_positions", through=e2v, source=verts["rest_position"])
        e = (en.bending(pos, rp, 1e-3) if kind == "bending"
             else en.baraff_witkin(rp, pos, 1e3, 1e2, 1e-3, 1e-2))
        att = el.addAttribute(kind, computed_attribute=e)
        s1.addEnergy(att, projection_method=2)
        s1.addMinimizeTarget([verts["position"]])
        ms = _energy_compute_ms(s1)
        out[kind] = {"elements": int(conn.shape[0]), "ms": round(ms, 4),
                     "ns_per_element": round(1e6 * ms / conn.shape[0], 2),
                     "reference_rtx4090": ref}
    return out


def cpu_baseline(cs, values_host, st, tol):
    """The oracle on the host: SNH + inertia assembly timed on 1/8 of the tets (scaled x8)
    and block-Jacobi CG (scipy, one thread) timed in full on the same matrix."""
    from oracle import energies, newton, projection
    import scipy.sparse.linalg as spla
    import contextlib
    try:
        from threadpoolctl import threadpool_limits
        limiter = threadpool_limits(limits=1)
        threads = 1
    except Exception:  # threadpoolctl absent: numpy keeps its default thread count
        limiter = contextlib.nullcontext()
        threads = os.cpu_count() or 1
    with limiter:
        T = cs.tets.shape[0]
        frac = 8
        sub = slice(0, T // frac)
        t0 = time.perf_counter()
        rd = energies.snh_rest_data(cs.rest[cs.tets[sub]])
        g, H = energies.snh_gradient_hessian(cs.position[cs.tets[sub]], rd[0], rd[1], cs.mu,
                                             cs.lam, cs.dt)
        H = projection.project_batch(H, 2)
        t_asm = (time.perf_counter() - t0) * frac
        t0 = time.perf_counter()
        A = newton.bsr_matrix(st, values_host)
        diag = values_host[(st.vert_lookup[:, None] + np.arange(9)).reshape(-1)].reshape(-1, 3, 3)
        inv = projection.jacobi_block_inverse(diag)
        b = np.asarray(cs.extras["gradient"])
        M = spla.LinearOperator(A.shape, matvec=lambda v: np.einsum(
            "bij,bj->bi", inv, v.reshape(-1, 3)).reshape(-1), dtype=np.float64)
        x, info = spla.cg(A, b, rtol=tol, atol=0.0, maxiter=20000, M=M)
        t_cg = time.perf_counter() - t0
        rel = float(np.linalg.norm(b - A @ x) / np.linalg.norm(b))
    total = t_asm + t_cg
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((l.split(":", 1)[1].strip() for l in f if l.startswith("model name")), "")
    except OSError:
        pass
    return {"value": round(1.0 / total, 4), "unit": "Newton solves/s", "cores": threads,
            "kind": "port", "host_cores_available": len(os.sched_getaffinity(0)),
            "cpu_model": model, "omp_num_threads": os.environ.get("OMP_NUM_THREADS"),
            "threads_note": "numpy/BLAS limited to 1 thread (threadpoolctl); scipy CG and "
                            "SpMV are single-threaded",
            "sample": (f"n={cs.n} cube: SNH assembly+projection on 1/{frac} of the {T} tets "
                       f"timed ({t_asm / frac:.2f}s) and scaled x{frac}; block-Jacobi CG "
                       f"(scipy, tol {tol:g}) timed in full ({t_cg:.2f}s, info={info}, "
                       f"rel.res {rel:.2e}) on the GPU-assembled matrix"),
            "seconds_per_solve": round(total, 3)}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; rerun with --gpus 1"}))
        sys.exit(2)
    from yasps_amd.scenes import cube_scene, build_yasps_scene
    from yasps_amd import device as dev
    from yasps_amd.hessian import vector
    torch = dev.torch()

    cs = cube_scene(args.n, seed=args.seed)
    s0, hnd = build_yasps_scene(cs, threshold_mode="relative_residual")
    # setup: static sparsity layout, rest data, first solve
    s0.minimizeEnergy(tolerance=args.tol)
    torch.cuda.synchronize()
    if args.kernels_only:
        spmv_roofline(s0.minimizer.activeHessian, args.spmv_reps)
        calibration_copy()
        return
    elapsed, lat, iters = timed_solves(s0, args.steps, args.warmup, args.tol)

    hs = s0.minimizer.activeHessian
    sol = s0.minimizer.linearSolver.solution
    g = hs.gradient.value
    r = hs.spmv(vector(hs.rows, sol.clone())).value - g
    true_rel = float(torch.linalg.norm(r) / torch.linalg.norm(g))

    roof = spmv_roofline(hs, args.spmv_reps)
    value = args.steps / elapsed
    forms = None
    if not args.no_forms:
        x64 = sol.clone()
        forms = {
            "mixed_jacobi": form_leg(cs, args, x64, value, "jacobian", precision="mixed"),
            "mas_fp64": form_leg(cs, args, x64, value, "mas", threshold_mode="relative_residual",
                                 max_levels=4, coarsest_level_weight=1.0),
            "mixed_mas": form_leg(cs, args, x64, value, "mas", precision="mixed", max_levels=4,
                                  coarsest_level_weight=1.0),
            "mixed_mas_bf16_coarse": form_leg(cs, args, x64, value, "mas", precision="mixed",
                                              max_levels=4, coarsest_level_weight=1.0,
                                              coarse_bank="bf16"),
        }
    out = {
        "metric": "Newton solves/s at 1e-6 rel. residual (plain FP64)",
        "value": round(value, 3),
        "unit": "Newton solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": (f"tet cube n={args.n} ({cs.tets.shape[0]} tets, {cs.n_dof} unknowns, "
                                f"{hs.static.num_blocks} 3x3 blocks), SNH E=1e5 nu=0.4 + inertia "
                                "dt=1e-2, bottom fixed, block-Jacobi PCG to ||g-Hx||/||g|| <= "
                                f"{args.tol:g}"),
                   "n": args.n, "solver": "jacobian", "precision": "fp64"},
        "p50_ms": round(1e3 * float(np.percentile(lat, 50)), 3),
        "p99_ms": round(1e3 * float(np.percentile(lat, 99)), 3),
        "pcg_iterations": int(np.median(iters)),
        "true_rel_residual": true_rel,
        "roofline": roof,
        "forms": forms,
        "serving": None if args.no_serving else serving_leg(args),
        "adjoint": None if args.no_adjoint else adjoint_leg(args),
        "dynamic_contacts": None if args.no_dynamic else dynamic_leg(args),
        "element_kernels": None if args.no_elements else element_leg(args),
        "contact_mixed_dof": None if args.no_contact else contact_leg(args),
    }
    if not args.no_cpu:
        from oracle import newton
        st = newton.fast_structure(cs)
        vals = dev.to_host(hs.blocks_flattened)
        cs.extras["gradient"] = dev.to_host(g)
        out["cpu_baseline"] = cpu_baseline(cs, vals, st, args.tol)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
