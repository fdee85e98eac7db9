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
    ap.add_argument("--cold-reps", type=int, default=20,
                    help="rotations over the operator copies in the HBM-regime product timing")
    ap.add_argument("--no-pmc", action="store_true",
                    help="skip the hardware-counter traffic passes (rocprofv3 --pmc)")
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
    ap.add_argument("--no-large", action="store_true",
                    help="skip the n=64 cube (operator beyond the Infinity Cache)")
    ap.add_argument("--contact-pairs", type=int, default=1_500_000)
    ap.add_argument("--contact-steps", type=int, default=3)
    ap.add_argument("--adjoint-n", type=int, default=32)
    ap.add_argument("--adjoint-fd", type=int, default=64)
    ap.add_argument("--serving-requests", type=int, default=512)
    ap.add_argument("--kernels-only", action="store_true",
                    help="assembly, the PCG product launches and a calibration copy of a "
                         "known size, no JSON line (for hardware-counter runs)")
    return ap.parse_args()


# the PCG product kernel on the mirrored FP64 copy, as rocprofv3 names it
PRODUCT_KERNEL = "pcg3_spmv_kernel<ysp::Mirrored<double>"


def measure_traffic(args):
    """HBM bytes per launch of pcg3_spmv_kernel<Mirrored<double>> (the PCG loop's product kernel) from
    two hardware-counter runs, FETCH_SIZE then WRITE_SIZE, of ``bench.py --kernels-only``
    under rocprofv3, each counter scaled by its reading for a 256 MiB device copy of known
    size. Runs as child processes before this process touches the GPU. Returns
    (bytes, details); bytes is None when the counters could not be collected."""
    import csv
    import glob
    import shutil
    import subprocess
    import tempfile
    rocprof = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not os.path.exists(rocprof):
        return None, {"error": "rocprofv3 not found"}
    tmp = tempfile.mkdtemp(prefix="ysp_pmc_", dir="/tmp")
    env = dict(os.environ, TMPDIR="/tmp")
    detail = {"command": ("rocprofv3 --pmc <COUNTER> --output-format csv -- python3 bench.py "
                          f"--kernels-only --n {args.n} --seed {args.seed} --spmv-reps 5 --cold-reps 0")}
    total = 0.0
    try:
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            d = os.path.join(tmp, counter)
            cmd = ["timeout", "-s", "KILL", "180", rocprof, "--pmc", counter, "--output-format",
                   "csv", "-d", d, "-o", "run", "--", sys.executable, os.path.abspath(__file__),
                   "--kernels-only", "--n", str(args.n), "--seed", str(args.seed),
                   "--spmv-reps", "5", "--cold-reps", "0"]
            t0 = time.perf_counter()
            r = subprocess.run(cmd, cwd="/tmp", env=env, capture_output=True, text=True)
            files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
            print(f"traffic: {counter} pass rc={r.returncode} in {time.perf_counter() - t0:.1f} s",
                  file=sys.stderr, flush=True)
            if not files:
                return None, {**detail, "error": f"{counter}: rocprofv3 exit {r.returncode}, "
                                                 f"no counter file: {r.stderr[-300:]}"}
            prod, copy = [], []
            for fn in files:
                with open(fn, newline="") as fh:
                    for row in csv.DictReader(fh):
                        if row.get("Counter_Name") != counter:
                            continue
                        name = row.get("Kernel_Name", "")
                        if PRODUCT_KERNEL in name:
                            prod.append(float(row["Counter_Value"]))
                        elif "copyBuffer" in name:
                            copy.append(float(row["Counter_Value"]))
            if not prod or not copy:
                return None, {**detail, "error": f"{counter}: kernels missing in the counter file"}
            # the calibration copies (256 MiB each way) are the largest copy dispatches;
            # smaller ones are uploads of the set-up
            copy = [c for c in copy if c >= 0.5 * max(copy)]
            scale = 262144.0 / float(np.median(copy))
            kib = float(np.median(prod))
            total += kib * scale * 1024.0
            detail[counter] = {"product_kib_per_dispatch": round(kib, 1),
                               "copy_kib_per_dispatch": round(float(np.median(copy)), 1),
                               "scale": round(scale, 4), "dispatches": len(prod)}
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    return int(round(total)), detail


def calibration_copy(mib=256):
    """A plain device copy of a known size (the reference point for the counters)."""
    from yasps_amd import device as dev
    torch = dev.torch()
    a = torch.ones(mib * 131072, dtype=torch.float64, device=dev.device())
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()


def product_roofline(hs, reps, traffic=None, traffic_detail=None, cold_reps=20, replicas=4):
    """Roofline of pcg3_spmv_kernel<Mirrored<double>>, the kernel that dominates the timed PCG loop
    (one launch per iteration: q = A z + beta q, p = z + beta p and the partial sums of
    p.q over the mirrored 3x3 block-CSR), launched through ysp_pcg3_product_pass with the
    loop's own grid and timed by HIP events on the stream it runs on.

    Algorithmic bytes per launch: the upper 3x3 blocks read once (72 B values + 8 B (row,
    col)) plus five vector passes (z read, p read and written, q read and written). Two
    regimes, each a run of back-to-back launches between one pair of events:
    * in the loop: launches on one operator, which (134 MB at n=48) then stays in the
      256 MB Infinity Cache;
    * HBM (the headline): launches rotating over ``replicas`` copies of the operator and of
      the vectors (4 x 155 MB at n=48), so that each launch finds its operands evicted by
      the launches on the other copies and streams them from HBM."""
    from yasps_amd import _lib, device as dev
    torch = dev.torch()
    lib = _lib.load()
    f = hs.full
    n = hs.rows
    stream = torch.cuda.current_stream()

    def make(copy):
        if not copy:
            return (f.row_ptr, f.col, f.v64, dev.zeros(5 * n, "float64"), dev.zeros(12288, "float64"))
        return (f.row_ptr.clone(), f.col.clone(), f.v64.clone(), dev.zeros(5 * n, "float64"),
                dev.zeros(12288, "float64"))           # 12288 = YSP_PCG3_SCALARS

    def launch(op, prepare=0):
        rp, col, v, work, scal = op
        _lib.check(lib.ysp_pcg3_product_pass(dev.ptr(rp), dev.ptr(col), dev.ptr(v), 0, f.nnz, n,
                                             dev.ptr(work), dev.ptr(scal), prepare,
                                             stream.cuda_stream), "ysp_pcg3_product_pass")

    def timed(seq):
        start = torch.cuda.Event(enable_timing=True)
        end = torch.cuda.Event(enable_timing=True)
        start.record(stream)
        for op in seq:
            launch(op)
        end.record(stream)
        end.synchronize()
        return start.elapsed_time(end) / len(seq)

    ops = [make(False)]
    launch(ops[0], 1)
    for _ in range(3):
        launch(ops[0])
    ms_warm = timed([ops[0]] * reps)
    ms_cold = ms_warm
    if cold_reps:
        ops += [make(True) for _ in range(replicas - 1)]
        for op in ops[1:]:
            launch(op, 1)
        seq = [ops[i % replicas] for i in range(cold_reps * replicas)]
        for op in ops:                                  # one untimed rotation
            launch(op)
        ms_cold = timed(seq)
    del ops
    nb = hs.static.num_blocks
    bytes_alg = nb * (9 * 8 + 8) + 5 * n * 8
    bytes_full = f.nnz * (9 * 8 + 4) + (n // 3 + 1) * 4 + 5 * n * 8
    gbs_cold = bytes_alg / (ms_cold * 1e-3) / 1e9
    gbs_warm = bytes_alg / (ms_warm * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(gbs_cold, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(gbs_cold / HBM_PEAK_GBS, 4), "traffic": traffic,
            "kernel": "pcg3_spmv_kernel<Mirrored<double>>",
            "regime": (f"HBM: back-to-back launches rotating over {replicas} copies of the "
                       "operator and vectors (each launch streams its operands from HBM)"),
            "avg_us": round(ms_cold * 1e3, 2), "launches": cold_reps * replicas if cold_reps else 0,
            "bytes_per_launch": int(bytes_alg), "full_bsr_bytes_per_launch": int(bytes_full),
            "full_bsr_GBs": round(bytes_full / (ms_cold * 1e-3) / 1e9, 1),
            "in_loop": {"avg_us": round(ms_warm * 1e3, 2), "launches": reps,
                        "achieved": round(gbs_warm, 1),
                        "frac": round(gbs_warm / HBM_PEAK_GBS, 4),
                        "full_bsr_GBs": round(bytes_full / (ms_warm * 1e-3) / 1e9, 1),
                        "note": ("back-to-back launches on one operator as in the PCG loop: "
                                 "the operator (134 MB at n=48) stays in the 256 MB Infinity "
                                 "Cache")},
            "traffic_detail": traffic_detail}


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
    (capacity: the better of two such passes), then Poisson arrivals at 80% of that
    capacity; p50/p99 latency from
    arrival to solution. The leg's own figures are the mixed FP32/FP64 PCG + MAS server
    (the serving configuration's solver); FP64 block-Jacobi is reported beside it, with
    the default server (a loader thread and stream overlap request loading with the
    batched PCG chunks) and the sequential one (loads between chunks: larger batches in
    bursts, longer waits)."""
    from yasps_amd.serving import (SolveServer, attach_payloads, serving_stream,
                                   poisson_arrivals, summary)
    base = attach_payloads(serving_stream(count=args.serving_requests, seed=args.seed))

    def fresh(arrivals=False):
        return [type(r)(id=r.id, n=r.n, young=r.young, seed=r.seed, positions=r.positions,
                        arrival=r.arrival if arrivals else 0.0) for r in base]

    def run(overlap, solver="jacobian", precision="fp64"):
        srv = SolveServer(max_active=32, passes=8, tolerance=args.tol, overlap_loads=overlap,
                          solver=solver, precision=precision)
        t0 = time.perf_counter()
        srv.warm([r.n for r in base])
        # the first closed-loop pass grows the slot pool to the concurrency it needs (static
        # structures of meshes that are in flight more than once); the second one measures
        srv.serve(fresh())
        warm_s = time.perf_counter() - t0
        # capacity: the better of two closed-loop passes (one pass varies by a quarter
        # between runs with the host's scheduling of the loader thread)
        cap = max((summary(srv.serve(fresh())) for _ in range(2)), key=lambda c: c["solves_per_s"])
        rate = 0.8 * cap["solves_per_s"]
        reqs = poisson_arrivals(fresh(), rate, seed=args.seed + 1)
        srv.batch_sizes.clear()
        st = summary(srv.serve(reqs))
        done = [r for r in reqs if r.status is not None]
        res = [r.rel_residual for r in done if r.rel_residual is not None]
        return {"solver": solver, "precision": precision,
                "capacity_solves_per_s": round(cap["solves_per_s"], 2),
                "arrival_rate_per_s": round(rate, 2), "value": round(st["solves_per_s"], 2),
                "p50_ms": round(st["p50_ms"], 3), "p99_ms": round(st["p99_ms"], 3),
                "admission_wait_p50_ms": round(st["wait_p50_ms"], 3),
                "admission_wait_p99_ms": round(st["wait_p99_ms"], 3),
                "converged": st["converged"], "mean_pcg_iterations": round(st["mean_iterations"], 1),
                "mean_batch": round(float(np.mean(srv.batch_sizes)), 2),
                "max_batch": int(max(srv.batch_sizes)), "warm_seconds": round(warm_s, 2),
                "slots": int(sum(len(v) for v in srv.pool.values())),
                "max_true_fp64_rel_residual": (float(max(res)) if res else None),
                "mean_refinement_steps": (round(float(np.mean([r.outer_iterations for r in done])), 2)
                                          if precision == "mixed" else None)}

    # config 4's solver: mixed FP32/FP64 PCG + MAS (FP32 values and banks in the batched
    # Krylov loop, FP64 refinement to the true residual); FP64 block-Jacobi beside it
    out = {"requests": len(base), **run(True, "mas", "mixed"),
           "solver_note": ("batched PCG + MAS (4 levels, 32-DoF domains), FP32 values and "
                           "MAS banks in the Krylov loop, FP64 vectors and dots, FP64 "
                           "iterative refinement to ||g-Hx||/||g|| <= %g" % args.tol)}
    out["fp64_jacobi"] = run(True)
    out["fp64_jacobi_sequential_loads"] = run(False)
    return out


def adjoint_leg(args):
    """Material inverse on an n^3 cube (bottom fixed, gravity, theta = log E per tet ~
    N(log 1e5, 0.3)): equilibrium, loss against the equilibrium of a second theta, adjoint
    gradient in FP64 and mixed precision, central differences (h = 1e-6) on random
    entries; and the gradients w.r.t. E and nu per tet through the generalised mixed
    Jacobian (diff2), checked against the log-E adjoint and central differences in nu."""
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
    # dL/dE and dL/dnu per tet through the generalised mixed Jacobian d g / d (E, nu)
    xbar = prob.loss_grad_x(x, target)
    prob.set_theta(theta)
    t5 = time.perf_counter()
    gE, gnu = prob.vjp_material(x, xbar)
    t6 = time.perf_counter()
    E = np.exp(theta)
    nu0 = np.full(prob.num_tets, prob.poisson)
    err_nu = []
    for i in idx[:8]:
        npl, nmi = nu0.copy(), nu0.copy()
        npl[i] += h
        nmi[i] -= h
        fd = (prob.loss(prob.equilibrium_material(E, npl, x0=x, tol=1e-13), target)
              - prob.loss(prob.equilibrium_material(E, nmi, x0=x, tol=1e-13), target)) / (2 * h)
        err_nu.append(abs(fd - gnu[i]))
    material = {"vjp_E_nu_s": round(t6 - t5, 4),
                "E_dLdE_vs_log_E_adjoint_rel": float(np.abs(E * gE - g).max() / scale),
                "fd_nu_entries": len(err_nu),
                "fd_nu_max_abs_err_rel_to_max_grad": float(max(err_nu) / np.abs(gnu).max())}
    return {"n": args.adjoint_n, "tets": prob.num_tets, "unknowns": 3 * prob.n_free,
            "material_gradients": material,
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
    MAS (32-DoF leaf domains, 4 levels, nodes typed by target: 9-DoF affine matrices,
    3-DoF translations and vertices; the leg's value)
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
    sets = side_leg(lambda _: contact_sets_timing(cs), None)
    return {"contact_sets": sets,
            "bodies": int(cs.n_bodies + cs.n_free_bodies), "affine_bodies": int(cs.n_bodies),
            "tets": int(cs.tets.shape[0]), "cloth_triangles": int(cs.triangles.shape[0]),
            "hinges": int(cs.hinges.shape[0]), "point_triangle_pairs": int(cs.pt.shape[0]),
            "edge_edge_pairs": int(cs.ee.shape[0]), "unknowns": int(hs.rows),
            "static_blocks": int(hs.static.num_blocks),
            "dynamic_blocks": int(hs.dynamic.num_blocks),
            "split_3x3_entries": int(hs.split.nnz) if hs.split is not None else None,
            "value": legs["mas"]["value"], "unit": "Newton solves/s", "solvers": legs,
            "scene_generation_s": round(scene_s, 2)}


def large_operator_leg(args, n=64, steps=5):
    """The FP64 block-Jacobi Newton solve on a cube whose mirrored block-CSR (285 MB at
    n=64) does not stay in the 256 MB Infinity Cache, with the solve loop reading the
    mirrored copy and reading the upper block storage through the entries' source words
    (the automatic choice for such an operator). Same iterations; ms per solve."""
    from yasps_amd import device as dev
    from yasps_amd.scenes import cube_scene, build_yasps_scene
    torch = dev.torch()
    cs = cube_scene(n, seed=args.seed)
    s0, _ = build_yasps_scene(cs, threshold_mode="relative_residual")
    s0.minimizeEnergy(tolerance=args.tol)
    hs = s0.minimizer.activeHessian
    out = {"n": n, "tets": int(cs.tets.shape[0]), "unknowns": int(hs.rows),
           "mirrored_entries": int(hs.full.nnz),
           "mirrored_value_MB": round(72 * hs.full.nnz / 1e6, 1)}
    for source in ("mirrored", "auto"):
        hs.value_source = source
        s0.minimizeEnergy(tolerance=args.tol)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            s0.minimizeEnergy(tolerance=args.tol)
        torch.cuda.synchronize()
        ms = 1e3 * (time.perf_counter() - t0) / steps
        st = s0.minimizer.linearSolver.statistics
        out[source] = {"ms_per_solve": round(ms, 3), "solves_per_s": round(1e3 / ms, 2),
                       "pcg_iterations": int(st["iterations"]),
                       "values_read": "upper storage" if hs.solve_values()[1] == 2 else "mirrored copy"}
    print(f"large operator n={n}: mirrored {out['mirrored']['ms_per_solve']} ms, "
          f"auto {out['auto']['ms_per_solve']} ms", file=sys.stderr, flush=True)
    return out


def side_leg(fn, args):
    """A leg beside the headline solve: its error is reported in its own JSON field (and
    its traceback on stderr) so that the headline line is still printed."""
    try:
        return fn(args)
    except Exception as e:  # noqa: BLE001
        import traceback
        traceback.print_exc()
        return {"error": f"{type(e).__name__}: {e}"}


def contact_sets_timing(cs, reps=5):
    """The collision module (yasps_amd.ccd, csrc/ccd.hip) on the contact scene's surfaces:
    the balls' boundary triangles and the cloth (each ball its own mesh id, the cloth
    another, so no pair lies within one body). Times a proximity detection at dhat
    (cd: BVHs, queries, distance types), a sweep of the balls falling by 0.01 (ccd) and its
    additive-CCD step bound, each a full synchronous call, median of reps."""
    from yasps_amd import device as dev
    from yasps_amd.ccd import CCD
    from yasps_amd.scenes import boundary_faces
    x = cs.positions()
    nb = cs.n_bodies + cs.n_free_bodies
    per = cs.tets.shape[0] // nb
    mesh = np.ones(x.shape[0], dtype=np.uint32)
    mesh[cs.tets.reshape(-1)] = 2 + np.repeat(np.arange(cs.tets.shape[0]) // per, 4)
    faces = np.concatenate([boundary_faces(cs.tets), cs.triangles])
    e = np.sort(np.concatenate([faces[:, [0, 1]], faces[:, [1, 2]], faces[:, [0, 2]]]), axis=1)
    edges = np.unique(e, axis=0)
    surface = np.unique(faces)
    # reach sqrt(dhat) ~ 0.03 spans several cloth cells, so the proximity sets hold tens of
    # millions of pairs: capacities up to the int32 limit (buffers grow only to what is found)
    c = CCD(len(surface), x.shape[0], max_cd_pairs=2_000_000_000, max_ccd_pairs=2_000_000_000,
            mesh_indices=mesh, print_timings=False)
    xd = dev.to_device(x.reshape(-1), "float64")
    c.init_faces(xd, faces.reshape(-1), surface, len(faces))
    c.init_edges(xd, xd, edges.reshape(-1), len(edges))
    d = np.zeros_like(x)
    d[mesh > 1, 1] = 0.01                        # x - t d: every ball falls by 0.01
    dd = dev.to_device(d.reshape(-1), "float64")
    t_cd, t_ccd, t_step = [], [], []
    for _ in range(reps + 1):
        t0 = time.perf_counter()
        counts = c.cd(xd, cs.dhat)
        t1 = time.perf_counter()
        n = c.ccd(xd, cs.dhat, dd, 1.0)
        t2 = time.perf_counter()
        step = c.compute_largest_step_size(0.8, xd, dd)
        t3 = time.perf_counter()
        t_cd.append(t1 - t0)
        t_ccd.append(t2 - t1)
        t_step.append(t3 - t2)
    med = lambda v: round(1e3 * float(np.median(v[1:])), 3)  # noqa: E731
    return {"faces": int(len(faces)), "edges": int(len(edges)), "surface_vertices": int(len(surface)),
            "proximity_pairs": {"pp": counts[0], "pe": counts[1], "pt": counts[2], "ee": counts[3]},
            "swept_candidates": int(n), "step_bound": float(step),
            "cd_ms": med(t_cd), "ccd_ms": med(t_ccd), "step_bound_ms": med(t_step),
            "note": "synchronous calls (BVH builds for faces and edges, queries, narrow phase)"}


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


class CpuPath:
    """The oracle's CPU path of one Newton linear solve of the tet cube: SNH gradients and
    Hessians with their PSD projection (method 2) on chunks of tets spread over `threads`
    threads (numpy releases the GIL in its array loops; every BLAS/LAPACK call is limited
    to one thread), inertia, scatter into the 3x3 upper block storage, the scipy BSR
    matrix of the full operator, then block-Jacobi CG (scipy, single-threaded) to
    ||g - Hx|| / ||g|| <= tol. The static layout (rest data, block lookups, BSR order) is
    built before timing, as on the GPU."""

    def __init__(self, cs, threads):
        from oracle import energies, newton
        self.cs = cs
        self.threads = threads
        st = newton.fast_structure(cs)
        self.st = st
        T = cs.tets.shape[0]
        self.rest = energies.snh_rest_data(cs.rest[cs.tets])
        nchunk = 4 * threads
        self.chunks = [slice(i * T // nchunk, (i + 1) * T // nchunk) for i in range(nchunk)]
        sel = st.tet_lookup_index >= 0                          # [T, 4, 4]
        per_tet = sel.sum(axis=(1, 2))
        self.blk_start = np.concatenate([[0], np.cumsum([per_tet[c].sum() for c in self.chunks])])
        self.sel = sel
        self.dst = (st.tet_lookup_index[sel][:, None] + np.arange(9)).reshape(-1)
        nf = cs.n_fixed
        self.gidx = np.where(cs.tets >= nf, 3 * (cs.tets - nf), -1)
        self.gmask = self.gidx >= 0
        self.grows = (self.gidx[:, :, None] + np.arange(3))[self.gmask].reshape(-1)
        # BSR of the full operator from the upper storage: fixed block order
        pos = st.layout.block_positions.astype(np.int64).reshape(-1, 2) // 3
        self.off = pos[:, 0] != pos[:, 1]
        rows = np.concatenate([pos[:, 0], pos[self.off, 1]])
        cols = np.concatenate([pos[:, 1], pos[self.off, 0]])
        self.order = np.lexsort((cols, rows))
        self.cols = cols[self.order]
        nb = st.n_dof // 3
        indptr = np.zeros(nb + 1, dtype=np.int64)
        np.add.at(indptr, rows + 1, 1)
        self.indptr = np.cumsum(indptr)
        self.H = np.empty((T, 12, 12))
        self.G = np.empty((T, 12))
        self.BLK = np.empty((int(self.blk_start[-1]), 9))

    def _chunk(self, i):
        from oracle import energies, projection
        cs, c = self.cs, self.chunks[i]
        g, H = energies.snh_gradient_hessian(cs.position[cs.tets[c]], self.rest[0][c],
                                             self.rest[1][c], cs.mu, cs.lam, cs.dt)
        H = projection.project_batch(H, 2)
        self.G[c] = g
        Hb = H.reshape(-1, 4, 3, 4, 3).transpose(0, 1, 3, 2, 4)    # [t, a, b, 3, 3]
        sel = self.sel[c]
        blk = Hb[sel]
        tr = self.st.tet_transpose[c][sel]
        blk = np.where(tr[:, None, None], np.transpose(blk, (0, 2, 1)), blk)
        self.BLK[self.blk_start[i]:self.blk_start[i + 1]] = blk.reshape(-1, 9)

    def solve(self, tol):
        from concurrent.futures import ThreadPoolExecutor
        from oracle import energies, projection
        import scipy.sparse as sp
        import scipy.sparse.linalg as spla
        cs, st = self.cs, self.st
        nf = cs.n_fixed
        t0 = time.perf_counter()
        with ThreadPoolExecutor(self.threads) as ex:
            list(ex.map(self._chunk, range(len(self.chunks))))
        t1 = time.perf_counter()
        grad = np.zeros(st.n_dof)
        np.add.at(grad, self.grows, self.G.reshape(-1, 4, 3)[self.gmask].reshape(-1))
        values = np.zeros(st.layout.total_size)
        np.add.at(values, self.dst, self.BLK.reshape(-1))
        gi, Hi = energies.inertia_gradient_hessian(cs.position[nf:], cs.last_position[nf:],
                                                   cs.velocity[nf:], cs.mass[nf:], cs.dt)
        grad += gi.reshape(-1)
        dsti = (st.vert_lookup[:, None] + np.arange(9)).reshape(-1)
        np.add.at(values, dsti, Hi.reshape(-1))
        B = values.reshape(-1, 3, 3)
        data = np.concatenate([B, np.transpose(B[self.off], (0, 2, 1))])[self.order]
        A = sp.bsr_matrix((data, self.cols, self.indptr), shape=(st.n_dof, st.n_dof))
        t2 = time.perf_counter()
        inv = projection.jacobi_block_inverse(values[dsti].reshape(-1, 3, 3))
        M = spla.LinearOperator(A.shape, matvec=lambda v: np.einsum(
            "bij,bj->bi", inv, v.reshape(-1, 3)).reshape(-1), dtype=np.float64)
        it = [0]

        def count(_):
            it[0] += 1
        x, info = spla.cg(A, grad, rtol=tol, atol=0.0, maxiter=20000, M=M, callback=count)
        t3 = time.perf_counter()
        rel = float(np.linalg.norm(grad - A @ x) / np.linalg.norm(grad))
        return {"seconds": t3 - t0, "eval_project_s": t1 - t0, "scatter_bsr_s": t2 - t1,
                "cg_s": t3 - t2, "cg_iterations": it[0], "info": int(info), "rel_residual": rel}


def cpu_baseline(cs, tol, threads):
    """One full Newton linear solve of the benchmark scene on the host (CpuPath), timed in
    full; threadpoolctl keeps every BLAS/LAPACK call on one thread so that the thread
    count is the one stated."""
    import contextlib
    try:
        from threadpoolctl import threadpool_limits
        limiter = threadpool_limits(limits=1)
    except Exception:  # threadpoolctl absent: numpy keeps its default BLAS threads
        limiter = contextlib.nullcontext()
    with limiter:
        path = CpuPath(cs, threads)
        r = path.solve(tol)
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((l.split(":", 1)[1].strip() for l in f if l.startswith("model name")), "")
    except OSError:
        pass
    return {"value": round(1.0 / r["seconds"], 4), "unit": "Newton solves/s", "cores": threads,
            "kind": "port", "host_cores_available": len(os.sched_getaffinity(0)),
            "cpu_model": model, "omp_num_threads": os.environ.get("OMP_NUM_THREADS"),
            "threads_note": (f"element evaluation and projection on {threads} threads (numpy, "
                             "one BLAS/LAPACK thread each); scatter, scipy BSR build and "
                             "scipy CG on one thread"),
            "sample": (f"one full Newton linear solve of the n={cs.n} cube ({cs.tets.shape[0]} "
                       f"tets, {cs.n_dof} unknowns): evaluation + projection "
                       f"{r['eval_project_s']:.2f} s, scatter + BSR {r['scatter_bsr_s']:.2f} s, "
                       f"block-Jacobi CG {r['cg_s']:.2f} s ({r['cg_iterations']} iterations, "
                       f"info={r['info']}, rel.res {r['rel_residual']:.2e})"),
            "seconds_per_solve": round(r["seconds"], 3)}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; rerun with --gpus 1"}))
        sys.exit(2)
    # hardware-counter passes run as child processes before this one touches the GPU
    traffic, traffic_detail = None, None
    if not args.kernels_only and not args.no_pmc:
        traffic, traffic_detail = measure_traffic(args)
    from yasps_amd.scenes import cube_scene, build_yasps_scene
    from yasps_amd import device as dev
    from yasps_amd.hessian import vector
    torch = dev.torch()

    cs = cube_scene(args.n, seed=args.seed)
    s0, hnd = build_yasps_scene(cs, threshold_mode="relative_residual")
    if args.kernels_only:
        # assembly only (the operator of the loop), then the product launches (back to back
        # on one operator, then rotating over its copies) and the calibration copy: the only
        # pcg3_spmv_kernel launches are the measured ones
        s0.minimizer.computeNumericValue()
        roof = product_roofline(s0.minimizer.activeHessian, args.spmv_reps,
                                cold_reps=args.cold_reps)
        print(json.dumps({"kernels_only": {k: roof[k] for k in ("avg_us", "launches", "in_loop")}}),
              flush=True)
        calibration_copy()
        return
    # setup: static sparsity layout, rest data, first solve
    s0.minimizeEnergy(tolerance=args.tol)
    torch.cuda.synchronize()
    elapsed, lat, iters = timed_solves(s0, args.steps, args.warmup, args.tol)

    hs = s0.minimizer.activeHessian
    sol = s0.minimizer.linearSolver.solution
    g = hs.gradient.value
    r = hs.spmv(vector(hs.rows, sol.clone())).value - g
    true_rel = float(torch.linalg.norm(r) / torch.linalg.norm(g))

    roof = product_roofline(hs, args.spmv_reps, traffic, traffic_detail, args.cold_reps)
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
        "serving": None if args.no_serving else side_leg(serving_leg, args),
        "adjoint": None if args.no_adjoint else side_leg(adjoint_leg, args),
        "dynamic_contacts": None if args.no_dynamic else side_leg(dynamic_leg, args),
        "element_kernels": None if args.no_elements else side_leg(element_leg, args),
        "contact_mixed_dof": None if args.no_contact else side_leg(contact_leg, args),
        "large_operator": None if args.no_large else side_leg(large_operator_leg, args),
    }
    if not args.no_cpu:
        threads = max(1, min(16, len(os.sched_getaffinity(0))))
        out["cpu_baseline"] = cpu_baseline(cs, args.tol, threads)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
