"""Continuous-batching server of Newton linear solves (device-resident, one GPU).

Requests (a tet-cube mesh of n^3 hexes, a Young's modulus and a noise seed) arrive over
time. The server admits an arrived request when a slot is free and its HBM footprint
fits the budget, keeps its mesh, Hessian and Krylov state on the device, and advances all
admitted requests together: every chunk of PCG passes is one sequence of batched
launches (``ysp_pcg3_batch_run``) in which each workgroup owns a chunk of one request's
system and every request reduces into its own scalar state. Requests that arrive while
others iterate join at the next chunk; finished ones leave and free their slot.

Meshes of one size share a pool of slots whose static sparsity structure (index
pipeline, block layout, full block-CSR, and with ``solver="mas"`` the MAS hierarchy and its
device tables) is built once; a request only reloads positions and material, evaluates
the elements and assembles (``hessian.compute``), then inverts its block-Jacobi blocks or
its MAS domains. By default a loader thread does that on a second stream, so the element
kernels of arriving requests overlap the batched PCG chunks of the admitted ones.

Preconditioners: block-Jacobi or MAS (``solver/mas/pcg.py:23-102``: every system of a
batch carries its own MAS, applied in batched restriction / local-solve / prolongation
launches). Precision: ``"fp64"`` runs the Krylov loop on the FP64 operator to the
tolerance; ``"mixed"`` runs it on FP32 block values (and FP32 / BF16 MAS banks, FP32
Jacobi blocks) with FP64 vectors and dots, and an FP64 iterative refinement around it: when
a system's inner solve ends, the true residual b - A x is formed with the FP64 values, and
if it is above the tolerance the system restarts on A d = r (correction) within the
running batch, until ||b - A x|| <= tol ||b||.
"""
from __future__ import annotations

import ctypes as C
import math
import queue
import threading
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

from . import _lib, device as dev
from .scenes import cube_scene, build_yasps_scene, lame_parameters

_MODES = {"reference": 0, "relative_residual": 1}


@dataclass
class SolveRequest:
    id: int
    n: int
    young: float
    seed: int
    arrival: float = 0.0          # seconds after the stream starts
    noise: float = 0.02
    # moving-vertex positions (flat, pinned host tensor) carried by the request; when None
    # the server derives them from the seed
    positions: Optional[object] = None
    # filled by the server
    admitted: Optional[float] = None
    finished: Optional[float] = None
    status: Optional[int] = None
    iterations: int = 0               # inner PCG iterations (all refinement steps)
    outer_iterations: int = 0         # FP64 refinement steps (mixed precision)
    rel_residual: Optional[float] = None   # final true FP64 residual (mixed precision)
    solution: Optional[np.ndarray] = None

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


class _Slot:
    """One device-resident problem of mesh size n (reused across requests)."""

    def __init__(self, n: int, precision: str, solver: str = "jacobian", coarse_bank: str = "fp32"):
        self.n = n
        self.precision = precision
        self.solver = solver
        cs = cube_scene(n)
        self.scene, self.h = build_yasps_scene(cs, threshold_mode="relative_residual")
        self.cs = cs
        self.scene.minimizer.computeNumericValue()       # builds the static structure
        hs = self.scene.minimizer.activeHessian
        if hs.full is None:
            raise RuntimeError("serving needs the 3x3 full block-CSR path")
        self.hs = hs
        m = hs.rows
        self.solution = dev.zeros(m, "float64")
        self.work = dev.zeros(5 * m, "float64")
        self.scalars = dev.zeros(12288, "float64")        # YSP_PCG3_SCALARS
        self.minv32 = dev.empty(hs.diagonal_blocks_inverse.numel(), "float32") \
            if precision == "mixed" else None
        f = hs.full
        staging = getattr(hs.static, "staging", None)
        self.footprint = int(
            8 * (hs.static.values.numel() + 9 * f.nnz + 5 * m + 2 * hs.diagonal_blocks.numel())
            + (4 * 9 * f.nnz if precision == "mixed" else 0)
            + 4 * (2 * f.nnz + m // 3)
            + (staging.numel() * staging.element_size() if staging is not None else 0))
        self.mas = None
        if solver == "mas":
            from .mas import masSolver
            # 4 levels, coarsest weight 1 (the configuration the benchmark legs run)
            self.mas = masSolver(max_levels=4, coarsest_level_weight=1.0,
                                 threshold_mode="relative_residual", coarse_bank=coarse_bank)
            shared = _MAS_HIERARCHIES.get(n)
            if shared is not None:                        # slots of one mesh size share it
                self.mas._hierarchy = shared
                self.mas._hierarchy_key = hs.full.key[:2]
            self.mas._ensure(hs, m)                       # hierarchy and device tables
            _MAS_HIERARCHIES[n] = self.mas._hierarchy
            self.footprint += _device_bytes(self.mas._device)
        elif solver != "jacobian":
            raise ValueError("solver must be 'jacobian' or 'mas'")
        if precision == "mixed":
            self.x = dev.zeros(m, "float64")             # the refined solution
            self.r = dev.zeros(m, "float64")             # FP64 residual b - A x
            self.footprint += 16 * m
        self.system = _lib.Pcg3System()
        self.pos_dev = dev.empty(3 * (cs.rest.shape[0] - cs.n_fixed), "float64")
        self.request: Optional[SolveRequest] = None
        self.outer = 0
        self.inner_iterations = 0
        self.rel_residual = None
        self.failed = None

    def load(self, req: SolveRequest, tolerance: float, max_iterations: int):
        """Positions and material of the request, element evaluation and assembly, block
        Jacobi inverse, and the first Krylov state."""
        if req.positions is not None:
            # asynchronous upload from the request's pinned payload into the slot's buffer
            self.pos_dev.copy_(req.positions, non_blocking=True)
            self.h["moving"]["position"].updateValue(self.pos_dev)
        else:
            self.h["moving"]["position"].updateValue(request_positions(self.cs, req))
        mu, lam = lame_parameters(req.young, self.cs.poisson)
        self.h["mu"].updateValue([mu])
        self.h["lam"].updateValue([lam])
        hs = self.scene.minimizer.computeNumericValue()
        lib = _lib.load()
        s = dev.stream()
        f = hs.full
        y = self.system
        mixed = self.precision == "mixed"
        minv = None
        if self.mas is not None:
            self.mas._device.update(f)                 # coarse operators, domain inverses
            y.mas = C.addressof(self.mas._device.args)
        else:
            counts = [w.numInstances for w in hs.wrt]
            sizes = [w.size for w in hs.wrt]
            _lib.check(lib.ysp_invert_diagonal_blocks(
                dev.ptr(hs.diagonal_blocks), dev.ptr(hs.diagonal_blocks_inverse),
                _lib.u32_array(hs.diagonal_blocks_start_cpu[:-1]), _lib.u32_array(counts),
                _lib.u32_array(sizes), len(hs.wrt), s), "ysp_invert_diagonal_blocks")
            minv = hs.diagonal_blocks_inverse
            if mixed:
                nd = hs.diagonal_blocks_inverse.numel()
                _lib.check(lib.ysp_cast_f32(dev.ptr(hs.diagonal_blocks_inverse), dev.ptr(self.minv32),
                                            nd, s), "ysp_cast_f32")
                minv = self.minv32
            y.mas = None
        values = hs.full_values_f32() if mixed else f.v64
        self._keep = (values, minv)
        y.row_ptr, y.col, y.values, y.nnz = dev.ptr(f.row_ptr), dev.ptr(f.col), dev.ptr(values), f.nnz
        y.diagonal_inverse = dev.ptr(minv)
        y.n = hs.rows
        y.work, y.cg_scalars = dev.ptr(self.work), dev.ptr(self.scalars)
        y.max_iterations = int(max_iterations)
        self.request = req
        self.tolerance = float(tolerance)
        self.max_iterations = int(max_iterations)
        self.outer = 0
        self.inner_iterations = 0
        self.rel_residual = None
        self.failed = None
        if mixed:
            # first inner solve on b itself (r = b, x = 0), to a slightly tighter tolerance
            self.x.zero_()
            self._start(hs.gradient.value, self.x, _SAFETY * self.tolerance)
        else:
            self._start(hs.gradient.value, self.solution, self.tolerance)

    def _start(self, rhs, solution, tol_in):
        y = self.system
        y.rhs = dev.ptr(rhs)
        y.solution = dev.ptr(solution)
        y.max_iterations = max(1, self.max_iterations - self.inner_iterations)
        mixed = self.precision == "mixed"
        _lib.check(_lib.load().ysp_pcg3_batch_start(C.byref(y), float(tol_in), 1,
                                                    int(mixed and self.mas is None), dev.stream()),
                   "ysp_pcg3_batch_start")

    def finish_inner(self) -> bool:
        """The inner solve of this slot ended (status in self.system). Returns True when the
        request is done; otherwise a correction solve has been started (mixed precision)."""
        y = self.system
        self.inner_iterations += max(0, int(y.iterations))
        if self.mas is not None and int(self.mas._device.status.item()) & 1:
            self.failed = "MAS local inversion failed (singular or indefinite domain)"
            return True
        if self.precision != "mixed":
            return True             # status >= 0: ||b - A x|| <= tol ||b|| (FP64 recurrence)
        lib = _lib.load()
        hs = self.hs
        f = hs.full
        rr, bb = C.c_double(0.0), C.c_double(0.0)
        first = self.outer == 0
        # x += d (after a correction solve), r = b - A x with the FP64 values
        _lib.check(lib.ysp_bsr3_refine(dev.ptr(f.row_ptr), dev.ptr(f.col), dev.ptr(f.v64), 0, f.nnz,
                                       hs.rows, dev.ptr(hs.gradient.value), dev.ptr(self.x),
                                       None if first else dev.ptr(self.solution), dev.ptr(self.r),
                                       C.byref(rr), C.byref(bb), dev.stream()), "ysp_bsr3_refine")
        self.outer += 1
        rel = math.sqrt(rr.value / bb.value) if bb.value > 0 else 0.0
        self.rel_residual = rel
        if (rel <= self.tolerance or not math.isfinite(rel) or self.outer > _MAX_OUTER
                or self.inner_iterations >= self.max_iterations or (y.status < 0 and y.status != -4)):
            return True
        self._start(self.r, self.solution, min(0.5, max(1e-14, _SAFETY * self.tolerance / rel)))
        return False

    @property
    def result(self):
        """The solution of the request (refined x in mixed precision)."""
        return self.x if self.precision == "mixed" else self.solution


# MAS hierarchy of the cube mesh of each size (a function of the sparsity structure only)
_MAS_HIERARCHIES: Dict[int, object] = {}
_SAFETY = 0.7        # inner tolerance factor of the refinement (as refine.refine_solve)
_MAX_OUTER = 8


def _device_bytes(obj) -> int:
    """Bytes of the device tensors an object holds (attributes and lists of tensors)."""
    total = 0
    for v in vars(obj).values():
        for t in (v if isinstance(v, (list, tuple)) else [v]):
            if isinstance(t, (list, tuple)):
                items = t
            else:
                items = [t]
            for u in items:
                if hasattr(u, "is_cuda") and u.is_cuda:
                    total += int(u.numel() * u.element_size())
    return total


class SolveServer:
    """Admission by HBM footprint and slot count, batched PCG chunks, retirement."""

    def __init__(self, hbm_budget_bytes: int = 64 << 30, max_active: int = 32, passes: int = 8,
                 tolerance: float = 1e-6, max_iterations: int = 20000, precision: str = "fp64",
                 keep_solutions: bool = False, overlap_loads: bool = True,
                 solver: str = "jacobian", coarse_bank: str = "fp32"):
        if precision not in ("fp64", "mixed"):
            raise ValueError("precision must be 'fp64' or 'mixed'")
        if solver not in ("jacobian", "mas"):
            raise ValueError("solver must be 'jacobian' or 'mas'")
        if coarse_bank not in ("fp32", "bf16"):
            raise ValueError("coarse_bank must be 'fp32' or 'bf16'")
        self.solver = solver
        self.coarse_bank = coarse_bank
        self.overlap_loads = bool(overlap_loads)
        self._lock = threading.Lock()           # the slot pool, shared with the loader
        self.budget = int(hbm_budget_bytes)
        self.max_active = int(max_active)
        self.passes = int(passes)
        self.tolerance = float(tolerance)
        self.max_iterations = int(max_iterations)
        self.precision = precision
        self.keep_solutions = keep_solutions
        self.pool: Dict[int, List[_Slot]] = {}
        self.chunks = 0
        self.batch_sizes: List[int] = []

    # ------------------------------------------------------------ slots
    def _new_slot(self, n: int) -> _Slot:
        return _Slot(n, self.precision, self.solver, self.coarse_bank)

    def _in_use(self) -> int:
        return sum(s.footprint for slots in self.pool.values() for s in slots if s.request is not None)

    def warm(self, sizes):
        """Build one idle slot per mesh size (static structures) before serving."""
        for n in sorted(set(int(x) for x in sizes)):
            if not self.pool.get(n):
                self.pool.setdefault(n, []).append(self._new_slot(n))
        dev.torch().cuda.synchronize()

    def _acquire(self, n: int) -> Optional[_Slot]:
        for s in self.pool.get(n, []):
            if s.request is None:
                return s
        slot = self._new_slot(n)
        self.pool.setdefault(n, []).append(slot)
        return slot

    def _fits(self, n: int) -> bool:
        idle = [s for s in self.pool.get(n, []) if s.request is None]
        need = idle[0].footprint if idle else self._estimate(n)
        return self._in_use() + need <= self.budget

    def _estimate(self, n: int) -> int:
        known = [s.footprint for s in self.pool.get(n, [])]
        if known:
            return known[0]
        # ~1.5 KB per tet (values, staging, vectors): conservative until a slot exists
        return int(1536 * 6 * n ** 3)

    # ------------------------------------------------------------ loop
    def serve(self, requests: List[SolveRequest], clock=time.perf_counter) -> List[SolveRequest]:
        """Serve a request stream; returns the finished requests (arrival, admission and
        finish times, status, PCG iterations)."""
        if self.overlap_loads:
            return self._serve_overlapped(requests, clock)
        return self._serve_sequential(requests, clock)

    def _retire(self, active, arr, t_end, done):
        torch = dev.torch()
        keep = []
        for sl, y in zip(active, arr):
            sl.system = y
            if y.status != _lib.PCG3_RUNNING and sl.finish_inner():
                req = sl.request
                req.finished = t_end
                req.status = -6 if sl.failed else int(y.status)
                req.iterations = sl.inner_iterations
                req.outer_iterations = sl.outer
                req.rel_residual = sl.rel_residual
                if self.precision == "mixed" and sl.rel_residual is not None and \
                        not sl.rel_residual <= self.tolerance and req.status >= 0:
                    req.status = -4          # the refinement did not reach the tolerance
                if self.keep_solutions:
                    torch.cuda.synchronize()
                    req.solution = dev.to_host(sl.result).copy()
                with self._lock:
                    sl.request = None
                done.append(req)
            else:
                keep.append(sl)
        return keep

    def _run_chunk(self, active, s):
        lib = _lib.load()
        mixed = int(self.precision == "mixed")
        arr = (_lib.Pcg3System * len(active))(*[sl.system for sl in active])
        _lib.check(lib.ysp_pcg3_batch_run(arr, len(active), self.passes, mixed, mixed, s),
                   "ysp_pcg3_batch_run")
        self.chunks += 1
        self.batch_sizes.append(len(active))
        return arr

    def _serve_overlapped(self, requests, clock):
        """Admission hands requests to a loader thread that evaluates and assembles them on
        its own stream (element kernels of new requests overlap the batched PCG chunks);
        a loaded request joins the next chunk after the solver stream waits on its load
        event."""
        torch = dev.torch()
        pending = sorted(requests, key=lambda r: r.arrival)
        active: List[_Slot] = []
        done: List[SolveRequest] = []
        todo: "queue.Queue" = queue.Queue()
        ready = []
        reserved = {}                           # request id -> footprint held while loading
        load_stream = torch.cuda.Stream()
        main = torch.cuda.current_stream()

        def loader():
            with torch.cuda.stream(load_stream):
                while True:
                    req = todo.get()
                    if req is None:
                        return
                    try:
                        with self._lock:
                            slot = next((sl for sl in self.pool.get(req.n, [])
                                         if sl.request is None), None)
                            if slot is not None:
                                slot.request = req
                        if slot is None:              # a new slot, built outside the lock
                            slot = self._new_slot(req.n)
                            slot.request = req
                            with self._lock:
                                self.pool.setdefault(req.n, []).append(slot)
                        slot.load(req, self.tolerance, self.max_iterations)
                        ev = torch.cuda.Event()
                        ev.record(load_stream)
                        item = (slot, req, ev, None)
                    except BaseException as e:         # reported by the serving loop
                        item = (None, req, None, e)
                    with self._lock:
                        ready.append(item)

        th = threading.Thread(target=loader, name="yasps-loader", daemon=True)
        t0 = clock()
        th.start()
        s = main.cuda_stream
        try:
            while pending or active or reserved:
                now = clock() - t0
                refused = False                 # an arrived request exceeded the budget
                while pending and pending[0].arrival <= now and \
                        len(active) + len(reserved) < self.max_active:
                    n = pending[0].n
                    with self._lock:
                        idle = [sl for sl in self.pool.get(n, []) if sl.request is None]
                        need = idle[0].footprint if idle else self._estimate(n)
                        if self._in_use() + sum(reserved.values()) + need > self.budget:
                            refused = True
                            break
                    req = pending.pop(0)
                    req.admitted = clock() - t0
                    reserved[req.id] = need
                    todo.put(req)
                with self._lock:
                    got, ready[:] = list(ready), []
                for slot, req, ev, err in got:
                    if err is not None:
                        raise err
                    main.wait_event(ev)
                    del reserved[req.id]
                    active.append(slot)
                if not active:
                    if reserved:
                        time.sleep(2e-5)
                        continue
                    if refused:
                        raise ValueError(f"request {pending[0].id} (n={pending[0].n}) does not "
                                         "fit the HBM budget of an empty server")
                    if pending:
                        wait = pending[0].arrival - (clock() - t0)
                        if wait > 0:
                            time.sleep(wait)
                    continue
                arr = self._run_chunk(active, s)
                active = self._retire(active, arr, clock() - t0, done)
        finally:
            todo.put(None)
            th.join()
        return done

    def _serve_sequential(self, requests, clock):
        pending = sorted(requests, key=lambda r: r.arrival)
        active: List[_Slot] = []
        done: List[SolveRequest] = []
        t0 = clock()
        s = dev.stream()
        while pending or active:
            now = clock() - t0
            refused = False                     # an arrived request exceeded the budget
            while pending and pending[0].arrival <= now and len(active) < self.max_active:
                if not self._fits(pending[0].n):
                    refused = True
                    break
                req = pending.pop(0)
                slot = self._acquire(req.n)
                slot.load(req, self.tolerance, self.max_iterations)
                req.admitted = clock() - t0
                active.append(slot)
            if not active:
                if refused:
                    raise ValueError(f"request {pending[0].id} (n={pending[0].n}) does not "
                                     "fit the HBM budget of an empty server")
                if pending:
                    wait = pending[0].arrival - (clock() - t0)
                    if wait > 0:
                        time.sleep(wait)
                continue
            arr = self._run_chunk(active, s)
            active = self._retire(active, arr, clock() - t0, done)
        return done


def request_positions(cs, req: SolveRequest) -> np.ndarray:
    """Moving-vertex positions of a request: rest + N(0, (noise h)^2), seeded."""
    rng = np.random.default_rng(req.seed)
    nf = cs.n_fixed
    h = 1.0 / cs.n
    return cs.rest[nf:] + rng.normal(0.0, req.noise * h, size=(cs.rest.shape[0] - nf, 3))


def attach_payloads(requests: List[SolveRequest]) -> List[SolveRequest]:
    """Give every request its positions as a pinned host tensor (the data a client would
    send), so that the server's loop only uploads them."""
    torch = dev.torch()
    scenes = {}
    for r in requests:
        if r.n not in scenes:
            scenes[r.n] = cube_scene(r.n)
        pos = request_positions(scenes[r.n], r).reshape(-1)
        r.positions = torch.from_numpy(np.ascontiguousarray(pos)).pin_memory()
    return requests


def serving_stream(count: int = 512, n_min: int = 6, n_max: int = 48, young_min: float = 1e4,
                   young_max: float = 1e6, seed: int = 0) -> List[SolveRequest]:
    """Requests with n log-uniform in [n_min, n_max] and E log-uniform in [young_min,
    young_max] (arrival times set by ``poisson_arrivals``)."""
    rng = np.random.default_rng(seed)
    ns = np.exp(rng.uniform(math.log(n_min), math.log(n_max + 1), count)).astype(int)
    ns = np.clip(ns, n_min, n_max)
    young = np.exp(rng.uniform(math.log(young_min), math.log(young_max), count))
    return [SolveRequest(id=i, n=int(ns[i]), young=float(young[i]), seed=int(seed * 100000 + i))
            for i in range(count)]


def poisson_arrivals(requests: List[SolveRequest], rate: float, seed: int = 1):
    """Exponential inter-arrival times at ``rate`` requests per second."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, len(requests)))
    for r, a in zip(requests, t):
        r.arrival = float(a)
    return requests


def summary(done: List[SolveRequest]) -> dict:
    lat = np.array([r.latency for r in done])
    wait = np.array([r.admitted - r.arrival for r in done])
    span = max(r.finished for r in done) - min(r.arrival for r in done)
    return {"requests": len(done), "solves_per_s": len(done) / span,
            "p50_ms": 1e3 * float(np.percentile(lat, 50)),
            "p99_ms": 1e3 * float(np.percentile(lat, 99)),
            "wait_p50_ms": 1e3 * float(np.percentile(wait, 50)),
            "wait_p99_ms": 1e3 * float(np.percentile(wait, 99)),
            "converged": int(sum(1 for r in done if r.status is not None and r.status >= 0)),
            "mean_iterations": float(np.mean([r.iterations for r in done]))}
