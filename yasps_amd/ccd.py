"""Contact sets on the GPU with the interface of the reference's collision class
(``examples/ccd/ccd.py`` ``CCD``): a linear-BVH broad phase, proximity pairs by distance
type (``cd``), swept candidates (``ccd``) and the additive-CCD step bound
(``compute_largest_step_size``), over ``ysp_ccd_candidates`` / ``ysp_ccd_active_pairs`` /
``ysp_ccd_step_bound`` (``csrc/ccd.hip``).

As in the reference: distances and ``dhat`` are squared; a step moves x to x - t d;
``mesh_indices`` (one id per vertex) suppresses pairs whose four vertices share one
nonzero id; ``separated_counts`` are (pp, pe, pt, ee) and ``pp`` / ``pe`` / ``pt`` /
``ee`` are flat uint32 device arrays of 2 / 3 / 4 / 4 ids per pair (the order within a
case is not fixed); candidates are compact int32 pairs (-(p + 1), face) or (edge, edge).
Device arrays are torch tensors (numpy arrays are uploaded). The BVHs are rebuilt for every
query; buffers grow on demand up to ``max_cd_pairs`` / ``max_ccd_pairs`` and a query that
needs more raises ``OverflowError``.

Not carried over: ``vertices_rest`` of ``init_edges`` is accepted and unused (the
reference reads it only for a parallel-edge marker that its packed output drops);
vertices marked fixed (``btype``, never set by the reference's Python class) do not exist.
"""
from __future__ import annotations

import ctypes as C
import time
from typing import List, Optional, Sequence

import numpy as np

from . import _lib, device as dev

_CASE_WIDTHS = (2, 3, 4, 4)
_INT32_MAX = np.iinfo(np.int32).max
_UINT32_MAX = np.iinfo(np.uint32).max
_INITIAL_CAPACITY = 1 << 20


def _is_int(v) -> bool:
    return isinstance(v, (int, np.integer)) and not isinstance(v, (bool, np.bool_))


class CCD:
    """GPU collision detection (``examples/ccd/ccd.py:109-1027``)."""

    def __init__(self, num_vertices: int, all_vertices: int, max_cd_pairs: int = 10_000_000,
                 max_ccd_pairs: int = 100_000_000, mesh_indices: Optional[Sequence[int]] = None,
                 print_timings: bool = True):
        for name, value in (("num_vertices", num_vertices), ("all_vertices", all_vertices),
                            ("max_cd_pairs", max_cd_pairs), ("max_ccd_pairs", max_ccd_pairs)):
            if not _is_int(value):
                raise TypeError(f"{name} must be an integer")
        if not 0 <= num_vertices <= all_vertices:
            raise ValueError("num_vertices must be between 0 and all_vertices")
        if not 1 <= all_vertices <= _INT32_MAX:
            raise ValueError(f"all_vertices must be in [1, {_INT32_MAX}]")
        for name, value in (("max_cd_pairs", max_cd_pairs), ("max_ccd_pairs", max_ccd_pairs)):
            if not 0 < value <= _INT32_MAX:
                raise ValueError(f"{name} must be in [1, {_INT32_MAX}]")
        self._num_vertices = int(num_vertices)
        self._all_vertices = int(all_vertices)
        self._max_cd = int(max_cd_pairs)
        self._max_ccd = int(max_ccd_pairs)
        self._print = bool(print_timings)
        mesh = self._validate_mesh_indices(mesh_indices, all_vertices)
        self._mesh = dev.to_device(mesh, "uint32") if mesh.any() else None
        self._faces = None
        self._surface = None
        self._edges = None
        self._face_num = 0
        self._edge_num = 0
        self._cand = dev.empty(2 * min(self._max_ccd, 1024), "int32")
        self._cases = [dev.empty(w * min(self._max_cd, 1024), "uint32") for w in _CASE_WIDTHS]
        self._counts = (0, 0, 0, 0)
        self._candidate_count = 0
        self._closed = False
        self._invalidate_cache()

    # ------------------------------------------------------------ validation
    @staticmethod
    def _validate_mesh_indices(mesh_indices, all_vertices):
        if mesh_indices is None or len(mesh_indices) == 0:
            return np.zeros(all_vertices, dtype=np.uint32)
        values = np.asarray(mesh_indices)
        if values.ndim != 1 or values.size != all_vertices:
            raise ValueError("Length of mesh_indices must be equal to all_vertices")
        if values.dtype.kind not in "iu":
            raise TypeError("mesh_indices must contain integers")
        if np.any(values < 0) or np.any(values > _UINT32_MAX):
            raise ValueError("mesh_indices values must fit in uint32")
        return np.ascontiguousarray(values, dtype=np.uint32)

    def _vertices(self, name, v):
        t = v if hasattr(v, "is_cuda") else dev.to_device(np.asarray(v, dtype=np.float64), "float64")
        if str(t.dtype) != "torch.float64":
            raise TypeError(f"{name} must have dtype float64")
        if t.numel() != 3 * self._all_vertices:
            raise ValueError(f"{name} must contain {3 * self._all_vertices} values; got {t.numel()}")
        if not t.is_contiguous():
            raise ValueError(f"{name} must be C-contiguous")
        return t

    def _indices(self, name, a, expected):
        if hasattr(a, "is_cuda"):
            host = dev.to_host(a).astype(np.int64).reshape(-1)
        else:
            host = np.asarray(a).astype(np.int64).reshape(-1)
        if host.size != expected:
            raise ValueError(f"{name} must contain {expected} values; got {host.size}")
        if host.size and (host.min() < 0 or host.max() >= self._all_vertices):
            raise ValueError(f"{name} indices must be in [0, {self._all_vertices - 1}]; got "
                             f"[{host.min()}, {host.max()}]")
        return dev.to_device(host.astype(np.uint32), "uint32")

    def _ensure_open(self):
        if self._closed:
            raise RuntimeError("CCD instance is closed")

    @staticmethod
    def _validate_dhat(dhat):
        if not np.isfinite(dhat) or dhat < 0.0:
            raise ValueError("dhat must be finite and nonnegative")

    def _validate_sweep(self, vertices, directions, alpha, dhat):
        self._validate_dhat(dhat)
        if not np.isfinite(alpha) or not 0.0 < alpha <= 1.0:
            raise ValueError("alpha must be finite and in (0, 1]")
        return self._vertices("vertices", vertices), self._vertices("moving_directions", directions)

    # ------------------------------------------------------------ set-up
    def init_faces(self, vertices, faces, surface_vertices, face_num: int) -> None:
        self._ensure_open()
        if self._faces is not None:
            raise RuntimeError("Face BVH is already initialized")
        if not _is_int(face_num):
            raise TypeError("face_num must be an integer")
        if not 0 <= face_num <= _INT32_MAX:
            raise ValueError(f"face_num must be in [0, {_INT32_MAX}]")
        self._vertices("vertices", vertices)
        self._faces = self._indices("faces", faces, 3 * face_num)
        self._surface = self._indices("surface_vertices", surface_vertices,
                                      self._num_vertices if face_num else
                                      (surface_vertices.numel() if hasattr(surface_vertices, "numel")
                                       else len(surface_vertices)))
        self._face_num = int(face_num)

    def init_edges(self, vertices, vertices_rest, edges, edge_num: int) -> None:
        self._ensure_open()
        if self._edges is not None:
            raise RuntimeError("Edge BVH is already initialized")
        if not _is_int(edge_num):
            raise TypeError("edge_num must be an integer")
        if not 0 <= edge_num <= _INT32_MAX:
            raise ValueError(f"edge_num must be in [0, {_INT32_MAX}]")
        self._vertices("vertices", vertices)
        if edge_num:
            self._vertices("vertices_rest", vertices_rest)
        self._edges = self._indices("edges", edges, 2 * edge_num)
        self._edge_num = int(edge_num)

    def _require_faces(self):
        self._ensure_open()
        if self._faces is None:
            raise RuntimeError("init_faces must be called first")

    def _require_edges(self):
        self._ensure_open()
        if self._edges is None:
            raise RuntimeError("init_edges must be called first")

    def construct_faces(self, vertices) -> None:
        """The BVHs are built inside every query; this validates and drops the cache."""
        self._require_faces()
        self._vertices("vertices", vertices)
        self._invalidate_cache()

    def construct_edges(self, vertices) -> None:
        self._require_edges()
        self._vertices("vertices", vertices)
        self._invalidate_cache()

    def construct_full_ccd_faces(self, vertices, moving_directions, alpha: float) -> None:
        self._require_faces()
        self._validate_sweep(vertices, moving_directions, alpha, 0.0)
        self._invalidate_cache()

    def construct_full_ccd_edges(self, vertices, moving_directions, alpha: float) -> None:
        self._require_edges()
        self._validate_sweep(vertices, moving_directions, alpha, 0.0)
        self._invalidate_cache()

    # ------------------------------------------------------------ state
    def _invalidate_cache(self):
        self._cache_valid = False
        self._cached_faces = self._cached_edges = False
        self._cached_dhat = self._cached_alpha = 0.0
        self._cached_x = self._cached_d = None

    @property
    def collision_pairs(self):
        """The active pairs as full-vertex int32 rows [pairs, 4] (point-based rows carry
        -(p + 1) first and -1 in unused slots)."""
        torch = dev.torch()
        rows = []
        for c, (w, n) in enumerate(zip(_CASE_WIDTHS, self._counts)):
            if not n:
                continue
            a = self._cases[c][:w * n].view(torch.int32).reshape(n, w).clone()
            if c < 3:
                a[:, 0] = -a[:, 0] - 1
            if w < 4:
                a = torch.cat([a, torch.full((n, 4 - w), -1, dtype=torch.int32, device=a.device)], 1)
            rows.append(a)
        if not rows:
            return dev.empty(0, "int32").reshape(0, 4)
        return torch.cat(rows)

    @property
    def collision_pairs_ccd(self):
        """The swept candidates as full-vertex int32 rows [candidates, 4]."""
        self._ensure_open()
        if not self._cache_valid:
            raise RuntimeError("ccd() must create a candidate cache first")
        torch = dev.torch()
        n = self._candidate_count
        c = self._cand[:2 * n].reshape(n, 2).long()
        out = torch.empty((n, 4), dtype=torch.int32, device=c.device)
        pf = c[:, 0] < 0
        if self._faces is not None and bool(pf.any()):
            f = self._faces.view(torch.int32).reshape(-1, 3).long()
            out[pf, 0] = c[pf, 0].int()
            out[pf, 1:] = f[c[pf, 1]].int()
        if self._edges is not None and bool((~pf).any()):
            e = self._edges.view(torch.int32).reshape(-1, 2).long()
            out[~pf, :2] = e[c[~pf, 0]].int()
            out[~pf, 2:] = e[c[~pf, 1]].int()
        return out

    @property
    def ccd_candidates(self):
        """Compact int32 candidate pairs [candidate_count, 2] of the current sweep."""
        return self._cand[:2 * self._candidate_count].reshape(-1, 2)

    @property
    def candidate_count(self) -> int:
        return self._candidate_count

    @property
    def cp_num(self):
        """[total, pp, pe, pt, ee] of the last collision detection (device uint32)."""
        return dev.to_device(np.array([sum(self._counts), *self._counts], dtype=np.uint32), "uint32")

    @property
    def separated_counts(self) -> List[int]:
        return list(self._counts)

    @property
    def case_offsets(self):
        pp, pe, pt, _ = self._counts
        return (0, 2 * pp, 2 * pp + 3 * pe, 2 * pp + 3 * pe + 4 * pt)

    @property
    def case_storage(self):
        """pp, pe, pt and ee back to back (case_offsets), as the reference stores them."""
        torch = dev.torch()
        return torch.cat([self._cases[c][:w * n] for c, (w, n) in
                          enumerate(zip(_CASE_WIDTHS, self._counts))])

    def _case(self, c):
        return self._cases[c][:_CASE_WIDTHS[c] * self._counts[c]]

    @property
    def pp(self):
        return self._case(0)

    @property
    def pe(self):
        return self._case(1)

    @property
    def pt(self):
        return self._case(2)

    @property
    def ee(self):
        return self._case(3)

    # ------------------------------------------------------------ queries
    def _candidates(self, x, d, alpha, dhat, faces: bool, edges: bool) -> int:
        lib = _lib.load()
        nf = self._face_num if faces else 0
        ne = self._edge_num if edges else 0
        count = C.c_uint32(0)
        while True:
            cap = self._cand.numel() // 2
            _lib.check(lib.ysp_ccd_candidates(
                dev.ptr(x), dev.ptr(d), float(alpha), self._all_vertices, dev.ptr(self._faces), nf,
                dev.ptr(self._surface), self._surface.numel() if nf else 0, dev.ptr(self._edges), ne,
                dev.ptr(self._mesh), float(dhat), dev.ptr(self._cand), cap, C.byref(count),
                dev.stream()), "ysp_ccd_candidates")
            if count.value <= cap:
                return int(count.value)
            if count.value > self._max_ccd:
                self._invalidate_cache()
                raise OverflowError("CCD broad-phase candidate capacity exceeded: needed more "
                                    f"than {self._max_ccd} entries")
            self._cand = dev.empty(2 * min(self._max_ccd, max(count.value, _INITIAL_CAPACITY)), "int32")

    def _active(self, x, dhat, num_candidates):
        lib = _lib.load()
        counts = (C.c_uint32 * 4)()
        while True:
            cap = min(t.numel() // w for t, w in zip(self._cases, _CASE_WIDTHS))
            _lib.check(lib.ysp_ccd_active_pairs(
                dev.ptr(x), dev.ptr(self._faces), dev.ptr(self._edges), dev.ptr(self._cand),
                num_candidates, float(dhat), *[dev.ptr(t) for t in self._cases], cap, counts,
                dev.stream()), "ysp_ccd_active_pairs")
            found = tuple(int(v) for v in counts)
            if max(found) <= cap:
                self._counts = found
                return found
            if sum(found) > self._max_cd:
                raise OverflowError("Active collision-pair capacity exceeded: needed more than "
                                    f"{self._max_cd} entries")
            need = min(self._max_cd, max(max(found), _INITIAL_CAPACITY))
            self._cases = [dev.empty(w * need, "uint32") for w in _CASE_WIDTHS]

    def _run_fresh_cd(self, vertices, dhat, faces: bool, edges: bool):
        self._ensure_open()
        if faces:
            self._require_faces()
        if edges:
            self._require_edges()
        if not faces and not edges:
            raise ValueError("At least one primitive type must be included")
        x = self._vertices("vertices", vertices)
        self._validate_dhat(dhat)
        self._invalidate_cache()
        n = self._candidates(x, None, 0.0, dhat, faces, edges)
        self._candidate_count = n
        return self._active(x, dhat, n)

    def cd_faces(self, vertices, dhat: float) -> List[int]:
        t0 = time.perf_counter()
        out = self._run_fresh_cd(vertices, dhat, True, False)
        self._elapsed("Face collision detection", t0)
        return list(out)

    def cd_edges(self, vertices, dhat: float) -> List[int]:
        t0 = time.perf_counter()
        out = self._run_fresh_cd(vertices, dhat, False, True)
        self._elapsed("Edge collision detection", t0)
        return list(out)

    def cd(self, vertices, dhat: float) -> List[int]:
        """Proximity pairs below dhat at the given positions (fresh broad phase)."""
        t0 = time.perf_counter()
        out = self._run_fresh_cd(vertices, dhat, self._faces is not None, self._edges is not None)
        self._elapsed("Collision detection", t0)
        return list(out)

    def _run_ccd(self, vertices, dhat, directions, alpha, faces: bool, edges: bool) -> int:
        self._ensure_open()
        if faces:
            self._require_faces()
        if edges:
            self._require_edges()
        if not faces and not edges:
            raise ValueError("At least one primitive type must be included")
        x, d = self._validate_sweep(vertices, directions, alpha, dhat)
        self._invalidate_cache()
        self._counts = (0, 0, 0, 0)
        n = self._candidates(x, d, alpha, dhat, faces, edges)
        self._candidate_count = n
        self._cache_valid = True
        self._cached_faces, self._cached_edges = faces, edges
        self._cached_dhat, self._cached_alpha = float(dhat), float(alpha)
        self._cached_x, self._cached_d = x, d
        return n

    def ccd_faces(self, vertices, dhat: float, moving_directions, alpha: float) -> int:
        t0 = time.perf_counter()
        n = self._run_ccd(vertices, dhat, moving_directions, alpha, True, False)
        self._elapsed("Face continuous collision detection", t0)
        return n

    def ccd_edges(self, vertices, dhat: float, moving_directions, alpha: float) -> int:
        t0 = time.perf_counter()
        n = self._run_ccd(vertices, dhat, moving_directions, alpha, False, True)
        self._elapsed("Edge continuous collision detection", t0)
        return n

    def ccd(self, vertices, dhat: float, moving_directions, alpha: float) -> int:
        """Swept candidates of the motion x -> x - alpha d (boxes grown by sqrt(dhat))."""
        t0 = time.perf_counter()
        n = self._run_ccd(vertices, dhat, moving_directions, alpha, self._faces is not None,
                          self._edges is not None)
        self._elapsed("Continuous collision detection", t0)
        return n

    def cd_from_cached_ccd(self, vertices, dhat: float, alpha: float) -> List[int]:
        """Proximity pairs at trial positions on the cached sweep (x - t d, t <= alpha): every
        such pair below the cached dhat is among the swept candidates."""
        self._ensure_open()
        if not self._cache_valid:
            raise RuntimeError("ccd() must create a candidate cache before cached filtering")
        if (self._face_num > 0 and not self._cached_faces) or \
                (self._edge_num > 0 and not self._cached_edges):
            raise RuntimeError("Cached filtering requires a sweep over every initialized "
                               "primitive type; call ccd() before cached collision filtering")
        x = self._vertices("vertices", vertices)
        self._validate_dhat(dhat)
        if alpha is not None and (not np.isfinite(alpha) or alpha < 0.0 or alpha > self._cached_alpha):
            raise ValueError(f"Trial alpha {alpha} is outside cached sweep [0, {self._cached_alpha}]")
        if dhat > self._cached_dhat:
            raise ValueError(f"Trial dhat {dhat} exceeds cached dhat {self._cached_dhat}")
        t0 = time.perf_counter()
        out = self._active(x, dhat, self._candidate_count)
        self._elapsed("Cached collision filtering", t0)
        return list(out)

    def compute_largest_step_size(self, slackness: float, vertices, moving_directions) -> float:
        """Additive-CCD bound of the step along x - t d over the cached candidates, at most
        the swept alpha (``examples/ccd/ccd.py:917-951``)."""
        self._ensure_open()
        if not self._cache_valid:
            raise RuntimeError("ccd() must be called before compute_largest_step_size()")
        if not np.isfinite(slackness) or not 0.0 <= slackness < 1.0:
            raise ValueError("slackness must be finite and in [0, 1)")
        x = self._vertices("vertices", vertices)
        d = self._vertices("moving_directions", moving_directions)
        if x.data_ptr() != self._cached_x.data_ptr():
            raise ValueError("vertices must be the same GPU array used to build the CCD cache")
        if d.data_ptr() != self._cached_d.data_ptr():
            raise ValueError("moving_directions must be the same GPU array used to build the CCD cache")
        t0 = time.perf_counter()
        step = C.c_double(1.0)
        _lib.check(_lib.load().ysp_ccd_step_bound(
            dev.ptr(x), dev.ptr(d), dev.ptr(self._faces), dev.ptr(self._edges), dev.ptr(self._cand),
            self._candidate_count, float(slackness), None, C.byref(step), dev.stream()),
            "ysp_ccd_step_bound")
        self._elapsed("Computing largest step size", t0)
        return min(float(step.value), self._cached_alpha)

    def candidate_times(self, slackness: float):
        """Additive-CCD time of every cached candidate (device FP64), the terms of
        compute_largest_step_size's minimum."""
        if not self._cache_valid:
            raise RuntimeError("ccd() must be called first")
        times = dev.empty(max(1, self._candidate_count), "float64")
        step = C.c_double(1.0)
        _lib.check(_lib.load().ysp_ccd_step_bound(
            dev.ptr(self._cached_x), dev.ptr(self._cached_d), dev.ptr(self._faces),
            dev.ptr(self._edges), dev.ptr(self._cand), self._candidate_count, float(slackness),
            dev.ptr(times), C.byref(step), dev.stream()), "ysp_ccd_step_bound")
        return times[:self._candidate_count]

    def reset(self) -> None:
        self._ensure_open()
        self._counts = (0, 0, 0, 0)
        self._candidate_count = 0
        self._invalidate_cache()

    def _scene_size(self, ids, vertices):
        x = vertices.reshape(-1, 3)[ids.long()]
        if x.numel() == 0:
            return 0.0
        ext = x.max(dim=0).values - x.min(dim=0).values
        return float((ext * ext).sum().item())

    def get_scene_size_faces(self, vertices=None) -> float:
        """Squared diagonal of the faces' bounding box (``mlbvh.cu:1417-1425``)."""
        self._require_faces()
        if not self._face_num:
            return 0.0
        v = self._vertices("vertices", vertices) if vertices is not None else self._cached_x
        if v is None:
            raise RuntimeError("pass the vertices: no positions are cached")
        return self._scene_size(self._faces.view(dev.torch().int32), v)

    def get_scene_size_edges(self, vertices=None) -> float:
        self._require_edges()
        if not self._edge_num:
            return 0.0
        v = self._vertices("vertices", vertices) if vertices is not None else self._cached_x
        if v is None:
            raise RuntimeError("pass the vertices: no positions are cached")
        return self._scene_size(self._edges.view(dev.torch().int32), v)

    def _elapsed(self, label, t0):
        if self._print:
            print(f"{label} took {(time.perf_counter() - t0) * 1e3:.2f} ms")

    def close(self) -> None:
        if getattr(self, "_closed", True):
            return
        self._faces = self._edges = self._surface = self._mesh = None
        self._cand = None
        self._cases = [None] * 4
        self._invalidate_cache()
        self._counts = (0, 0, 0, 0)
        self._candidate_count = 0
        self._closed = True

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
