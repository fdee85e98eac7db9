"""ysp_pcg3_solve_noalloc (work vectors and scalar state in the library's device array)
against ysp_pcg3_solve (caller-provided work buffers) through the C ABI: two calls in a
row, the second on a smaller system, each matching the allocating entry point's solution
and iteration count, and the size limit rejected before any device work."""
import ctypes as C

import numpy as np
import pytest

from yasps_amd.scenes import cube_scene, build_yasps_scene

pytestmark = pytest.mark.gpu


def _system(n, seed):
    """Assembled tet-cube operator (n^3 cells), its inverse diagonal blocks and the
    allocating solver's solution at a relative residual of 1e-10."""
    from yasps_amd import device as dev
    cs = cube_scene(n, seed=seed)
    s0, _ = build_yasps_scene(cs, threshold_mode="relative_residual")
    x_ref = dev.to_host(s0.minimizeEnergy(tolerance=1e-10)[0]).copy()
    stats = s0.minimizer.linearSolver.statistics
    assert stats["result"] >= 0
    return s0.minimizer.activeHessian, x_ref, stats["iterations"]


def _noalloc(h, tol):
    from yasps_amd import device as dev, _lib
    lib = _lib.load()
    f = h.full
    n = h.rows
    x = dev.zeros(n, "float64")
    x.fill_(np.nan)                 # every entry must be written by the solve
    st = _lib.PcgStats()
    rc = lib.ysp_pcg3_solve_noalloc(20000, tol, 1, dev.ptr(f.row_ptr), dev.ptr(f.col),
                                    dev.ptr(f.v64), 0, f.nnz, dev.ptr(h.diagonal_blocks_inverse),
                                    0, dev.ptr(h.gradient.value), n, dev.ptr(x), None, 1,
                                    C.byref(st), dev.stream())
    _lib.check(rc, "ysp_pcg3_solve_noalloc")
    return rc, st, dev.to_host(x).copy()


def test_noalloc_pcg_matches_allocating_entry_point_twice():
    from oracle import newton
    results = []
    for n, seed in ((10, 3), (5, 4)):            # 3,630 then 540 unknowns
        print(f"no-alloc PCG: cube n={n} seed={seed}", flush=True)
        h, x_ref, it_ref = _system(n, seed)
        rc, st, x = _noalloc(h, 1e-10)
        assert rc == st.iterations and st.status == rc
        assert abs(st.iterations - it_ref) <= 1
        assert np.all(np.isfinite(x))
        assert np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref) < 1e-8
        # the residual against the oracle's assembled operator
        cs = cube_scene(n, seed=seed)
        stc = newton.fast_structure(cs)
        asm = newton.evaluate(cs, stc)
        A = newton.bsr_matrix(stc, asm.values)
        rel = np.linalg.norm(asm.gradient - A @ x) / np.linalg.norm(asm.gradient)
        assert rel < 1.5e-10
        results.append(st.iterations)
    assert len(results) == 2


def test_noalloc_pcg_rejects_sizes_beyond_the_array():
    from yasps_amd import _lib
    lib = _lib.load()
    st = _lib.PcgStats()
    too_big = 393216 + 3
    rc = lib.ysp_pcg3_solve_noalloc(10, 1e-6, 1, None, None, None, 0, 0, None, 0, None, too_big,
                                    None, None, 1, C.byref(st), None)
    assert rc == -2
    assert b"YSP_PCG3_NOALLOC_MAX_N" in lib.ysp_last_error()
