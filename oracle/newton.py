"""One Newton linear solve of the tet-cube scene on the CPU.

Oracle module (test infrastructure only; also the timed CPU baseline).

The scene is the Newton step of ``examples/one_bunny/one_bunny.py:112-176``: Stable
Neo-Hookean on tets that JOIN the union [fixed, moving] vertices (projection method 2)
plus inertia on the moving vertices, minimisation target = moving positions, then
``minimizer.__solveLinearSystem`` (``minimizer/minimizer.pyx:363-378``) solves H x = g
from a zero initial guess.

Two structure builders are provided:
* ``general_structure`` runs the full integer restatement of ``oracle.indices`` (slow,
  small meshes);
* ``fast_structure`` builds the same lookups and layout with vectorised numpy for meshes
  without repeated vertices inside an element; tests check the two agree bit for bit.
"""
from __future__ import annotations

from dataclasses import dataclass
import time
import numpy as np

from . import energies, projection, indices as ix, assembly, pcg


def cube_programs(scene):
    """Index programs of the two terms (see ``oracle.indices``)."""
    union = ix.Union(counts=[scene.n_fixed, scene.n_moving],
                     members=[[ix.Leaf(3, None)], [ix.Leaf(3, 0)]])
    tet_prog = ix.Concat([ix.Join(scene.tets.astype(np.uint32), [union])])
    inertia_prog = ix.Concat([ix.Leaf(3, 0)])
    return tet_prog, inertia_prog


@dataclass
class Structure:
    layout: ix.BlockLayout
    tet_lookup_index: np.ndarray    # [T, 4, 4] value offset of pair (a, b), -1 if none
    tet_transpose: np.ndarray       # [T, 4, 4] bool: stored as H[b, a]
    vert_lookup: np.ndarray         # [Vm] value offset of the diagonal block
    n_dof: int
    n_fixed: int


def general_structure(scene):
    tet_prog, inertia_prog = cube_programs(scene)
    wrt_start = [0, scene.n_dof]
    t_tet = ix.term_indices(tet_prog, scene.tets.shape[0], wrt_start)
    t_in = ix.term_indices(inertia_prog, scene.n_moving, wrt_start)
    layout = ix.build_layout([t_tet, t_in])
    return t_tet, t_in, layout


def fast_structure(scene) -> Structure:
    tets = scene.tets
    nf = scene.n_fixed
    T = tets.shape[0]
    g = np.where(tets >= nf, 3 * (tets - nf), -1)                  # global DoF start
    a_idx, b_idx = np.triu_indices(4)
    ga, gb = g[:, a_idx], g[:, b_idx]
    valid = (ga >= 0) & (gb >= 0)
    row = np.minimum(ga, gb)
    col = np.maximum(ga, gb)
    nd = scene.n_dof
    vdiag = np.arange(scene.n_moving) * 3
    keys = np.concatenate([(row[valid] * nd + col[valid]), vdiag * nd + vdiag])
    ukeys = np.unique(keys)
    block_offsets = np.arange(len(ukeys)) * 9
    pos = np.stack([ukeys // nd, ukeys % nd], axis=1).reshape(-1).astype(np.uint32)
    layout = ix.BlockLayout(pos, [0, 9 * len(ukeys)], [len(ukeys)], [3, 3], [],
                            num_unique=len(ukeys), block_offsets=block_offsets.astype(np.uint32))
    look = np.full((T, 4, 4), -1, dtype=np.int64)
    trans = np.zeros((T, 4, 4), dtype=bool)
    lk = np.full(row.shape, -1, dtype=np.int64)
    lk[valid] = np.searchsorted(ukeys, row[valid] * nd + col[valid]) * 9
    look[:, a_idx, b_idx] = lk
    trans[:, a_idx, b_idx] = ga > gb
    vert_lookup = np.searchsorted(ukeys, vdiag * nd + vdiag) * 9
    return Structure(layout, look, trans, vert_lookup, nd, nf)


@dataclass
class Assembled:
    gradient: np.ndarray
    values: np.ndarray
    diagonal_blocks: np.ndarray
    energy: float


def evaluate(scene, st: Structure, method: int = 2, rest_data=None) -> Assembled:
    """Element gradients/Hessians (projected), scattered into the block storage."""
    tets = scene.tets
    if rest_data is None:
        rest_data = energies.snh_rest_data(scene.rest[tets])
    dinv, vol = rest_data
    x = scene.position[tets]
    g_t, H_t = energies.snh_gradient_hessian(x, dinv, vol, scene.mu, scene.lam, scene.dt)
    H_t = projection.project_batch(H_t, method)
    nf = scene.n_fixed
    nd = scene.n_dof
    grad = np.zeros(nd)
    gidx = np.where(tets >= nf, 3 * (tets - nf), -1)
    m = gidx >= 0
    rows = (gidx[:, :, None] + np.arange(3))[m]
    np.add.at(grad, rows.reshape(-1), g_t.reshape(-1, 4, 3)[m].reshape(-1))
    values = np.zeros(st.layout.total_size)
    Hb = H_t.reshape(-1, 4, 3, 4, 3).transpose(0, 1, 3, 2, 4)      # [T, a, b, 3, 3]
    sel = st.tet_lookup_index >= 0
    blk = Hb[sel]
    tr = st.tet_transpose[sel]
    blk = np.where(tr[:, None, None], np.transpose(blk, (0, 2, 1)), blk)
    dst = st.tet_lookup_index[sel][:, None] + np.arange(9)
    np.add.at(values, dst.reshape(-1), blk.reshape(-1))
    # inertia (Hessian m I is PSD, projection leaves it unchanged)
    mv = scene.mass[nf:]
    gi, Hi = energies.inertia_gradient_hessian(scene.position[nf:], scene.last_position[nf:],
                                               scene.velocity[nf:], mv, scene.dt)
    grad += gi.reshape(-1)
    dst = st.vert_lookup[:, None] + np.arange(9)
    np.add.at(values, dst.reshape(-1), Hi.reshape(-1))
    # diagonal blocks of the single wrt (moving positions): the (v, v) block values
    diag = values[(st.vert_lookup[:, None] + np.arange(9)).reshape(-1)].copy()
    e = float(energies.snh_energy(x, dinv, vol, scene.mu, scene.lam, scene.dt).sum()
              + energies.inertia_energy(scene.position[nf:], scene.last_position[nf:],
                                        scene.velocity[nf:], mv, scene.dt).sum())
    return Assembled(grad, values, diag, e)


def bsr_matrix(st: Structure, values):
    """scipy BSR (3x3) of the full symmetric matrix built from the upper storage."""
    import scipy.sparse as sp
    pos = st.layout.block_positions.astype(np.int64).reshape(-1, 2) // 3
    B = values.reshape(-1, 3, 3)
    off = pos[:, 0] != pos[:, 1]
    rows = np.concatenate([pos[:, 0], pos[off, 1]])
    cols = np.concatenate([pos[:, 1], pos[off, 0]])
    data = np.concatenate([B, np.transpose(B[off], (0, 2, 1))])
    nb = st.n_dof // 3
    coo = sp.coo_matrix((np.ones(len(rows)), (rows, cols)), shape=(nb, nb)).tocsr()
    # order blocks as CSR
    order = np.lexsort((cols, rows))
    indptr = np.zeros(nb + 1, dtype=np.int64)
    np.add.at(indptr, rows + 1, 1)
    indptr = np.cumsum(indptr)
    del coo
    return sp.bsr_matrix((data[order], cols[order], indptr), shape=(st.n_dof, st.n_dof))


def newton_solve(scene, st: Structure = None, tol: float = 1e-8, method: int = 2,
                 max_iterations: int = 20000, use_scipy: bool = True, rest_data=None):
    """One Newton linear solve: evaluate, assemble, block-Jacobi, PCG to ``tol``
    (true relative residual ||g - H x|| / ||g||). Returns a dict of results."""
    t0 = time.perf_counter()
    if st is None:
        st = fast_structure(scene)
    asm = evaluate(scene, st, method, rest_data)
    t1 = time.perf_counter()
    inv = projection.jacobi_block_inverse(asm.diagonal_blocks.reshape(-1, 3, 3))
    A = bsr_matrix(st, asm.values)
    t2 = time.perf_counter()
    b = asm.gradient
    M = lambda v: pcg.block_jacobi_apply(inv, v)
    if use_scipy:
        import scipy.sparse.linalg as spla
        op = spla.LinearOperator(A.shape, matvec=M, dtype=np.float64)
        iters = [0]

        def cb(_):
            iters[0] += 1
        x, info = spla.cg(A, b, rtol=tol, atol=0.0, maxiter=max_iterations, M=op, callback=cb)
        status, it = info, iters[0]
    else:
        status, x, it, _ = pcg.pcg(lambda v: A @ v, M, b, tol * tol, max_iterations)
    t3 = time.perf_counter()
    rel = pcg.relative_residual(lambda v: A @ v, x, b)
    return dict(x=x, gradient=b, values=asm.values, status=int(status), iterations=int(it),
                rel_residual=rel, energy=asm.energy, t_assemble=t1 - t0, t_setup=t2 - t1,
                t_solve=t3 - t2, t_total=t3 - t0, structure=st)
