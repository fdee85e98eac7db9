"""Element scatter into upper block storage, dense reconstruction and block SpMV.

Oracle module (test infrastructure only).

``scatter_term`` restates the element epilogue of
``kernel/Hessian/hessianKernelNoProject.pyx:63-264``:
* gradient: every slot with index >= 2 adds its segment of the *uncompressed* local
  gradient (running offset over all slots, padding included) at index - 2 (``:63-88``);
* Hessian: slots are grouped by |perm| in first-occurrence order (``:93-124``); for each
  pair of groups i <= j whose first slot has index >= 2 the block is the sum over all
  slot pairs of the two groups (``:185-251``), stored at ``lookups[coordStart + k]`` with
  k counting valid pairs in that order. If the row group's global index is smaller the
  block is H[i-rows, j-cols] row-major (h = size_i), otherwise it is stored transposed
  (h = size_j). Diagonal pairs (i == j) also go to ``diagonal_blocks`` at
  ``diag_start[attr] + instance * size^2`` (``:187-200,253-258``).

``spmv`` restates ``matrixAndVector/matrix/matrix.pyx:21-86`` for symmetric storage:
y_r += B x_c and, for r != c, y_c += B^T x_r.
"""
from __future__ import annotations

import numpy as np

from .indices import TermIndices, BlockLayout


def diagonal_layout(wrt_sizes, wrt_counts):
    """(gradient segment starts, diagonal block starts) of ``hessian.pyx:110-127``."""
    gs = [0]
    ds = [0]
    for s, c in zip(wrt_sizes, wrt_counts):
        gs.append(gs[-1] + s * c)
        ds.append(ds[-1] + s * s * c)
    return gs, ds


def scatter_term(local_g, local_H, term: TermIndices, lookups, gradient, blocks,
                 diagonal_blocks, grad_starts, diag_starts):
    """Accumulate one term. ``local_H`` is the dense (symmetric) local Hessian
    [N, n, n] over the uncompressed local layout, ``local_g`` [N, n]."""
    idx_all = term.indices.astype(np.int64)
    size_all = term.sizes.astype(np.int64)
    perm_all = term.compression.permutations.astype(np.int64)
    cstart = term.compression.coordinates_outer.astype(np.int64)
    N, K = idx_all.shape
    num_attr = len(grad_starts) - 1
    for e in range(N):
        idx, size, perm = idx_all[e], size_all[e], perm_all[e]
        g = local_g[e]
        H = local_H[e]
        off = 0
        for i in range(K):
            if idx[i] >= 2:
                p = idx[i] - 2
                gradient[p:p + size[i]] += g[off:off + size[i]]
            off += size[i]
        if H is None:
            continue
        # column offset of each slot over the uncompressed layout
        col = np.concatenate([[0], np.cumsum(size)[:-1]])
        groups = []          # (placement, [slots])
        for i in range(K):
            if perm[i] == 0:
                continue
            p = abs(perm[i]) - 1
            for grp in groups:
                if grp[0] == p:
                    grp[1].append(i)
                    break
            else:
                groups.append((p, [i]))
        count = 0
        for gi in range(len(groups)):
            si = groups[gi][1]
            ii = idx[si[0]]
            if ii < 2:
                continue
            hi = size[si[0]]
            for gj in range(gi, len(groups)):
                sj = groups[gj][1]
                jj = idx[sj[0]]
                if jj < 2:
                    continue
                wj = size[sj[0]]
                blk = np.zeros((hi, wj))
                for a in si:
                    for b in sj:
                        blk += H[col[a]:col[a] + hi, col[b]:col[b] + wj]
                place = int(lookups[cstart[e] + count])
                if ii < jj:
                    blocks[place:place + hi * wj] += blk.reshape(-1)
                else:
                    blocks[place:place + hi * wj] += blk.T.reshape(-1)
                    if gi == gj:
                        seg = ii - 2
                        attr = 0
                        while attr < num_attr and seg >= grad_starts[attr + 1]:
                            attr += 1
                        inst = (seg - grad_starts[attr]) // hi
                        d0 = diag_starts[attr] + inst * hi * hi
                        diagonal_blocks[d0:d0 + hi * hi] += blk.T.reshape(-1)
                count += 1


def dense_from_blocks(layout: BlockLayout, values, n):
    """Symmetric dense matrix from upper block storage (mirrors off-diagonal blocks)."""
    dense = np.zeros((n, n))
    pos = np.asarray(layout.block_positions, dtype=np.int64).reshape(-1, 2)
    k = 0
    for cat, cnt in enumerate(layout.block_counts):
        h, w = layout.block_dimensions[2 * cat:2 * cat + 2]
        for b in range(cnt):
            r, c = pos[k]
            o = layout.blocks_start_indices[cat] + b * h * w
            blk = np.asarray(values[o:o + h * w]).reshape(h, w)
            dense[r:r + h, c:c + w] += blk
            if r != c:
                dense[c:c + w, r:r + h] += blk.T
            k += 1
    return dense


def spmv(layout: BlockLayout, values, x, transpose=False):
    """Symmetric-storage block SpMV (the transpose flag only matters for
    rectangular blocks; for a symmetric matrix both give H x)."""
    x = np.asarray(x, dtype=np.float64)
    y = np.zeros_like(x)
    pos = np.asarray(layout.block_positions, dtype=np.int64).reshape(-1, 2)
    k = 0
    for cat, cnt in enumerate(layout.block_counts):
        h, w = layout.block_dimensions[2 * cat:2 * cat + 2]
        o0 = layout.blocks_start_indices[cat]
        B = np.asarray(values[o0:o0 + cnt * h * w]).reshape(cnt, h, w)
        r = pos[k:k + cnt, 0]
        c = pos[k:k + cnt, 1]
        xc = x[c[:, None] + np.arange(w)]
        xr = x[r[:, None] + np.arange(h)]
        if not transpose:
            np.add.at(y, r[:, None] + np.arange(h), np.einsum("bij,bj->bi", B, xc))
            off = r != c
            np.add.at(y, c[off][:, None] + np.arange(w), np.einsum("bij,bi->bj", B[off], xr[off]))
        else:
            np.add.at(y, c[:, None] + np.arange(w), np.einsum("bij,bi->bj", B, xr))
            off = r != c
            np.add.at(y, r[off][:, None] + np.arange(h), np.einsum("bij,bj->bi", B[off], xc[off]))
        k += cnt
    return y
