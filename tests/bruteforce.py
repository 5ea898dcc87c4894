"""O(N^2) numpy DBSCAN with the reference CPU path's exact rules.

Independent of any sort, cell table or union-find: it states each rule of
csrc/cpu/dbscan_grid.hpp directly over all point pairs, so it pins the C
oracle (and through it the HIP path) on small inputs.

  cells     kx = floor(f32(x) * inv_r), inv_r = f32(sqrt 2) / f32(eps)   (:153,162-163)
  window    |kx_i - kx_j| <= 2 and |ky_i - ky_j| <= 2                     (:239-252)
  neighbour same cell (no distance test), or in the window with
            fl(fl(ex*ex) + fl(ey*ey)) <= f32(eps)^2                        (:291-305)
  core      neighbour count (self included) >= min_samples               (:271-307)
  edges     two cores in one cell, or two cores in the window within eps  (:351-387)
  numbering components ranked by the smallest signed int64 cell key
            (u32(kx) << 32 | u32(ky)) among the cells holding their cores  (:176-178,392-398)
  border    argmin over window cores with d2 <= eps_sq and d2 < eps_sq + 1
            of (d2, off), off = (dkx + 2) * 5 + (dky + 2)                 (:419-445)

Finite inputs only (the cell rule uses numpy's float->int cast).
"""

from __future__ import annotations

import numpy as np
from scipy.sparse import csr_matrix
from scipy.sparse.csgraph import connected_components


def bruteforce_dbscan(X: np.ndarray, eps: float, min_samples: int) -> np.ndarray:
    X = np.ascontiguousarray(X, dtype=np.float32)
    n = X.shape[0]
    labels = np.full(n, -1, dtype=np.int32)
    if n == 0:
        return labels
    eps_f = np.float32(eps)
    inv_r = np.float32(np.sqrt(2.0)) / eps_f
    eps_sq = eps_f * eps_f
    best_init = np.float32(eps_sq + np.float32(1.0))
    kx = np.floor(X[:, 0] * inv_r).astype(np.int64)
    ky = np.floor(X[:, 1] * inv_r).astype(np.int64)

    ex = X[None, :, 0] - X[:, None, 0]
    ey = X[None, :, 1] - X[:, None, 1]
    d2 = ex * ex + ey * ey                      # float32, each op rounded
    assert d2.dtype == np.float32
    dkx = kx[None, :] - kx[:, None]
    dky = ky[None, :] - ky[:, None]
    same = (dkx == 0) & (dky == 0)
    window = (np.abs(dkx) <= 2) & (np.abs(dky) <= 2)
    within = window & (d2 <= eps_sq)
    neigh = same | within
    core = neigh.sum(axis=1) >= min_samples

    adj = (same | within) & core[:, None] & core[None, :]
    ncomp, comp = connected_components(csr_matrix(adj), directed=False)
    refkey = (kx << 32) | (ky & 0xFFFFFFFF)     # signed int64 view of the packed key
    core_idx = np.flatnonzero(core)
    if core_idx.size:
        comp_min = np.full(ncomp, np.iinfo(np.int64).max, dtype=np.int64)
        np.minimum.at(comp_min, comp[core_idx], refkey[core_idx])
        used = np.unique(comp[core_idx])
        order = used[np.argsort(comp_min[used], kind="stable")]
        rank = np.full(ncomp, -1, dtype=np.int64)
        rank[order] = np.arange(order.size)
        labels[core_idx] = rank[comp[core_idx]]

    off = (dkx + 2) * 5 + (dky + 2)
    for i in np.flatnonzero(~core):
        cand = np.flatnonzero(core & within[i] & (d2[i] < best_init))
        if cand.size:
            best = cand[np.lexsort((off[i, cand], d2[i, cand]))[0]]
            labels[i] = labels[best]
    return labels


def canonical(labels: np.ndarray) -> np.ndarray:
    """Relabel clusters by first appearance (noise stays -1)."""
    out = np.full_like(labels, -1)
    seen: dict[int, int] = {}
    for i, lab in enumerate(labels.tolist()):
        if lab == -1:
            continue
        if lab not in seen:
            seen[lab] = len(seen)
        out[i] = seen[lab]
    return out
