"""Fill-reducing nested-dissection order for the face system matrix K.

Faces carry coordinates (centroids), so the separator tree is built by recursive
coordinate bisection of the adjacency graph of K: split a vertex set at the median
of its longer axis, take as separator the vertices of one half that touch the
other half, recurse on both remainders. Subsets of at most ``leaf_size`` vertices
become leaf blocks.

Every tree node with rows is one dense "block" of the factor. Its stage is its
height above the leaves (leaves = 0). In the returned order, rows are sorted by
stage, so

* a block's factor rows only reference rows of the same block and of its
  descendants, which all sit in lower stages;
* blocks of one stage are never ancestor and descendant, so they are solved in
  parallel.

This is what lets the GPU triangular solves synchronise once per stage rather than
once per dependency level of the elimination tree.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse


def _split(idx, xy, adj, mark):
    """Return (separator, left_rest, right) for vertex subset idx."""
    pts = xy[idx]
    ext = pts.max(axis=0) - pts.min(axis=0)
    axes = (0, 1) if ext[0] >= ext[1] else (1, 0)
    for ax in axes:
        c = pts[:, ax]
        med = np.median(c)
        left = c <= med
        if left.all():
            left = c < med
        if left.any() and not left.all():
            break
    else:
        left = np.zeros(len(idx), dtype=bool)
        left[: len(idx) // 2] = True
        ax = axes[0]
    L, R = idx[left], idx[~left]
    # boundary of L towards R
    mark[R] = True
    sub = adj[L]
    rows = np.repeat(np.arange(len(L)), np.diff(sub.indptr))
    touch = np.zeros(len(L), dtype=bool)
    if len(sub.indices):
        touch[rows[mark[sub.indices]]] = True
    mark[R] = False
    sep = L[touch]
    other = 1 - ax
    sep = sep[np.argsort(xy[sep, other], kind="stable")]
    return sep, L[~touch], R


def nested_dissection(adj, xy, leaf_size=16):
    """Order the vertices of a symmetric sparsity pattern.

    Parameters
    ----------
    adj : scipy sparse (n, n), symmetric pattern (diagonal ignored)
    xy : (n, 2) coordinates
    leaf_size : int, largest subset kept as one leaf block

    Returns
    -------
    perm : (n,) int array, perm[new] = old
    blocks : (nb, 3) int array of (row_start, row_end, stage) in the new order,
             sorted by stage
    """
    adj = scipy.sparse.csr_matrix(adj)
    n = adj.shape[0]
    xy = np.asarray(xy, dtype=np.double)
    mark = np.zeros(n, dtype=bool)
    # iterative post-order: nodes = (rows, children list)
    nodes_rows = []
    nodes_children = []
    stack = [(np.arange(n), -1, False)]
    parent_of = []
    while stack:
        idx, parent, _ = stack.pop()
        me = len(nodes_rows)
        parent_of.append(parent)
        nodes_children.append([])
        if parent >= 0:
            nodes_children[parent].append(me)
        if len(idx) <= leaf_size:
            order = np.lexsort((xy[idx, 0], xy[idx, 1])) if len(idx) else np.zeros(0, dtype=int)
            nodes_rows.append(idx[order])
            continue
        sep, a, b = _split(idx, xy, adj, mark)
        nodes_rows.append(sep)
        for part in (b, a):
            if len(part):
                stack.append((part, me, False))
    nn = len(nodes_rows)
    height = np.zeros(nn, dtype=np.int64)
    # children are created after parents, so a reverse sweep sees children first
    for v in range(nn - 1, -1, -1):
        ch = nodes_children[v]
        if ch:
            hmax = max(height[c] for c in ch)
            height[v] = hmax + 1 if len(nodes_rows[v]) else hmax
    # stage-major order; within a stage, creation (pre-)order
    keep = [v for v in range(nn) if len(nodes_rows[v])]
    keep.sort(key=lambda v: (height[v], v))
    perm = np.concatenate([nodes_rows[v] for v in keep]) if keep else np.zeros(0, dtype=np.int64)
    sizes = np.array([len(nodes_rows[v]) for v in keep], dtype=np.int64)
    ends = np.cumsum(sizes)
    blocks = np.stack([ends - sizes, ends, height[keep]], axis=1) if keep else np.zeros((0, 3), np.int64)
    assert len(perm) == n and len(np.unique(perm)) == n
    return perm.astype(np.int64), blocks
