"""Host preparation of one circuit for the GPU step kernel.

Builds, once per run (the reference's set-up, time_evolution.py:466-521):

* the face matrix K = A (L + diag(1/(Rv+Cv))) A^T (time_evolution.py:476, 505),
  with the same scipy expression as the reference;
* a nested-dissection order and its stage/block structure (ordering.py);
* the Cholesky factor L of the reordered K (native, libjja);
* the stage schedule consumed by the kernel: off-block rows of L and L^T and the
  packed inverses of the dense diagonal blocks;
* cycle-matrix index arrays and per-junction coefficients.

Everything here is host work that runs once; the per-step work is on the GPU.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import scipy.linalg
import scipy.sparse

from . import _native
from .ordering import nested_dissection

STENCILS = {
    3: ((1.0, -1.0, 0.0), (1.0, -2.0, 1.0)),
    4: ((1.5, -2.0, 0.5, 0.0), (2.0, -5.0, 4.0, -1.0)),
    5: ((11.0 / 6, -3.0, 1.5, -1.0 / 3, 0.0), (35.0 / 12, -26.0 / 3, 9.5, -14.0 / 3, 11.0 / 12)),
}


def cholesky(K):
    """Sparse Cholesky K = L L^T (lower, CSC) with the native up-looking factoriser."""
    K = scipy.sparse.csc_matrix(K)
    K.sum_duplicates()
    K.sort_indices()
    n = K.shape[0]
    colptr = K.indptr.astype(np.int64)
    rowidx = K.indices.astype(np.int32)
    val = K.data.astype(np.float64)
    lib = _native.lib()
    handle = C.c_void_p()
    nnz = C.c_int64()
    rc = lib.jja_cholesky(n, _native.ptr(colptr, C.c_int64), _native.ptr(rowidx, C.c_int32),
                          _native.ptr(val, C.c_double), C.byref(handle), C.byref(nnz))
    _native.check(rc, factor=True)
    try:
        Lp = np.empty(n + 1, dtype=np.int64)
        Li = np.empty(nnz.value, dtype=np.int32)
        Lx = np.empty(nnz.value, dtype=np.float64)
        _native.check(lib.jja_factor_get(handle, _native.ptr(Lp, C.c_int64),
                                         _native.ptr(Li, C.c_int32), _native.ptr(Lx, C.c_double)),
                      factor=True)
    finally:
        lib.jja_factor_free(handle)
    return scipy.sparse.csc_matrix((Lx, Li, Lp), shape=(n, n))


def face_system_matrix(circuit, dt):
    """K exactly as the reference builds it (time_evolution.py:470-476, 505)."""
    A = circuit.get_cycle_matrix()
    Rv = 1 / (dt * circuit._R()[:, None])
    Cv = circuit._C()[:, None] / (dt ** 2)
    Cnext = Cv + Rv
    return A @ (circuit._L() + scipy.sparse.diags(1.0 / Cnext[:, 0], 0)) @ A.T


def _pack_inverses(Ldense_blocks):
    """Column-major and row-major packed lower-triangular inverses of each block."""
    cm_parts, rm_parts, offs = [], [], []
    off = 0
    for Lb in Ldense_blocks:
        B = Lb.shape[0]
        inv = scipy.linalg.solve_triangular(Lb, np.eye(B), lower=True)
        rows, cols = np.tril_indices(B)
        # row-major packed: row c contains inv(c, 0..c)
        rm_parts.append(inv[rows, cols])
        # column-major packed: column c contains inv(c..B-1, c)
        cr, cc = np.triu_indices(B)            # (cc >= cr): transposed traversal
        cm_parts.append(inv[cc, cr])
        offs.append(off)
        off += B * (B + 1) // 2
    if not cm_parts:
        return np.zeros(0), np.zeros(0), np.zeros(0, dtype=np.int64)
    return np.concatenate(cm_parts), np.concatenate(rm_parts), np.array(offs, dtype=np.int64)


class SolvePlan:
    """Stage schedule of the forward/backward triangular solves of K."""

    def __init__(self, K, face_xy, leaf_size=16):
        K = scipy.sparse.csc_matrix(K)
        n = K.shape[0]
        self.n = n
        if n == 0:
            self.perm = np.zeros(0, dtype=np.int64)
            self.blocks = np.zeros((0, 3), dtype=np.int64)
            self._build_empty()
            return
        pattern = (abs(K) > 0).astype(np.int8)
        perm, blocks = nested_dissection(pattern, face_xy, leaf_size=leaf_size)
        self.perm = perm
        self.blocks = blocks
        Kp = K[perm][:, perm]
        L = cholesky(Kp)
        self.L = L
        self.nnz_L = L.nnz
        pos = np.empty(n, dtype=np.int64)
        pos[perm] = np.arange(n)
        self.face_pos = pos.astype(np.int32)
        stages = blocks[:, 2]
        n_stages = int(stages.max()) + 1
        # stage boundaries (blocks sorted by stage and contiguous)
        stage_ptr = np.zeros(n_stages + 1, dtype=np.int64)
        for s in range(n_stages):
            sel = blocks[stages == s]
            stage_ptr[s + 1] = sel[:, 1].max() if len(sel) else stage_ptr[s]
        for s in range(1, n_stages + 1):
            stage_ptr[s] = max(stage_ptr[s], stage_ptr[s - 1])
        self.stage_ptr = stage_ptr.astype(np.int32)
        self.n_stages = n_stages
        sizes = blocks[:, 1] - blocks[:, 0]
        row_block = np.repeat(blocks[:, 0], sizes).astype(np.int32)
        row_bsize = np.repeat(sizes, sizes).astype(np.int32)
        row_bend = row_block.astype(np.int64) + row_bsize
        self.row_block, self.row_bsize = row_block, row_bsize
        # off-block lower part, row-wise
        Lr = L.tocsr()
        Lr.sort_indices()
        rr = np.repeat(np.arange(n), np.diff(Lr.indptr))
        keep_lo = Lr.indices < row_block[rr]
        self.lo_ptr = np.concatenate(([0], np.cumsum(np.bincount(rr[keep_lo], minlength=n)))).astype(np.int64)
        self.lo_col = Lr.indices[keep_lo].astype(np.int32)
        self.lo_val = Lr.data[keep_lo].astype(np.float64)
        # off-block part of L^T row-wise = columns of L below the block
        Lc = L.tocsc()
        Lc.sort_indices()
        cc = np.repeat(np.arange(n), np.diff(Lc.indptr))
        keep_up = Lc.indices >= row_bend[cc]
        self.up_ptr = np.concatenate(([0], np.cumsum(np.bincount(cc[keep_up], minlength=n)))).astype(np.int64)
        self.up_col = Lc.indices[keep_up].astype(np.int32)
        self.up_val = Lc.data[keep_up].astype(np.float64)
        # dense diagonal blocks and their inverses
        dense = []
        for b0, b1, _ in blocks:
            dense.append(L[b0:b1, b0:b1].toarray())
        self.inv_cm, self.inv_rm, boffs = _pack_inverses(dense)
        self.row_boff = np.repeat(boffs, sizes).astype(np.int64)
        self.max_stage_rows = int(np.max(np.diff(stage_ptr)))

    def _build_empty(self):
        self.face_pos = np.zeros(0, dtype=np.int32)
        self.stage_ptr = np.zeros(1, dtype=np.int32)
        self.n_stages = 0
        self.row_block = self.row_bsize = np.zeros(0, dtype=np.int32)
        self.row_boff = np.zeros(0, dtype=np.int64)
        self.lo_ptr = self.up_ptr = np.zeros(1, dtype=np.int64)
        self.lo_col = self.up_col = np.zeros(0, dtype=np.int32)
        self.lo_val = self.up_val = np.zeros(0)
        self.inv_cm = self.inv_rm = np.zeros(0)
        self.max_stage_rows = 0
        self.nnz_L = 0

    # -------------------------------------------------------- host emulation
    def solve(self, b):
        """Solve K x = b with exactly the staged algorithm the kernel runs (host
        emulation, used by tests and by the adjoint's host checks)."""
        b = np.asarray(b, dtype=np.float64)
        squeeze = b.ndim == 1
        if squeeze:
            b = b[:, None]
        z = np.empty_like(b)
        z[self.face_pos] = b
        n = self.n
        bs = self.row_bsize
        for s in range(self.n_stages):
            r0, r1 = self.stage_ptr[s], self.stage_ptr[s + 1]
            t = z[r0:r1].copy()
            for r in range(r0, r1):
                e0, e1 = self.lo_ptr[r], self.lo_ptr[r + 1]
                t[r - r0] -= self.lo_val[e0:e1] @ z[self.lo_col[e0:e1]]
            for r in range(r0, r1):
                b0, B, a = self.row_block[r], bs[r], r - self.row_block[r]
                off = self.row_boff[r]
                c = np.arange(a + 1)
                idx = off + c * B - c * (c - 1) // 2 + (a - c)
                z[r] = self.inv_cm[idx] @ t[b0 + c - r0]
        for s in range(self.n_stages - 1, -1, -1):
            r0, r1 = self.stage_ptr[s], self.stage_ptr[s + 1]
            t = z[r0:r1].copy()
            for r in range(r0, r1):
                e0, e1 = self.up_ptr[r], self.up_ptr[r + 1]
                t[r - r0] -= self.up_val[e0:e1] @ z[self.up_col[e0:e1]]
            for r in range(r0, r1):
                b0, B, a = self.row_block[r], bs[r], r - self.row_block[r]
                off = self.row_boff[r]
                c = np.arange(a, B)
                idx = off + c * (c + 1) // 2 + a
                z[r] = self.inv_rm[idx] @ t[b0 + c - r0]
        x = z[self.face_pos]
        assert x.shape[0] == n
        return x[:, 0] if squeeze else x

    def factor_bytes(self):
        """Bytes of factor data one column streams per step (both passes)."""
        return int(12 * (len(self.lo_val) + len(self.up_val)) + 8 * (len(self.inv_cm) + len(self.inv_rm))
                   + 4 * 4 * self.n)


def cycle_codes(A, Nf, Nj):
    """Face-wise and junction-wise index arrays of the cycle matrix.

    Face rows list junctions ascending (the order scipy's CSC product visits
    them); duplicate entries of one junction in one face (a dangling junction
    walked both ways) cancel and are dropped.
    """
    A = scipy.sparse.csr_matrix(A, dtype=np.float64)
    A.sum_duplicates()
    A.eliminate_zeros()
    A.sort_indices()
    face_ptr = A.indptr.astype(np.int32)
    neg = (A.data < 0).astype(np.int32)
    face_code = (2 * A.indices.astype(np.int64) + neg).astype(np.int32)
    At = A.T.tocsr()
    At.sort_indices()
    counts = np.diff(At.indptr)
    if np.any(counts > 2):
        raise ValueError("a junction borders more than two faces; the graph is not planar")
    junc_code = -np.ones((Nj, 2), dtype=np.int32)
    rows = np.repeat(np.arange(Nj), counts)
    slot = np.arange(len(At.indices)) - np.repeat(At.indptr[:-1], counts)
    junc_code[rows, slot] = 2 * At.indices + (At.data < 0)
    return face_ptr, face_code, junc_code.ravel()


class CircuitPlan:
    """All host arrays of one circuit for ``jja_plan_create``."""

    def __init__(self, circuit, dt, stencil_width=3, leaf_size=16):
        if stencil_width not in STENCILS:
            raise ValueError(f"stencil width must be 3, 4 or 5 (equals {stencil_width})")
        self.Nj, self.Nf = circuit._Nj(), circuit._Nf()
        self.stencil_width = stencil_width
        self.dt = dt
        A = circuit.get_cycle_matrix()
        self.face_ptr, self.face_code, self.junc_code = cycle_codes(A, self.Nf, self.Nj)
        Rv = 1 / (dt * circuit._R()[:, None])
        Cv = circuit._C()[:, None] / (dt ** 2)
        S1, S2 = STENCILS[stencil_width]
        self.coef = np.ascontiguousarray(
            np.stack([(S1[k] * Rv + S2[k] * Cv)[:, 0] for k in range(stencil_width)]))
        self.Rv = np.ascontiguousarray(Rv[:, 0])
        self.Ic = np.ascontiguousarray(circuit._Ic(), dtype=np.float64)
        self.K = face_system_matrix(circuit, dt)
        if self.Nf:
            cx, cy = circuit.get_face_centroids()
            xy = np.stack([cx, cy], axis=1)
        else:
            xy = np.zeros((0, 2))
        self.solve_plan = SolvePlan(self.K, xy, leaf_size=leaf_size)

    def desc(self):
        """ctypes ProblemDesc pointing at this object's arrays (keep self alive)."""
        sp = self.solve_plan
        P = _native.ptr
        d = _native.ProblemDesc()
        d.n_junctions, d.n_faces = self.Nj, self.Nf
        d.stencil_width, d.n_stages = self.stencil_width, sp.n_stages
        d.face_ptr = P(self.face_ptr, C.c_int32)
        d.face_code = P(self.face_code, C.c_int32)
        d.junc_code = P(self.junc_code, C.c_int32)
        d.Ic = P(self.Ic, C.c_double)
        d.Rv = P(self.Rv, C.c_double)
        d.coef = P(self.coef, C.c_double)
        d.face_pos = P(sp.face_pos, C.c_int32)
        d.stage_ptr = P(sp.stage_ptr, C.c_int32)
        d.lo_ptr = P(sp.lo_ptr, C.c_int64)
        d.lo_col = P(sp.lo_col, C.c_int32)
        d.lo_val = P(sp.lo_val, C.c_double)
        d.up_ptr = P(sp.up_ptr, C.c_int64)
        d.up_col = P(sp.up_col, C.c_int32)
        d.up_val = P(sp.up_val, C.c_double)
        d.row_block = P(sp.row_block, C.c_int32)
        d.row_bsize = P(sp.row_bsize, C.c_int32)
        d.row_boff = P(sp.row_boff, C.c_int64)
        d.n_inv = len(sp.inv_cm)
        d.inv_cm = P(sp.inv_cm, C.c_double)
        d.inv_rm = P(sp.inv_rm, C.c_double)
        return d

    def lds_bytes(self):
        return 8 * (self.Nf + self.solve_plan.max_stage_rows)
