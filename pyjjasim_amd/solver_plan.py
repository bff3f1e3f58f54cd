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

import copy
import ctypes as C
import heapq

import numpy as np
import scipy.linalg
import scipy.sparse
import scipy.sparse.csgraph

from . import _native
from .lattice_solve import lattice_solve
from .ordering import nested_dissection

MAX_ITER = 8          # entries per lane and solve record (kernel register budget)
FACE_ITER = 4         # junction codes per thread and face record
LDS_BYTES = 163840    # dynamic LDS one workgroup may use on MI355X (gfx950)
WAVE = 64             # lanes of one wavefront: solve records are per wave
SOLVERS = ("auto", "lattice", "cholesky")

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

    def row_work(self):
        """Factor entries each row's records stream per solve (both passes)."""
        a = np.arange(self.n) - self.row_block.astype(np.int64)
        return np.diff(self.lo_ptr) + np.diff(self.up_ptr) + (a + 1) + (self.row_bsize - a)

    def subtrees(self, n_low):
        """Group label of every row of the lowest ``n_low`` stages: blocks of those
        stages joined by a factor entry between their rows share a label. Below
        stage n_low the solve reads and writes only rows of the row's own group
        (its lower entries point into the group; its upper entries point into the
        group or into higher stages), so groups can be solved independently."""
        R = int(self.stage_ptr[n_low])
        sizes = self.blocks[:, 1] - self.blocks[:, 0]
        nb = int(np.count_nonzero(self.blocks[:, 2] < n_low))
        bid = np.repeat(np.arange(len(self.blocks)), sizes)
        if not hasattr(self, "_block_pairs"):
            rr = np.repeat(np.arange(self.n), np.diff(self.lo_ptr))
            key = np.unique(bid[rr] * np.int64(len(self.blocks)) + bid[self.lo_col])
            self._block_pairs = np.stack([key // len(self.blocks), key % len(self.blocks)], axis=1)
        p = self._block_pairs
        p = p[(p[:, 0] < nb) & (p[:, 1] < nb)]
        g = scipy.sparse.coo_matrix((np.ones(len(p)), (p[:, 0], p[:, 1])), shape=(nb, nb))
        _, label = scipy.sparse.csgraph.connected_components(g, directed=False)
        return label[bid[:R]]

    def factor_bytes(self):
        """Bytes of factor data one column streams per step (both passes)."""
        return int(12 * (len(self.lo_val) + len(self.up_val)) + 8 * (len(self.inv_cm) + len(self.inv_rm))
                   + 4 * 4 * self.n)


def _pow2ceil(x):
    g = 1
    while g < x:
        g *= 2
    return g


def sliced(row_ptr, idx, val, BS, row_ids, target_iter=4, max_group=64, fixed_group=None):
    """Sliced layout of a set of rows (CSR) for workgroups of BS threads.

    A row of length n gets G lanes, the smallest power of two (<= max_group) with
    at most ``target_iter`` entries per lane; rows of equal G are taken longest
    first, BS/G per slice; lane l of a row takes its entries l, l+G, ...; entry `it`
    of thread tid sits at base + it*BS + tid. Row slots without a row hold -1.

    Returns dict with per-slice base, iters, group, rowbase, and the flat arrays
    rows (row id per slot), idx, val (padding: index 0, value 0).
    """
    row_ptr = np.asarray(row_ptr, dtype=np.int64)
    nrows = len(row_ptr) - 1
    lens = np.diff(row_ptr)
    # lanes per row depend on the row's own length only, so the summation order
    # of a row (and the bits of the result) does not depend on BS
    if fixed_group:
        gsel = np.full(nrows, fixed_group, dtype=np.int64)
    else:
        need = np.maximum(1, -(-lens // target_iter))
        gsel = np.minimum(max_group, 2 ** np.ceil(np.log2(need)).astype(np.int64))
    order = np.lexsort((-lens, -gsel))
    bases, iters, groups, rowbases, rows = [], [], [], [], []
    row_slice = np.zeros(nrows, dtype=np.int64)
    row_slot = np.zeros(nrows, dtype=np.int64)
    row_G = np.ones(nrows, dtype=np.int64)
    base = 0
    rb = 0
    i = 0
    while i < nrows:
        G = int(gsel[order[i]])
        R = BS // G
        take = order[i:i + R]
        take = take[gsel[take] == G]
        L0 = int(lens[take[0]])
        it = -(-L0 // G)
        k = len(bases)
        row_slice[take] = k
        row_slot[take] = np.arange(len(take))
        row_G[take] = G
        slot_rows = -np.ones(R, dtype=np.int64)
        slot_rows[:len(take)] = np.asarray(row_ids)[take]
        bases.append(base)
        iters.append(it)
        groups.append(G)
        rowbases.append(rb)
        rows.append(slot_rows)
        base += it * BS
        rb += R
        i += len(take)
    idx_out = np.zeros(base, dtype=np.int64)
    val_out = np.zeros(base, dtype=np.float64)
    nnz = int(row_ptr[-1] - row_ptr[0]) if nrows else 0
    if nnz:
        r = np.repeat(np.arange(nrows), lens)
        kk = np.arange(nnz) - np.repeat(row_ptr[:-1] - row_ptr[0], lens)
        G = row_G[r]
        pos = np.asarray(bases, dtype=np.int64)[row_slice[r]] + (kk // G) * BS + row_slot[r] * G + kk % G
        idx_out[pos] = idx
        val_out[pos] = val
    return {"base": np.asarray(bases, dtype=np.int64), "iters": np.asarray(iters, dtype=np.int64),
            "group": np.asarray(groups, dtype=np.int64), "rowbase": np.asarray(rowbases, dtype=np.int64),
            "rows": np.concatenate(rows) if rows else np.zeros(0, np.int64),
            "idx": idx_out, "val": val_out}


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
    """All host arrays of one circuit for ``jja_plan_create``, laid out for
    workgroups of ``block_size`` threads; with ``parts`` > 1 a column is worked
    by that many workgroups (the layouts are built for parts * block_size
    threads, and the face store is global).

    ``local_stages``: the lowest stages are solved wave-locally: their blocks
    fall into independent subtrees, each wave solves whole subtrees, and waves
    meet once per pass for those stages instead of twice per stage. "auto"
    chooses the count from a cost estimate for split columns and 0 (every stage
    in step) for one workgroup per column. Above them, a split column's stages
    with at least one subtree per part are part-local (``part_stages``): each
    workgroup solves its own subtrees between workgroup barriers; 0 turns both
    off. Results do not depend on either: a row's summation order depends only
    on the row."""

    def __init__(self, circuit, dt, stencil_width=3, leaf_size=16, block_size=1024, target_iter=8,
                 face_store="auto", parts=1, local_stages="auto", solver="auto"):
        if face_store not in ("auto", "lds", "global"):
            raise ValueError("face_store must be 'auto', 'lds' or 'global'")
        if solver not in SOLVERS:
            raise ValueError("solver must be 'auto', 'lattice' or 'cholesky'")
        if parts < 1 or parts & (parts - 1) or parts > 32:
            raise ValueError("parts must be a power of two from 1 to 32")
        if parts > 1 and face_store == "lds":
            raise ValueError("a column split over workgroups needs the global face store")
        if local_stages != "auto" and (int(local_stages) != local_stages or local_stages < 0):
            raise ValueError("local_stages must be 'auto' or a count >= 0")
        self.face_store = face_store
        self.parts = parts
        self.local_option = local_stages
        self.solver = solver
        self.column_pairs = False
        if stencil_width not in STENCILS:
            raise ValueError(f"stencil width must be 3, 4 or 5 (equals {stencil_width})")
        self.Nj, self.Nf = circuit._Nj(), circuit._Nf()
        self.stencil_width = stencil_width
        self.dt = dt
        self.block_size = block_size
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
            self.circuit_xy = (np.asarray(cx), np.asarray(cy))
            xy = np.stack([cx, cy], axis=1)
        else:
            self.circuit_xy = None
            xy = np.zeros((0, 2))
        self.solve_plan = SolvePlan(self.K, xy, leaf_size=leaf_size)
        self.lattice = lattice_solve(self.K, xy) if self.Nf and solver != "cholesky" else None
        self.target_iter = target_iter
        self._build_layouts(target_iter)

    def relayout(self, block_size, parts=None, face_store=None, local_stages=None, solver=None):
        """The same circuit, factor and schedule laid out for another workgroup
        size, part count, face store, wave-local stage count or solver (the
        factorisation depends on none of them)."""
        other = copy.copy(self)
        other.block_size = block_size
        other.column_pairs = False
        if solver is not None:
            if solver not in SOLVERS:
                raise ValueError("solver must be 'auto', 'lattice' or 'cholesky'")
            if solver != "cholesky" and self.lattice is None and self.Nf:
                other.lattice = lattice_solve(self.K, self._face_xy())
            other.solver = solver
        if local_stages is not None:
            if local_stages != "auto" and (int(local_stages) != local_stages or local_stages < 0):
                raise ValueError("local_stages must be 'auto' or a count >= 0")
            other.local_option = local_stages
        if face_store is not None:
            if face_store not in ("auto", "lds", "global"):
                raise ValueError("face_store must be 'auto', 'lds' or 'global'")
            other.face_store = face_store
        if parts is not None:
            other.parts = parts
        if other.parts > 1 and other.face_store == "lds":
            raise ValueError("a column split over workgroups needs the global face store")
        other._build_layouts(self.target_iter)
        return other

    def paired(self):
        """This split plan for columns worked two at a time: workgroups of half
        the size, twice the parts, so the same threads per column and the same
        record programs (nothing is rebuilt)."""
        if self.parts < 2 or self.block_size not in (512, 1024) or 2 * self.parts > 32:
            raise ValueError("column pairs need a split plan (parts 2..16) with workgroups of 512 or 1024 threads")
        if self.uses_lattice:
            raise ValueError("column pairs share factor records; the lattice solve has none (solver='cholesky')")
        other = copy.copy(self)
        other.block_size = self.block_size // 2
        other.parts = 2 * self.parts
        other.column_pairs = True
        return other

    @property
    def uses_lattice(self):
        """The face system is solved by the lattice solve (a rectangular lattice
        with one junction weight per direction, one workgroup per column)."""
        return getattr(self, "_lattice_layout", False)

    def _face_xy(self):
        if not self.Nf:
            return np.zeros((0, 2))
        cx, cy = self.circuit_xy
        return np.stack([cx, cy], axis=1)

    @property
    def column_threads(self):
        """Threads that work on one column: the layouts are built for this many."""
        return self.block_size * self.parts

    # What one column barrier between workgroups costs, in factor entries one
    # wave streams meanwhile (about 14 us on MI355X for SquareArray(512,512) at
    # 16 parts, where a wave streams about 0.7 entries/ns).
    COLUMN_BARRIER_ENTRIES = 6000

    def _subtree_waves(self, NW, n_low):
        """Subtrees of the lowest n_low stages spread over NW waves, heaviest
        first, NW at a time: each round gives its heaviest subtree to the
        least-loaded wave, the next to the next least-loaded, and so on.
        Returns (factor entries per wave, owner wave of every row of those
        stages)."""
        sp = self.solve_plan
        label = sp.subtrees(n_low)
        if not hasattr(sp, "_row_work"):
            sp._row_work = sp.row_work()
        gw = np.bincount(label, weights=sp._row_work[:len(label)])
        order = np.lexsort((np.arange(len(gw)), -gw))
        loads = np.zeros(NW)
        wave_of = np.zeros(len(gw), dtype=np.int64)
        for i in range(0, len(order), NW):
            chunk = order[i:i + NW]
            waves = np.argsort(loads, kind="stable")[:len(chunk)]
            wave_of[chunk] = waves
            loads[waves] += gw[chunk]
        return loads, wave_of[label]

    def _choose_local(self, NW, max_rows):
        """Wave-local stage count with the lowest estimated cost: the slowest
        wave's excess over an even share of the wave-local work, minus the
        column barriers saved (4 per stage, less the leaves' skipped phase,
        plus the two that close the runs). max_rows bounds the scratch rows.

        With one workgroup per column the count is 0: a workgroup barrier costs
        less than the partly filled records that per-wave chains add (on
        SquareArray(100,100) every count from 2 to 7 was slower, by 4-14%)."""
        if self.parts == 1:
            return 0
        sp = self.solve_plan
        B = self.COLUMN_BARRIER_ENTRIES
        best, best_cost = 0, 0.0
        for k in range(1, sp.n_stages + 1):
            if max_rows is not None and max(sp.max_stage_rows, int(sp.stage_ptr[k])) > max_rows:
                break
            load, _ = self._subtree_waves(NW, k)
            cost = (load.max() - load.sum() / NW) - B * (4 * k - 3)
            if cost < best_cost:
                best, best_cost = k, cost
        return best

    def _wave_owner(self, NW, n_low):
        """Owner wave of every row of the lowest n_low stages (None for none)."""
        return self._subtree_waves(NW, n_low)[1] if n_low else None

    @staticmethod
    def _spread(weights, n):
        """Items heaviest first, n at a time, onto the least-loaded bins."""
        order = np.lexsort((np.arange(len(weights)), -weights))
        loads = np.zeros(n)
        bin_of = np.zeros(len(weights), dtype=np.int64)
        for i in range(0, len(order), n):
            chunk = order[i:i + n]
            bins = np.argsort(loads, kind="stable")[:len(chunk)]
            bin_of[chunk] = bins
            loads[bins] += weights[chunk]
        return loads, bin_of

    def _choose_part_local(self, NW, n_wave):
        """Stage count below which whole subtrees go to one workgroup (part) of
        a split column, so that those stages' phases end in workgroup barriers
        instead of column barriers: the one with the lowest estimated cost (the
        slowest part's excess over an even share, per wave, minus 4 column
        barriers per stage beyond the wave-local ones). 0 for unsplit columns."""
        if self.parts == 1 or self.local_option == 0:
            return 0
        sp = self.solve_plan
        if not hasattr(sp, "_row_work"):
            sp._row_work = sp.row_work()
        per_part = NW // self.parts
        B = self.COLUMN_BARRIER_ENTRIES
        best, best_cost = 0, 0.0
        for k in range(n_wave + 1, sp.n_stages + 1):
            label = sp.subtrees(k)
            gw = np.bincount(label, weights=sp._row_work[:len(label)])
            if len(gw) < self.parts:
                break
            load, _ = self._spread(gw, self.parts)
            cost = (load.max() - load.sum() / self.parts) / per_part - B * 4 * (k - n_wave)
            if cost < best_cost:
                best, best_cost = k, cost
        return best

    def _owners(self, NW, n_wave, n_part):
        """Owner wave of every row of the lowest n_wave stages and owner part of
        every row of the lowest n_part stages (None where none). Subtrees at
        level n_part go to parts; within a part, its subtrees at level n_wave go
        to the part's waves, so a wave's rows lie inside its part's."""
        if not n_part:
            return self._wave_owner(NW, n_wave), None
        sp = self.solve_plan
        if not hasattr(sp, "_row_work"):
            sp._row_work = sp.row_work()
        per_part = NW // self.parts
        label2 = sp.subtrees(n_part)
        gw2 = np.bincount(label2, weights=sp._row_work[:len(label2)])
        part_of = self._spread(gw2, self.parts)[1][label2]
        if not n_wave:
            return None, part_of
        label1 = sp.subtrees(n_wave)
        R1 = len(label1)
        gw1 = np.bincount(label1, weights=sp._row_work[:R1])
        grp_part = np.zeros(len(gw1), dtype=np.int64)
        grp_part[label1] = part_of[:R1]
        wave_of_grp = np.zeros(len(gw1), dtype=np.int64)
        for q in range(self.parts):
            gs = np.nonzero(grp_part == q)[0]
            if len(gs):
                wave_of_grp[gs] = q * per_part + self._spread(gw1[gs], per_part)[1]
        return wave_of_grp[label1], part_of

    def _build_layouts(self, target_iter=8):
        BS, sp, Nf, Nj = self.column_threads, self.solve_plan, self.Nf, self.Nj
        self._lattice_layout = False
        if Nf == 0:
            self._build_faceless()
            return
        lattice_ok = self.lattice is not None and self.solver != "cholesky"
        if self.solver == "lattice" and not lattice_ok:
            raise ValueError("the lattice solve needs a rectangular lattice with one junction weight per "
                             "direction (no mutual inductance)")
        if lattice_ok:
            self._build_lattice_layout()
            return
        # the face vector and stage scratch of a column: in the workgroup's LDS
        # when they fit (16-bit indices), else in a global per-column array.
        # Wave-local stages keep every scratch row of those stages at once.
        NW = BS // WAVE
        fixed = None if self.local_option == "auto" else min(int(self.local_option), sp.n_stages)
        scratch = sp.max_stage_rows if fixed is None else max(sp.max_stage_rows, int(sp.stage_ptr[fixed]))
        lds_rows = min(LDS_BYTES // 8, 65536) - Nf
        fits = scratch <= lds_rows
        if self.face_store == "lds" and not fits:
            raise ValueError(f"circuit with {Nf} faces needs {8 * (Nf + scratch)} bytes of face store per "
                             f"column; more than one workgroup's LDS ({LDS_BYTES})")
        self.global_store = self.face_store == "global" or not fits or self.parts > 1
        self.index_bytes = 4 if self.global_store else 2
        if fixed is None:
            fixed = self._choose_local(NW, None if self.global_store else lds_rows)
        self.local_stages = fixed
        self.part_stages = self._choose_part_local(NW, fixed)
        self.scratch_rows = max(sp.max_stage_rows, int(sp.stage_ptr[max(fixed, self.part_stages)]))
        owner, part_owner = self._owners(NW, fixed, self.part_stages)
        pos = sp.face_pos.astype(np.int64)
        # junctions: face positions of the (up to two) faces, ascending face index
        jc = self.junc_code.reshape(Nj, 2).astype(np.int64)
        jpos = np.where(jc >= 0, 2 * pos[np.maximum(jc, 0) >> 1] + (jc & 1), -1)
        self.junc_pos_code = jpos.astype(np.int32).ravel()
        self.face_of_pos = np.asarray(sp.perm, dtype=np.int32) if Nf else np.zeros(0, np.int32)
        rows_all = []
        n_rows = 0

        def records(L, ent_offset, r0, kind_flags, barrier, max_iter=MAX_ITER):
            """Split slices into <= max_iter-entry records (one row may span several)."""
            nonlocal n_rows
            out = []
            for k in range(len(L["base"])):
                it, G = int(L["iters"][k]), int(L["group"][k])
                R = BS // G
                live = int(np.count_nonzero(L["rows"][L["rowbase"][k]:L["rowbase"][k] + R] >= 0))
                n_chunks = max(1, -(-it // max_iter))
                for c in range(n_chunks):
                    flags = G | kind_flags | (live << 16)
                    if c == n_chunks - 1:
                        flags |= _native.SLICE_FINISH
                    out.append((int(L["base"][k]) + ent_offset + c * max_iter * BS,
                                int(L["rowbase"][k]) + n_rows, r0,
                                min(max_iter, it - c * max_iter), flags))
            if barrier and out:
                b, rb, rr, it, fl = out[-1]
                out[-1] = (b, rb, rr, it, fl | _native.SLICE_BARRIER)
            rows_all.append(L["rows"])
            n_rows += len(L["rows"])
            return out

        # faces in circuit order, one thread each, each written to its solve
        # position; entries: junction codes (sign in bit 0). Neighbouring faces
        # share junctions and have nearby junction indices, so the lanes of a
        # wave gather u from few cache lines.
        fp = self.face_ptr.astype(np.int64)
        pos_of_face = np.empty(Nf, dtype=np.int64)
        pos_of_face[self.face_of_pos] = np.arange(Nf)
        F = sliced(fp, self.face_code.astype(np.int64), np.ones(int(fp[-1]) if Nf else 0), BS,
                   pos_of_face, fixed_group=1, target_iter=FACE_ITER)
        self.face_code_sliced = np.where(F["val"] != 0, F["idx"], -1).astype(np.int32)
        self.face_slices = records(F, 0, 0, 0, False, max_iter=FACE_ITER)
        # solve program: forward stages ascending, backward stages descending
        rb, bsz, boff = sp.row_block.astype(np.int64), sp.row_bsize.astype(np.int64), sp.row_boff
        per_stage = []
        n_loc = self.local_stages
        n_part = self.part_stages
        for s in range(sp.n_stages):
            r0, r1 = int(sp.stage_ptr[s]), int(sp.stage_ptr[s + 1])
            # scratch row of row r: t[r - t0]; wave-local stages keep all their
            # rows' scratch at once (t0 = 0), the others reuse it per stage
            t0 = 0 if s < max(n_loc, n_part) else r0
            rows = np.arange(r0, r1)
            a = rows - rb[rows]
            B = bsz[rows]
            parts = [(sp.lo_ptr[r0:r1 + 1] - sp.lo_ptr[r0], sp.lo_col[sp.lo_ptr[r0]:sp.lo_ptr[r1]],
                      sp.lo_val[sp.lo_ptr[r0]:sp.lo_ptr[r1]])]
            cnt = a + 1
            ptr = np.concatenate(([0], np.cumsum(cnt)))
            rr = np.repeat(np.arange(r1 - r0), cnt)
            c = np.arange(ptr[-1]) - np.repeat(ptr[:-1], cnt)
            ar = a[rr]
            parts.append((ptr, (rb[rows][rr] - t0) + c, sp.inv_rm[boff[rows][rr] + ar * (ar + 1) // 2 + c]))
            parts.append((sp.up_ptr[r0:r1 + 1] - sp.up_ptr[r0], sp.up_col[sp.up_ptr[r0]:sp.up_ptr[r1]],
                          sp.up_val[sp.up_ptr[r0]:sp.up_ptr[r1]]))
            cnt = B - a
            ptr = np.concatenate(([0], np.cumsum(cnt)))
            rr = np.repeat(np.arange(r1 - r0), cnt)
            c = np.arange(ptr[-1]) - np.repeat(ptr[:-1], cnt)
            ar, Br = a[rr], B[rr]
            parts.append((ptr, (rb[rows][rr] - t0) + ar + c,
                          sp.inv_cm[boff[rows][rr] + ar * Br - ar * (ar - 1) // 2 + c]))
            per_stage.append((t0, rows, parts))
        prog = []
        for s in range(sp.n_stages):
            prog += [(s, 0), (s, 1)]
        for s in range(sp.n_stages - 1, -1, -1):
            prog += [(s, 2), (s, 3)]
        phases = []
        self.face_copy_rows = 0
        for s, kind in prog:
            t0, rows, parts = per_stage[s]
            pptr, pidx, pval = parts[kind]
            if (s, kind) == (0, 0) and pptr[-1] == pptr[0]:
                # leaves have no off-block entries: their scratch is b itself,
                # written by the face phase
                self.face_copy_rows = len(rows)
                continue
            # LDS indices in doubles: z at its position, stage scratch t at Nf + row - t0
            lds_base = Nf if kind % 2 else 0
            if s < n_loc:
                # a wave-local run ends at the last forward stage (then part-local
                # stages, if any, follow within each workgroup) and at the end of
                # the solve
                mode, own = "wave", owner[rows]
                if (s, kind) == (n_loc - 1, 1):
                    end = "wg" if n_part > n_loc else "col"
                else:
                    end = "col" if (s, kind) == (0, 3) else None
            elif s < n_part:
                mode, own = "part", part_owner[rows]
                last = (s, kind) == (n_part - 1, 1) or ((s, kind) == (0, 3))
                end = "col" if last else "wg"
            else:
                mode, own, end = "col", None, "col"
            phases.append((t0, rows, pptr, lds_base + np.asarray(pidx, dtype=np.int64), pval,
                           _native.SLICE_BLOCK if kind % 2 else 0, own, mode, end))
        self._build_wave_program(phases, target_iter)
        self.face_slices = [(b, 0, rb, r0, n, fl) for (b, rb, r0, n, fl) in self.face_slices]
        # face rows follow the solve rows (record row bases were taken before)
        face_rows = rows_all[0]
        self.face_slices = [(b, ib, rb + len(self.slice_rows), r0, n, fl)
                            for (b, ib, rb, r0, n, fl) in self.face_slices]
        self.slice_rows = np.concatenate((self.slice_rows, face_rows)).astype(np.int32)
        self._face_rec = self._slice_array(self.face_slices)
        self._solve_rec = self._slice_array(self.solve_slices)
        self._wave_ptr = np.ascontiguousarray(self.wave_ptr, dtype=np.int32)

    def _build_lattice_layout(self):
        """Face records by lattice position (r*C + c) and no solve program: the
        kernel solves K by the lattice solve (lattice_solve.py)."""
        lat, Nf, Nj, BS = self.lattice, self.Nf, self.Nj, self.column_threads
        scratch = lat.scratch_rows(Nf, self.parts)
        fits = Nf + scratch <= min(LDS_BYTES // 8, 65536)
        if self.face_store == "lds" and not fits:
            raise ValueError(f"circuit with {Nf} faces needs {8 * (Nf + scratch)} bytes of face store per "
                             f"column; more than one workgroup's LDS ({LDS_BYTES})")
        self.global_store = self.face_store == "global" or not fits or self.parts > 1
        self.index_bytes = 4 if self.global_store else 2
        self.local_stages = self.part_stages = 0
        self.scratch_rows = lat.scratch_rows(Nf, self.parts, self.global_store)
        pos = lat.pos_of_face
        jc = self.junc_code.reshape(Nj, 2).astype(np.int64)
        jpos = np.where(jc >= 0, 2 * pos[np.maximum(jc, 0) >> 1] + (jc & 1), -1)
        self.junc_pos_code = jpos.astype(np.int32).ravel()
        self.face_of_pos = lat.face_of_pos.astype(np.int32)
        fp = self.face_ptr.astype(np.int64)
        F = sliced(fp, self.face_code.astype(np.int64), np.ones(int(fp[-1])), BS, pos, fixed_group=1,
                   target_iter=FACE_ITER)
        self.face_code_sliced = np.where(F["val"] != 0, F["idx"], -1).astype(np.int32)
        recs = []
        for k in range(len(F["base"])):
            it = int(F["iters"][k])
            live = int(np.count_nonzero(F["rows"][F["rowbase"][k]:F["rowbase"][k] + BS] >= 0))
            n_chunks = max(1, -(-it // FACE_ITER))
            for c in range(n_chunks):
                flags = 1 | (live << 16)
                if c == n_chunks - 1:
                    flags |= _native.SLICE_FINISH
                recs.append((int(F["base"][k]) + c * FACE_ITER * BS, 0, int(F["rowbase"][k]), 0,
                             min(FACE_ITER, it - c * FACE_ITER), flags))
        self.face_slices, self.solve_slices = recs, []
        self.slice_rows = F["rows"].astype(np.int32)
        self.ent_val = np.zeros(0)
        self.ent_off = np.zeros(0, np.uint32 if self.global_store else np.uint16)
        self.face_copy_rows = 0
        self.n_slices = self.n_solve_exec = 0
        self.wave_ptr = np.zeros(BS // WAVE + 1, dtype=np.int64)
        self._face_rec = self._slice_array(self.face_slices)
        self._solve_rec = self._slice_array([])
        self._wave_ptr = np.ascontiguousarray(self.wave_ptr, dtype=np.int32)
        self._lattice_layout = True

    def _build_faceless(self):
        """A circuit without faces (a tree): K is 0 x 0 and the step is
        theta = -x / c0; no face records, no solve program."""
        self.global_store = self.parts > 1
        self.index_bytes = 4 if self.global_store else 2
        self.junc_pos_code = -np.ones(2 * self.Nj, dtype=np.int32)
        self.face_of_pos = np.zeros(0, np.int32)
        self.face_code_sliced = np.zeros(0, np.int32)
        self.face_slices, self.solve_slices = [], []
        self.slice_rows = np.zeros(0, np.int32)
        self.ent_val = np.zeros(0)
        self.ent_off = np.zeros(0, np.uint32 if self.global_store else np.uint16)
        self.face_copy_rows = 0
        self.local_stages = self.part_stages = self.scratch_rows = 0
        self.n_slices = self.n_solve_exec = 0
        self.wave_ptr = np.zeros(self.column_threads // WAVE + 1, dtype=np.int64)
        self._face_rec = self._slice_array([])
        self._solve_rec = self._slice_array([])
        self._wave_ptr = np.ascontiguousarray(self.wave_ptr, dtype=np.int32)

    def _build_wave_program(self, phases, target_iter):
        """Per-wave solve programs.

        Each phase (rows of one stage and kind) is cut into chains of 64-lane
        records: rows of equal G (the smallest power of two <= 64 with at most
        ``target_iter`` entries per lane, from the row's own length), 64/G rows
        per chain, longest first for an LDS face store and in row order for a
        global one (neighbouring rows read neighbouring cache lines:
        SquareArray(512,512) 10% faster; in LDS it was 2% slower); a chain is as
        long as its longest row; lane j of a row takes entries j, j+G, ...
        and a record holds 8 of them per lane (a longer row continues in the next
        record of its chain). Chains go to the least-loaded wave (longest first);
        the last record of every wave in a phase carries BARRIER (a wave without
        work gets a no-op record), so all waves meet at every phase boundary.
        A phase with an owner wave per row (a wave-local stage) is cut into chains
        of one owner's rows, each chain goes to its owner, and no barrier follows
        it: the owner's later records see its writes because one wave's LDS and
        vector-memory accesses are performed in order. The last phase of a run
        of wave-local phases ends with a barrier in every wave. A phase with an
        owner part per row (a part-local stage of a split column) spreads each
        part's chains over that part's waves and ends with a workgroup barrier
        (WG_BARRIER); a column barrier (BARRIER) closes the run.
        Record data: value `it` of lane l at base + (it/2)*128 + 2l + it%2, LDS
        index `it` (in doubles) of lane l at ibase + 8l + it, row of lane l at
        rowbase + l/G. Block 0 of every array is zero (the no-op record); record
        r of the data arrays has base = ibase = 512 r.
        """
        NW = self.column_threads // WAVE
        vals = [np.zeros(512)]
        offs = [np.zeros(512, dtype=np.int64)]
        narrow = [np.zeros(1, dtype=bool)]     # per record: 16-bit index deltas
        rows_out = [-np.ones(WAVE, dtype=np.int64)]
        n_rec, nr = 1, WAVE
        noop = (0, 0, 0, 0, 0, 1)
        progs = [[] for _ in range(NW)]
        region = [None] * NW    # where each wave's current wave-local run began
        per_part = NW // self.parts
        for (r0, rows, ptr, idx, val, kflag, owner, mode, end) in phases:
            ptr = np.asarray(ptr, dtype=np.int64)
            rows = np.asarray(rows, dtype=np.int64)
            lens = np.diff(ptr)
            need = np.maximum(1, -(-lens // target_iter))
            Gsel = np.minimum(WAVE, 2 ** np.ceil(np.log2(need)).astype(np.int64))
            key_all = np.zeros(len(rows), dtype=np.int64) if owner is None else np.asarray(owner, np.int64)
            chains = []   # (g, rows, records, first record, rowbase, iterations, owner)
            for g in sorted(set(Gsel.tolist())):
                sel = np.nonzero(Gsel == g)[0]
                if self.global_store:
                    # by owner, in row order: neighbouring rows gather from few
                    # cache lines of the global face store
                    sel = sel[np.lexsort((sel, key_all[sel]))]
                else:
                    sel = sel[np.lexsort((-lens[sel], key_all[sel]))]   # by owner, longest first
                key = key_all[sel]
                R = WAVE // g
                starts = np.concatenate(([0], np.nonzero(np.diff(key))[0] + 1))
                gsize = np.diff(np.concatenate((starts, [len(sel)])))
                gid = np.repeat(np.arange(len(starts)), gsize)
                inpos = np.arange(len(sel)) - starts[gid]
                gch = -(-gsize // R)
                chain_of = (np.cumsum(gch) - gch)[gid] + inpos // R
                slot = inpos % R
                n_ch = int(gch.sum())
                head = slot == 0                           # the first row of each chain
                cmax = np.zeros(n_ch, dtype=np.int64)
                np.maximum.at(cmax, chain_of, lens[sel])
                n_it = -(-cmax // g)
                nch = np.maximum(1, -(-n_it // MAX_ITER))
                rec0 = np.cumsum(nch) - nch
                v = np.zeros((int(nch.sum()), MAX_ITER, WAVE))
                o = np.zeros(v.shape, dtype=np.int64)
                L = lens[sel]
                who = np.repeat(np.arange(len(sel)), L)
                e = np.arange(int(L.sum())) - np.repeat(np.cumsum(L) - L, L)
                src = np.repeat(ptr[sel], L) + e
                k = e // g
                rec = rec0[chain_of[who]] + k // MAX_ITER
                ln = slot[who] * g + e % g
                v[rec, k % MAX_ITER, ln] = val[src]
                # record layouts: value it of lane l at (it/2)*128 + 2l + it%2
                vals.append(v.reshape(-1, MAX_ITER // 2, 2, WAVE).transpose(0, 1, 3, 2).reshape(-1))
                if kflag & _native.SLICE_BLOCK and self.index_bytes == 4:
                    offs.append(self._block_index_words(rec, k % MAX_ITER, ln, idx[src], g, v.shape[0]))
                    narrow.append(np.zeros(v.shape[0], dtype=bool))
                else:
                    # index it of lane l at 8l + it
                    o[rec, k % MAX_ITER, ln] = idx[src]
                    if self.index_bytes == 4:
                        words, nar = self._narrow_index_words(o, rec, k % MAX_ITER, ln)
                        offs.append(words)
                        narrow.append(nar)
                    else:
                        offs.append(o.transpose(0, 2, 1).reshape(-1))
                        narrow.append(np.zeros(v.shape[0], dtype=bool))
                rws = -np.ones(n_ch * R, dtype=np.int64)
                rws[chain_of * R + slot] = rows[sel]
                rows_out.append(rws)
                nrow = np.bincount(chain_of, minlength=n_ch)
                cown = key[head]
                for c in range(n_ch):
                    chains.append((g, int(nrow[c]), int(nch[c]), n_rec + int(rec0[c]), nr + c * R,
                                   int(n_it[c]), int(cown[c])))
                n_rec += int(nch.sum())
                nr += n_ch * R
            order = sorted(range(len(chains)), key=lambda c: -chains[c][2])
            assign = [[] for _ in range(NW)]
            if mode == "wave":
                for c in order:
                    assign[chains[c][6]].append(c)
            else:
                # least-loaded wave first, the lowest index among equals: over all
                # waves, or over the waves of the chain's owner part
                groups = [range(NW)] if mode == "col" else \
                    [range(q * per_part, (q + 1) * per_part) for q in range(self.parts)]
                heaps = [[(0, w) for w in ws] for ws in groups]
                for c in order:
                    h = heaps[0 if mode == "col" else chains[c][6]]
                    load, w = heapq.heappop(h)
                    assign[w].append(c)
                    heapq.heappush(h, (load + chains[c][2], w))
            bar = _native.SLICE_BARRIER if end == "col" else _native.SLICE_WG_BARRIER
            nar_all = np.concatenate(narrow)
            for w in range(NW):
                recs = []
                for c in assign[w]:
                    g, nrow, nch, first, rowbase, n_it, _ = chains[c]
                    for ch in range(nch):
                        flags = g | kflag | (nrow << 16)
                        if ch == nch - 1:
                            flags |= _native.SLICE_FINISH
                        if nar_all[first + ch]:
                            flags |= _native.SLICE_NARROW
                        b = 512 * (first + ch)
                        recs.append((b, b, rowbase, r0, min(MAX_ITER, n_it - ch * MAX_ITER), flags))
                if mode == "wave":
                    # wave-local: the wave's records follow each other with no
                    # barrier until the run of wave-local phases ends
                    if region[w] is None:
                        region[w] = len(progs[w])
                    progs[w] += recs
                    if not end:
                        continue
                    if len(progs[w]) == region[w]:
                        progs[w].append(noop)
                    recs, progs[w] = [progs[w][-1]], progs[w][:-1]
                    region[w] = None
                elif not recs:
                    recs.append(noop)
                b_, ib_, rb_, r0_, it_, fl_ = recs[-1]
                recs[-1] = (b_, ib_, rb_, r0_, it_, fl_ | bar)
                progs[w] += recs
        assert all(r is None for r in region), "a wave-local run without its closing barrier"
        solve, wave_ptr, n_exec = [], [], 0
        for w in range(NW):
            recs = progs[w]
            recs = recs + [noop] * (-len(recs) % 3)   # the kernel runs records in threes
            wave_ptr.append(len(solve))
            n_exec += len(recs)
            solve += recs + [noop] * 4
        wave_ptr.append(len(solve))
        self.solve_slices = solve
        self.wave_ptr = np.asarray(wave_ptr, dtype=np.int64)
        self.n_solve_exec = n_exec
        self.ent_val = np.concatenate(vals).astype(np.float64)
        self.ent_off = np.concatenate(offs).astype(np.uint32 if self.global_store else np.uint16)
        self.slice_rows = np.concatenate(rows_out).astype(np.int32)
        self.n_slices = len(solve)

    @staticmethod
    def _narrow_index_words(o, rec, it, lane):
        """Index data of off-block records of a global face store. A lane's
        entries are sorted columns of one row; when every lane of a record spans
        fewer than 2^16 positions the record is NARROW: slot l holds lane l's
        base (its smallest index), and slots 64 + 4l + q hold its entries 2q and
        2q+1 as 16-bit offsets from the base (low, high half); padding entries
        read the base (their value is zero). 1.25 KB of indices per record
        instead of 2 KB. Other records keep index it of lane l at slot 8l + it.
        Returns (words, narrow per record)."""
        n_rec = o.shape[0]
        real = np.zeros(o.shape, dtype=bool)
        real[rec, it, lane] = True
        big = np.iinfo(np.int64).max
        lo = np.where(real, o, big).min(axis=1)              # (n_rec, WAVE)
        hi = np.where(real, o, -1).max(axis=1)
        has = hi >= 0
        base = np.where(has, lo, 0)
        nar = np.all(~has | (hi - lo < 1 << 16), axis=1)
        out = o.transpose(0, 2, 1).reshape(n_rec, -1).copy()  # wide layout
        if nar.any():
            d = np.where(real, o - base[:, None, :], 0)[nar]   # (n_nar, 8, WAVE)
            words = np.zeros((int(nar.sum()), MAX_ITER * WAVE), dtype=np.int64)
            words[:, :WAVE] = base[nar]
            packed = d[:, 0::2, :] | (d[:, 1::2, :] << 16)     # (n_nar, 4, WAVE)
            words[:, WAVE:WAVE + 4 * WAVE] = packed.transpose(0, 2, 1).reshape(len(packed), -1)
            out[nar] = words
        return out.reshape(-1), nar

    def _block_index_words(self, rec, it, lane, ix, g, n_rec):
        """Index data of block-phase records of a global face store. A lane's
        entries in one record address consecutive scratch rows, so its 8
        indices are implicit: one word per lane (record slot l), base | len <<
        28, and entry it reads base + it*g for it < len, else index 0 (a
        position the face phase always writes; the value there is zero). The
        kernel loads 256 bytes of indices instead of 2 KB. LDS stores keep
        explicit indices (there the decode cost more than it saved)."""
        base = np.zeros((n_rec, WAVE), dtype=np.int64)
        cnt = np.zeros((n_rec, WAVE), dtype=np.int64)
        first = it == 0
        base[rec[first], lane[first]] = ix[first]
        np.add.at(cnt, (rec, lane), 1)
        if not np.array_equal(ix, base[rec, lane] + it * g) or cnt.max(initial=0) > MAX_ITER:
            raise AssertionError("block-phase entries of a lane are not consecutive scratch rows")
        if base.max(initial=0) >= 1 << 28:
            raise ValueError("face store too large for implicit block indices (2^28 doubles)")
        out = np.zeros((n_rec, MAX_ITER * WAVE), dtype=np.int64)
        out[:, :WAVE] = base | (cnt << 28)
        return out.reshape(-1)

    def record_indices(self, ibase, flags):
        """Face-store index of entry it of every lane of a solve record, (8, 64):
        read from the index array, or decoded from the lane words of a
        global-store block-phase record (_block_index_words)."""
        lane = np.arange(WAVE)
        if flags & _native.SLICE_NARROW:
            base = self.ent_off[ibase + lane].astype(np.int64)
            out = np.empty((MAX_ITER, WAVE), dtype=np.int64)
            for q in range(MAX_ITER // 2):
                w = self.ent_off[ibase + WAVE + 4 * lane + q].astype(np.int64)
                out[2 * q] = base + (w & 0xFFFF)
                out[2 * q + 1] = base + (w >> 16)
            return out
        if not (flags & _native.SLICE_BLOCK and self.index_bytes == 4):
            return np.stack([self.ent_off[ibase + 8 * lane + it].astype(np.int64) for it in range(MAX_ITER)])
        words = self.ent_off[ibase + lane].astype(np.int64)
        base, cnt = words & ((1 << 28) - 1), words >> 28
        it = np.arange(MAX_ITER)[:, None]
        return np.where(it < cnt[None, :], base[None, :] + it * (flags & 0xFF), 0)

    def wave_program(self, w):
        """Executed records of wave w (the four lookahead records stripped)."""
        return self.solve_slices[int(self.wave_ptr[w]):int(self.wave_ptr[w + 1]) - 4]

    def __getstate__(self):
        # the ctypes record arrays are rebuilt from the record lists
        state = dict(self.__dict__)
        state.pop("_face_rec", None)
        state.pop("_solve_rec", None)
        return state

    def __setstate__(self, state):
        self.__dict__.update(state)
        self._face_rec = self._slice_array(self.face_slices)
        self._solve_rec = self._slice_array(self.solve_slices)

    @staticmethod
    def _slice_array(recs):
        arr = (_native.Slice * max(len(recs), 1))()
        for i, (b, ib, rb, r0, it, fl) in enumerate(recs):
            arr[i].base, arr[i].ibase, arr[i].rowbase = b, ib, rb
            arr[i].r0, arr[i].n_iter, arr[i].flags = r0, it, fl
        return arr

    def desc(self, max_columns=1):
        """ctypes ProblemDesc pointing at this object's arrays (keep self alive),
        for runs of at most ``max_columns`` columns."""
        sp = self.solve_plan
        P = _native.ptr
        d = _native.ProblemDesc()
        d.n_junctions, d.n_faces = self.Nj, self.Nf
        d.stencil_width, d.block_size = self.stencil_width, self.column_threads
        d.parts = self.parts
        d.column_pairs = int(getattr(self, "column_pairs", False))
        d.max_columns = int(max_columns)
        d.Ic = P(self.Ic, C.c_double)
        d.Rv = P(self.Rv, C.c_double)
        d.coef = P(self.coef, C.c_double)
        d.junc_code = P(self.junc_pos_code, C.c_int32)
        d.face_of_pos = P(self.face_of_pos, C.c_int32)
        d.n_face_slices = len(self.face_slices)
        d.n_solve_slices = len(self.solve_slices)
        d.face_slices = C.cast(self._face_rec, C.c_void_p)
        d.solve_slices = C.cast(self._solve_rec, C.c_void_p)
        d.solve_wave_ptr = P(self._wave_ptr, C.c_int32)
        d.n_face_entries = len(self.face_code_sliced)
        d.face_code = P(self.face_code_sliced, C.c_int32)
        d.n_slice_rows = len(self.slice_rows)
        d.slice_rows = P(self.slice_rows, C.c_int32)
        d.n_entries = len(self.ent_val)
        d.ent_val = P(self.ent_val, C.c_double)
        d.n_index = len(self.ent_off)
        d.ent_off = self.ent_off.ctypes.data
        d.max_stage_rows = self.scratch_rows
        d.face_copy_rows = self.face_copy_rows
        d.index_bytes = self.index_bytes
        if self.uses_lattice:
            self.lattice.set_desc(d)
        return d

    def lds_bytes(self):
        """Bytes of the per-column face store (LDS or global)."""
        return 8 * (self.Nf + self.scratch_rows)

    # ---------------------------------------------------- host emulation
    def factor_bytes(self):
        """Bytes of solve-program data one column loads per step: every executed
        record moves 8 values and 8 LDS indices per lane, a row per lane and its
        32-byte descriptor."""
        return int(self.n_solve_exec * (512 * 8 + 512 * self.index_bytes + WAVE * 4 + 32))

    def emulate_face_and_solve(self, u):
        """Exactly the kernel's face phase and per-wave solve programs, on the host:
        u (Nj,) = x/c0 - theta_s  ->  J (Nf,) = K^-1 (A u)  (f = 0). Lanes
        accumulate with FMA-free numpy adds of products, so sums agree with the
        kernel to rounding, not bit for bit.

        Between two barriers each wave runs its records in order and sees its own
        writes at once; any face-store element that one wave writes and another
        wave reads or writes between the same two barriers raises AssertionError
        (a race on the GPU). A workgroup barrier (WG_BARRIER) joins only the
        waves of one part: within a column phase, an element that one part writes
        and another part touches raises too. Reads with a zero factor value are
        left out: they add an exact zero whatever they read."""
        BS, Nf = self.column_threads, self.Nf
        zs = np.zeros(Nf + max(self.scratch_rows, 1))
        tids = np.arange(BS)
        acc = np.zeros(BS)
        for (base, _, rowbase, r0, n_it, flags) in self.face_slices:
            for it in range(n_it):
                code = self.face_code_sliced[base + it * BS + tids]
                v = np.where(code >= 0, u[np.maximum(code, 0) >> 1], 0.0)
                acc = acc + np.where(code & 1, -v, v)
            if flags & _native.SLICE_FINISH:
                pos = self.slice_rows[rowbase + tids]
                ok = pos >= 0
                zs[pos[ok]] = acc[ok]
                cp = ok & (pos < self.face_copy_rows)
                zs[Nf + pos[cp]] = acc[cp]
                acc = np.zeros(BS)
        if self.uses_lattice:
            return self.lattice.emulate(zs[:Nf][self.lattice.pos_of_face], self.parts)
        return self._emulate_records(zs)

    def emulate_solve(self, b):
        """Exactly the solve kernel (jja_solve) on the host: b (Nf,) in face order
        is loaded by solve position (the first stage's scratch too), the
        per-wave record programs run as in emulate_face_and_solve, and K^-1 b
        comes back in face order."""
        if self.uses_lattice:
            return self.lattice.emulate(b, self.parts)
        Nf = self.Nf
        zs = np.zeros(Nf + max(self.scratch_rows, 1))
        pos = self.solve_plan.face_pos.astype(np.int64)
        zs[pos] = b
        cp = pos < self.face_copy_rows
        zs[Nf + pos[cp]] = np.asarray(b)[cp]
        return self._emulate_records(zs)

    def _emulate_records(self, zs):
        """The per-wave solve programs on the face store zs (emulate_face_and_solve)."""
        BS, Nf = self.column_threads, self.Nf
        # split every wave's program at its column barriers into phases, and a
        # phase at its workgroup barriers into sub-phases: phase p of all waves
        # runs before phase p+1; within a phase, the waves of one part meet at
        # each workgroup barrier, and different parts never meet
        NW = BS // WAVE
        per_part = NW // self.parts
        WGB, COLB = _native.SLICE_WG_BARRIER, _native.SLICE_BARRIER
        phases = []
        for w in range(NW):
            cur, sub, mine = [], [], []
            for rec in self.wave_program(w):
                cur.append(rec)
                if rec[5] & (WGB | COLB):
                    sub.append(cur)
                    cur = []
                    if rec[5] & COLB:
                        mine.append(sub)
                        sub = []
            assert not sub, "work after the last barrier"
            assert not cur or all(r[5] & 0xFFFF0000 == 0 for r in cur), "work after the last barrier"
            phases.append(mine)
        n_ph = len(phases[0])
        assert all(len(p) == n_ph for p in phases), "waves disagree on the barrier count"
        lane = np.arange(WAVE)
        n = len(zs)

        def run_wave(recs, mem):
            """One wave's records in order; returns its view and write/read masks."""
            mem = mem.copy()
            wmask = np.zeros(n, dtype=bool)
            rmask = np.zeros(n, dtype=bool)
            acc = np.zeros(WAVE)
            for (vbase, ibase, rowbase, r0, n_it, flags) in recs:
                G = flags & 0xFF
                offs = self.record_indices(ibase, flags)
                for it in range(MAX_ITER):
                    v = self.ent_val[vbase + (it // 2) * 128 + 2 * lane + it % 2]
                    off = offs[it]
                    rmask[off[v != 0]] = True
                    acc = acc + v * mem[off]
                if flags & _native.SLICE_FINISH:
                    red = acc.reshape(WAVE // G, G).sum(axis=1)
                    rows = self.slice_rows[rowbase + np.arange(WAVE // G)]
                    live = np.arange(WAVE // G) < (flags >> 16)
                    ok = live & (rows >= 0)
                    rows, red = rows[ok].astype(np.int64), red[ok]
                    if flags & _native.SLICE_BLOCK:
                        mem[rows] = red
                        wmask[rows] = True
                    else:
                        rmask[rows] = True
                        mem[Nf + rows - r0] = mem[rows] - red
                        wmask[Nf + rows - r0] = True
                    acc = np.zeros(WAVE)
            return mem, wmask, rmask

        def check(wrote, read, what):
            writers = np.sum(wrote, axis=0)
            assert writers.max(initial=0) <= 1, f"two {what}s write one face-store element between barriers"
            for i in range(len(wrote)):
                assert not np.any((writers - wrote[i])[read[i]]), \
                    f"{what} {i} reads an element another {what} writes between the same barriers"

        for ph in range(n_ph):
            part_mem, part_w, part_r = [], [], []
            for q in range(self.parts):
                waves = range(q * per_part, (q + 1) * per_part)
                n_sub = len(phases[waves[0]][ph])
                assert all(len(phases[w][ph]) == n_sub for w in waves), "waves of a part disagree on barriers"
                mem = zs.copy()
                tw = np.zeros(n, dtype=bool)
                tr = np.zeros(n, dtype=bool)
                for j in range(n_sub):
                    out = [run_wave(phases[w][ph][j], mem) for w in waves]
                    check([o[1] for o in out], [o[2] for o in out], "wave")
                    for view, wm, rm in out:
                        mem[wm] = view[wm]
                        tw |= wm
                        tr |= rm
                part_mem.append(mem)
                part_w.append(tw)
                part_r.append(tr)
            check(part_w, part_r, "part")
            for mem, wm in zip(part_mem, part_w):
                zs[wm] = mem[wm]
        return zs[self.solve_plan.face_pos]


class MatrixPlan(CircuitPlan):
    """Factor and solve program of one sparse symmetric positive definite matrix
    on its own (no circuit): the device solves K x = b for many right-hand
    sides with it (jja_solve), as scipy's ``factorized(K)`` does on the host for
    the reference's cycle and cut-space systems (josephson_circuit.py:93-128).
    ``xy`` (n, 2): coordinates of the unknowns, for the nested-dissection
    order. The descriptor has no junctions, so the time-evolution launch
    refuses it."""

    def __init__(self, K, xy, leaf_size=16, block_size=None, face_store="auto"):
        # a private copy: index sorting during the ordering must not reorder
        # arrays the caller's matrix shares
        K = scipy.sparse.csc_matrix(K, dtype=np.float64, copy=True)
        n = K.shape[0]
        if K.shape != (n, n) or n == 0:
            raise ValueError("MatrixPlan needs a non-empty square matrix")
        if face_store not in ("auto", "lds", "global"):
            raise ValueError("face_store must be 'auto', 'lds' or 'global'")
        self.face_store = face_store
        self.parts = 1
        self.local_option = "auto"
        self.solver = "cholesky"
        self.lattice = None
        self.circuit_xy = None
        self.column_pairs = False
        self.Nj, self.Nf = 0, n
        self.stencil_width = 3
        self.dt = 1.0
        if block_size is None:
            block_size = 1024
            for bs in (64, 128, 256, 512):
                if n <= 2 * bs:
                    block_size = bs
                    break
        self.block_size = block_size
        self.face_ptr = np.zeros(n + 1, dtype=np.int32)
        self.face_code = np.zeros(0, dtype=np.int32)
        self.junc_code = np.zeros(0, dtype=np.int32)
        self.coef = np.zeros((3, 0))
        self.Rv = np.zeros(0)
        self.Ic = np.zeros(0)
        self.K = K
        self.solve_plan = SolvePlan(K, np.asarray(xy, dtype=np.float64), leaf_size=leaf_size)
        self.target_iter = 8
        self._build_layouts(self.target_iter)
